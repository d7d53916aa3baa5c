// Device-driven Krylov solvers behind the InverseOperator interface.
#pragma once

#include <functional>
#include <memory>

#include "fem_op.hpp"

namespace femamd {

constexpr int GM_MAX = 64;   // maximum GMRES restart length

// scalars of one solve, resident in device memory; kernels read and update
// them so that the iteration loop needs no host round trip per step
struct KState {
    double res, prev, qh, alpha, beta, tol, bnorm2, scale;
    int it, done, maxit, criteria;
    int crit_res;      // stop on ||r||^2 instead of <r, M^-1 r>
    double rr;         // ||r||^2 of the current residual
    // GMRES
    double H[(GM_MAX + 1) * GM_MAX];
    double g[GM_MAX + 1], cs[GM_MAX], sn[GM_MAX], y[GM_MAX];
    double gtol, hinv;
    int jlast;       // inner steps done in the current cycle
    int converged;   // cycle ended by |g| < tol
    int cycle_done;  // stop inner loop (converged or maxit)
    // BiCGStab
    double nu, omega, bres;
    int fin;         // the iteration just reduced is the last one
};

struct Solver {
    Context *ctx = nullptr;
    fem_solver_params prm{};
    Operator *op = nullptr, *iop = nullptr;
    Precond *pre = nullptr, *ipre = nullptr;
    // preconditioners built at bind from fem.solver.preconditioning.method
    // when the caller passes none
    std::unique_ptr<Precond> own_pre, own_ipre;
    std::unique_ptr<Precond> widened;   // an FP32 preconditioner serving the FP64 solve
    DevBuf state, partials, counters;
    KState *hstate = nullptr;   // pinned mirror
    std::vector<DevBuf *> work;
    // CG: called with the device iterate and residual every snap_chunk steps
    std::function<void(const void *x, const void *r)> snapshot;
    int snap_chunk = 0;
    int fallback_used = 0;

    void init(Context *c, const fem_solver_params &p);
    void bind(Operator *o, Precond *p, Operator *io, Precond *ip);
    void unbind();
    void solve(const Vec &b, Vec &x, fem_solver_info &info);
    ~Solver();

    // one Krylov solve in precision of `o` on raw pointers; returns signed iterations
    int krylov(int method, Operator *o, Precond *p, const void *b, void *x, double tol, int criteria, int maxit,
               int restart, int64_t N, int prec);
    void *workbuf(int i, size_t bytes);
    void fetch_state();
    double true_residual(const Vec &b, const Vec &x, Vec &r);
};

}  // namespace femamd

struct fem_solver { femamd::Solver s; femamd::Context *held = nullptr; };
