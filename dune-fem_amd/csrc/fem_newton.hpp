// Newton's method over the matrix-free operators: the inverse operator of a
// nonlinear operator S, solving S[w] = u.
#pragma once

#include "fem_solver.hpp"

namespace femamd {

// the forcing-term sequence of Eisenstat and Walker (choice 2, with the
// safeguards of their paper): the relative tolerance of each linear solve
struct EisenstatWalker {
    double eta_max = 0.99, gamma = 0.1, tol = 1e-6;
    double prev_res = 0, prev_eta = -1;
    double next(double res);
};

struct Newton {
    Context *ctx = nullptr;
    fem_newton_params prm{};
    Solver lin;                       // the linear inverse operator (owned)
    Operator *op = nullptr;           // bound operator S
    Precond *pre = nullptr;           // optional preconditioner of every linear solve
    std::unique_ptr<Operator> jop;    // S'(w) of a nonlinear S
    std::unique_ptr<Operator> jop32;  // its FP32 twin (linear solver with refinement)
    DevBuf res, dw;
    EisenstatWalker ew;
    // state of the last solve
    double delta = 0;
    int iterations = 0, linear_iterations = 0;
    int linear_limit = 0;   // total linear iterations allowed in this solve
    bool step_completed = true;
    double timing[3] = { 0, 0, 0 };

    void init(Context *c, const fem_newton_params &p, const fem_solver_params &linear);
    void bind(Operator *o, Precond *p);
    void unbind();
    void solve(const Vec *u, Vec &w, fem_newton_info &info);
    int failed() const;

  private:
    int max_linear() const;
    void residual_vec(const Vec *u, const Vec &w, Vec &r) const;   // r = S[w] - u
    double residual(const Vec *u, const Vec &w, Vec &r) const;       // the same, returns ||r||_2
    int line_search(const Vec *u, Vec &w, const Vec &d, Vec &r);
    void update_linear_tolerance();
};

}  // namespace femamd

struct fem_newton { femamd::Newton n; femamd::Context *held = nullptr; };
