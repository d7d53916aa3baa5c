// Batched continuous-Q2 solves (many independent instances per launch).
#pragma once

#include "fem_op.hpp"

namespace femamd {

struct Batch {
    Context *ctx = nullptr;
    int count = 0;
    double creac = 0;
    std::vector<int> nx, ny;
    std::vector<int64_t> dof_off, elem_off;
    int64_t N = 0, E = 0;
    int shape = 0;                       // index into CG_SHAPES
    float tune_ms[CG_NSHAPES] = {};
    DevBuf inst_dev[CG_NSHAPES];
    int tiles[CG_NSHAPES] = {};
    DevBuf tab, kvol64, kvol32, dinv64, dinv32;
    DevBuf seg_inst, seg_lo, seg_hi, seg_first, seg_counter, seg_partials;
    int nseg = 0;
    DevBuf sdbl, sint, active;
    int *hactive = nullptr;
    DevBuf h64, r64, r32, d32, p32, q32, s32, h32;
    // one captured inner CG step; it holds the addresses of the work,
    // coefficient and Jacobi buffers, so it is dropped whenever one of them
    // is (re)allocated or the tile shape changes
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void drop_graph();
    std::vector<int> last_iters;
    std::vector<double> last_relres;

    void init(Context *c, int n, const int *nx, const int *ny, double creac);
    void set_kappa_qp(const double *kvol);
    void apply_raw(int prec, const void *u, void *w, const int *inst_done) const;
    void rhs(const double *fq, Vec &F) const;
    int autotune(int reps);
    void solve(const Vec &b, Vec &x, double tol, double inner_tol, int inner_maxit, int max_outer, int use_graph,
               fem_batch_info &info);
    ~Batch();
};

}  // namespace femamd

struct fem_batch { femamd::Batch b; femamd::Context *held = nullptr; };
