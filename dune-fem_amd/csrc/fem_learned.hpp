// Learned coarse-space preconditioner (CNN coarse correction + block Jacobi).
#pragma once

#include "fem_op.hpp"

namespace femamd {

struct CNN;

// Jacobi-preconditioned CG on the 7-point Galerkin coarse operator of a CNN
// grid, entirely on the device (scalars stay in device memory, no host sync
// per step); FP32 vectors, FP64 reductions with a deterministic finish.
// Chronopoulos-Gear recurrences: one kernel per step; each step reduces the
// previous step's block partials itself, so no atomics.
struct CoarseCG {
    const CNN *cnn = nullptr;
    int V = 0;
    DevBuf x, p, r[2], w[2], s[2], part[2], sc;
    int cur = 0;     // index of the current r, w, s and partial buffers
    int nstep = 0;   // steps since start
    void init(const CNN *c);
    // x = x0, r = b - A_c x0 with b = src[v * stride]; x0 = s0 gamma0 y if
    // y is given (s0 = red[0]), else 0
    template <typename T>
    void start(const T *src, int stride, const float *y, const float *red, float gamma0, hipStream_t s,
               const int *skip);
    void step(hipStream_t s, const int *skip);
    // host: current (r, D^-1 r) product and its start value
    void gammas(hipStream_t s, double &now, double &first) const;
};

struct Learned {
    Operator *op = nullptr;
    CNN *cnn = nullptr;
    // smoother kind: FEM_PRE_BLOCK_JACOBI (coarse correction first, then the
    // block-Jacobi smoother on the remaining residual) or FEM_PRE_MULTIGRID
    // (a multigrid V-cycle first, then the coarse correction of what it
    // leaves, then one block-Jacobi sweep of the fine level); FEM_LEARNED_OFF
    // applies the smoother alone (with the multigrid: the V-cycle and the sweep)
    int smoother = FEM_PRE_BLOCK_JACOBI;
    Precond bj;                 // exact element-block inverse, or the multigrid
    CoarseCG ccg;
    int coarse_steps = 5;
    int mode = FEM_LEARNED_NET;
    DevBuf z1, t1, w1, x0, yc, acc, red, partials, counter;
    DevBuf dparts;   // per-workgroup <z1, A z1> partials of the apply (not for BF16 applies)
    void init(Operator *o, CNN *c, int coarse_steps, int smoother = FEM_PRE_BLOCK_JACOBI);
    void apply_raw(const void *x, void *y, const int *skip);
    bool active() const { return mode != FEM_LEARNED_OFF; }
};

void cnn_set_coefficient(CNN &c, const double *kcell, double beta);
// `steps` coarse CG steps from x = 0 on device arrays [V]; rz[0] = final
// (r, D^-1 r), rz[1] = its start value
void cnn_coarse_solve(const CNN &c, const float *b, int steps, float *x, double rz[2]);
// runs the plain FP64 CG solve of A x = b from x = 0 with `pre64` (block
// Jacobi or multigrid) and stores coarse (residual, error) pairs of every
// `every`-th iterate; with the multigrid the stored residual is what one
// V-cycle leaves of it, r - A MG(r) (the input of the learned correction that
// follows the V-cycle); returns the count
int cnn_collect_pairs(CNN &c, Operator *op64, Precond *pre64, const Vec &b, int max_pairs, int every);
void cnn_train(CNN &c, int epochs, int batch, float lr, float val_frac, unsigned seed, double &train_loss,
               double &val_rel);

}  // namespace femamd
