// Learned coarse-space preconditioner (CNN coarse correction + block Jacobi).
#pragma once

#include "fem_op.hpp"

namespace femamd {

struct CNN;

struct Learned {
    Operator *op = nullptr;
    CNN *cnn = nullptr;
    Precond bj;                 // exact element-block inverse
    // 0 additive, 1 multiplicative two-level preconditioner; 2 restart mode:
    // the preconditioner is block Jacobi and the network corrects the iterate
    // from its true residual at every FGMRES restart
    int multiplicative = 1;
    bool enabled = true;        // false: block Jacobi only (fallback / ablation)
    DevBuf z1, t1, w1, x0, yc, red, partials, counter;
    void init(Operator *o, CNN *c, int multiplicative);
    void apply_raw(const void *x, void *y, const int *skip);
    // restart mode: x += P C(R r) for the residual r (device, operator precision)
    void correct(const void *r, void *x);
    bool restart_mode() const { return multiplicative == 2; }
};

void cnn_set_coefficient(CNN &c, const double *kcell, double beta);
// runs the plain FP64 CG + block-Jacobi solve of A x = b from x = 0 and stores
// coarse (residual, error) pairs of every `every`-th iterate; returns the count
int cnn_collect_pairs(CNN &c, Operator *op64, Precond *bj64, const Vec &b, int max_pairs, int every);
void cnn_train(CNN &c, int epochs, int batch, float lr, float val_frac, unsigned seed, double &train_loss,
               double &val_rel);

}  // namespace femamd
