// 3D CNN on a structured grid (coarse-space correction network): BF16 MFMA
// forward/backward kernels with FP32 accumulation, Adam training.
#pragma once

#include "fem_op.hpp"

namespace femamd {

constexpr int CNN_LAYERS = 4;
constexpr int CNN_HID = 32;

struct CNN {
    Context *ctx = nullptr;
    int n[3] = { 0, 0, 0 };
    int64_t V = 0;
    int ch[CNN_LAYERS + 1] = { 2, CNN_HID, CNN_HID, CNN_HID, 1 };
    int dil[CNN_LAYERS] = { 1, 2, 4, 1 };
    size_t woff[CNN_LAYERS + 1] = {};   // offsets of each layer's [27][ci][co] block
    size_t nparams = 0;
    int nchunks = 0;                      // voxel chunks of the weight-gradient kernels
    // weights: FP32 master, BF16-rounded FP32 copy ([27][ci][co]), BF16 MFMA
    // operand layouts for the hidden layers: fwd [27][co][ci], bwd [27][ci][co]
    DevBuf wmaster, wq, wfwd, wbwd, grad, adam_m, adam_v, wpart;
    // activations: input [V][2] FP32, hidden [V][32] BF16, output [V] FP32,
    // pre-activation gradients FP32
    DevBuf x0, a1, a2, a3, y, g1, g2, g3, g4, lossp;
    int adam_t = 0;
    // training data: coarse residual / coarse error pairs [P][V] (raw), and
    // the coefficient feature 1/kappa per voxel
    DevBuf ds_r, ds_e, kinv;
    int ds_cap = 0, ds_n = 0;
    float gamma0 = 1.0f;   // output scale: median ||A_c^-1 r_c|| / ||r_c|| of the training set
    DevBuf feat_red;       // feature normalisation scalars of the current input
    // Galerkin coarse operator A_c = P^T A P of the SIPG operator on element
    // means: 7-point stencil [V][7] = (diag, -x, +x, -y, +y, -z, +z) weights
    DevBuf stencil;
    void set_problem(const double *kcell, double beta);
    void coarse_apply(const float *c, float *out, hipStream_t s) const;

    void init(Context *c, const int n[3], const int dil[CNN_LAYERS]);
    void set_weights(const float *host);
    void get_weights(float *host) const;
    void refresh();   // master -> rounded and MFMA copies
    // forward of input x (device [V][2]) -> y (device [V]); skip: device flag
    void forward(const float *x, float *y, hipStream_t s, const int *skip = nullptr);
    // loss = weight * mean((y - t)^2) and accumulate dL/dW into grad (device t [V])
    void backward_accumulate(const float *t, hipStream_t s, float weight = 1.0f);
    // accumulate dL/dW for a given dL/dy already stored in g4
    void backward_from_g4(hipStream_t s);
    double loss_value(hipStream_t s);
    void zero_grad(hipStream_t s);
    void adam_step(float lr, float scale, hipStream_t s);
};

}  // namespace femamd

struct fem_cnn { femamd::CNN c; femamd::Context *held = nullptr; };
