// Multigrid preconditioner for SIPG DG operators with cellwise or constant
// kappa and no advection: p-levels, a conforming trilinear vertex level and
// geometric vertex h-levels (one symmetric V-cycle per application).
#pragma once

#include <memory>

#include "fem_op.hpp"

namespace femamd {

// one DG level: Q_p on the fine mesh, its exact element-block inverse
struct MGDGLevel {
    std::unique_ptr<Operator> own;   // coarser DG levels own their operator
    Operator *op = nullptr;          // level 0: the preconditioner's operator
    // level 0 of a BF16-MFMA operator: its FP32 twin forms the residual that
    // is restricted to the coarse levels (the smoothing sweeps keep BF16)
    std::unique_ptr<Operator> resop;
    Precond bj;
    int k = 0;                       // points/modes per direction, p + 1
    DevBuf x, r, t;                  // correction, residual, A x
};

// one conforming trilinear level on a vertex lattice of nv[0] x nv[1] x nv[2]
// vertices: 27-point stencil [V][27], offset s = (dx+1) + 3 (dy+1) + 9 (dz+1)
struct MGVertexLevel {
    int nv[3] = { 0, 0, 0 };
    int64_t V = 0;
    DevBuf A, dinv, x, r, t;
};

struct Multigrid {
    Operator *op = nullptr;
    int prec = FEM_F64;
    std::vector<std::unique_ptr<MGDGLevel>> dg;
    std::vector<std::unique_ptr<MGVertexLevel>> vx;
    DevBuf coarse_inv;               // dense inverse of the coarsest vertex level
    DevBuf corners;                  // [E][8] FP64: Q1 modes -> vertex functionals
    int nu_dg = 2, nu_vertex = 1;    // smoothing sweeps before and after the coarse correction
    double omega_dg = 0.8, omega_vertex = 0.7;
    fem_op_desc desc{};              // the finest operator's, at init
    bool reaction = false;           // built with a pointwise reaction cvol
    // set-up data kept for update(): the FP64 stencils per vertex level, the
    // element means of cvol, and scratch for the level data and the
    // coarsest inverse
    std::vector<std::unique_ptr<DevBuf>> A64;
    DevBuf cmean, d64, As64, gjX, gjY, gjp;

    // builds the hierarchy (levels, operators, buffers), then update()
    void init(Operator *fine);
    // computes everything that depends on the finest operator's coefficients
    // (kappa, c, cvol) in the existing buffers: the coarse operators'
    // coefficients, the smoother blocks, the vertex stencils and Jacobi
    // scalings, the coarsest inverse.  No allocation after the first call
    // (the host fallback of an indefinite coarsest level aside).
    void update();
    // the operator the hierarchy was built on, with the same structure
    // (a pointwise reaction present or not)
    bool updatable_for(const Operator *o) const;
    // z = V-cycle(r); `skip` (device flag, may be null) skips all work
    void apply_raw(const void *r, void *z, const int *skip);
    // one damped block-Jacobi sweep of the finest level: x += omega B^-1 (r - t)
    // with t = A x given
    void smooth_fine(const void *r, const void *t, void *x, const int *skip);

private:
    template <typename T> void cycle_dg(size_t l, const T *r, T *x, const int *skip);
    template <typename T> void cycle_vertex(size_t l, const T *r, T *x, const int *skip);
};

}  // namespace femamd
