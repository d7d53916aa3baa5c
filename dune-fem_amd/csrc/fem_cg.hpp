// Continuous Lagrange Q2 operator on structured quad meshes of [0,1]^2,
// single instance or a batch of independent instances in one launch.
#pragma once

#include "fem_common.hpp"

namespace femamd {

// one problem instance of a batch
struct CGInst {
    int nx, ny;
    int64_t dof_off;    // offset of its DoFs in the batched vectors
    int64_t elem_off;   // offset of its elements (kappa / f arrays: elem_off * 9)
    int tiles_x;        // tiles per row
    int tile_off;       // first workgroup of this instance
    int ntiles;
    double c;           // reaction coefficient
    double kappa_const; // used when kvol == nullptr
};

constexpr int CG_TX = 15, CG_TY = 15;                  // default tile: owned elements per tile
constexpr int CG_THREADS = (CG_TX + 1) * (CG_TY + 1);  // + lower halo row/column

// tile shapes available to the autotuner (owned elements per tile)
constexpr int CG_NSHAPES = 5;
constexpr int CG_SHAPES[CG_NSHAPES][2] = { { 15, 15 }, { 7, 7 }, { 15, 7 }, { 31, 7 }, { 7, 31 } };

struct CGTab {
    double L[3][3];   // Q2 basis i at Gauss point q: L[q][i]
    double Ld[3][3];  // derivative
    double w[3];
};

enum { CG_MODE_APPLY = 0, CG_MODE_DIAG = 1, CG_MODE_RHS = 2, CG_MODE_RHS_ADD = 3 };

// the reaction beyond the constant c of a single instance: a pointwise
// coefficient at the volume quadrature points [E][9] (qx + 3 qy) and the
// nonlinear law coef u^pw, both multiplying u v
template <typename T>
struct CGReaction {
    const T *cvol = nullptr;
    T coef = 0;
    int pw = 0;
};

// mode APPLY  : out = A in (Dirichlet rows: out = in)
// mode DIAG   : out = diag(A) (Dirichlet rows: 1), `in` unused
// mode RHS    : out = F from f at quadrature points (kvol slot carries f), Dirichlet rows 0
// mode RHS_ADD: out += F on the interior rows (Dirichlet rows unchanged)
template <typename T>
void cg_launch(int mode, const CGInst *d_inst, int ninst, int total_tiles, const CGTab *d_tab, const T *kvol,
               const T *in, T *out, const int *skip, hipStream_t s, const CGReaction<T> &rx = CGReaction<T>());
// same with an explicit tile shape (index into CG_SHAPES) and an optional
// per-instance mask: instances with inst_done[i] != 0 are skipped
template <typename T>
void cg_launch_shape(int shape, int mode, const CGInst *d_inst, int ninst, int total_tiles, const CGTab *d_tab,
                     const T *kvol, const T *in, T *out, const int *skip, const int *inst_done, hipStream_t s,
                     const CGReaction<T> &rx = CGReaction<T>());
// out[e*9 + qx + 3 qy] = base + coef u(x_q)^pw (base may be null) for a
// single instance of nx x ny elements
template <typename T>
void cg_eval_qp(int nx, int ny, const CGTab *d_tab, const T *u, const T *base, T coef, int pw, T *out, hipStream_t s);

CGTab make_cg_tab();

}  // namespace femamd
