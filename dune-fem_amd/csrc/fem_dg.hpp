// SIPG DG operator on structured hexahedral meshes: kernel arguments and launchers.
#pragma once

#include "fem_blas.hpp"

namespace femamd {

// division by a launch constant d >= 1 for 0 <= x < 2^31: with
// l = ceil(log2 d), m = ceil(2^(31+l) / d), x / d = (x m) >> (31 + l)
// (Granlund and Montgomery, 1994); wave-uniform operands stay scalar
struct UDiv {
    unsigned long long m = 1ull << 31;
    int sh = 31;
};
inline UDiv make_udiv(unsigned d)
{
    UDiv v;
    int l = 0;
    while ((1ull << l) < d) ++l;
    v.sh = 31 + l;
    v.m = ((1ull << v.sh) + d - 1) / d;
    return v;
}
__host__ __device__ inline unsigned udiv(unsigned x, UDiv v)
{
    return (unsigned)(((unsigned long long)x * v.m) >> v.sh);
}

template <typename T>
struct DGArgs {
    int n[3];
    T h[3];
    T ih[3];    // 1 / h
    T beta;
    T b[3];
    T c;
    int kappa_mode;
    T kappa_const;
    const T *kvol;
    const T *kface[3];
    const T *kcell;
    // reaction at the volume quadrature points added to c (DG operators with
    // constant or cellwise kappa): a pointwise coefficient cvol [E][k^3]
    // (x fastest, may be null) and the nonlinear law nl_coef u^nl_pow
    // evaluated from the input at each point (nl_pow = 0: none); the volume
    // term is int (c + cvol + nl_coef u^nl_pow) u v
    const T *cvol;
    T nl_coef;
    int nl_pow;
    const Tab1D<T> *tab;   // device copy of the 1D tables
    // per direction d, set on the host: beta / h_d, the face area of the
    // faces normal to d, |K| / h_d^2, b_d h_d, max(b_d, 0), max(-b_d, 0)
    T pen[3], farea[3], vsc[3], bh[3], bpos[3], bneg[3];
    // element traversal of the apply (order k-1, epb elements per
    // workgroup): x-rows cut into rpb = ceil(nx / epb) groups of epb
    // elements (the last group of a row padded with shadow lanes that
    // repeat the row's last element and write nothing), visited in slabs of
    // ty y-rows through all z; `slabs` = 0 selects plain element order
    int slabs, rpb, ty;
    UDiv dv_slab, dv_zrow, dv_rpb;   // by rpb ty nz, rpb ty, rpb
};

// the damped block-Jacobi sweep fused into the apply (multigrid smoothing):
// xout = u + omega B^-1 (r - A u) with B^-1 the exact element-block inverse
// from the eigen data Q [E][3][k][k], lam [E][3][k]
template <typename T>
struct DGSmooth {
    const T *r = nullptr;
    const T *Q = nullptr;
    const T *lam = nullptr;
    T omega = T(0);
    T *xout = nullptr;
};

// elements per workgroup of the apply for k points per direction
int dg_apply_epb(int k);

// fills the per-direction constants and the traversal of DGArgs from n, h,
// ih, beta, b and the number of 1D points k
template <typename T>
inline void dg_derived(DGArgs<T> &a, int k)
{
    const int epb = dg_apply_epb(k);
    a.slabs = 1;
    a.rpb = (a.n[0] + epb - 1) / epb;
    const int ny = a.n[1];
    a.ty = (ny % 16 == 0) ? 16 : (ny % 8 == 0) ? 8 : (ny % 4 == 0) ? 4 : (ny % 2 == 0) ? 2 : 1;
    if (a.slabs) {
        a.dv_slab = make_udiv((unsigned)(a.rpb * a.ty * a.n[2]));
        a.dv_zrow = make_udiv((unsigned)(a.rpb * a.ty));
        a.dv_rpb = make_udiv((unsigned)a.rpb);
    }
    for (int d = 0; d < 3; ++d) {
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        a.pen[d] = a.beta * a.ih[d];
        a.farea[d] = a.h[t1] * a.h[t2];
        a.vsc[d] = a.h[0] * a.h[1] * a.h[2] * (a.ih[d] * a.ih[d]);
        a.bh[d] = a.b[d] * a.h[d];
        a.bpos[d] = a.b[d] > T(0) ? a.b[d] : T(0);
        a.bneg[d] = a.b[d] < T(0) ? -a.b[d] : T(0);
    }
}

// w = A u.  `skip` (device flag, may be null): when *skip != 0 the launch
// does nothing (device-driven solver loops).  `dot` (may be null): each
// workgroup also writes its FP64 partial of <u, w> to dot[workgroup]
// (dg_blocks(k, E) entries), for the solver's <q, A q>.
void dg_apply_f64(const DGArgs<double> &a, int k, const double *u, double *w, const int *skip, hipStream_t s,
                  double *dot = nullptr);
void dg_apply_f32(const DGArgs<float> &a, int k, const float *u, float *w, const int *skip, hipStream_t s,
                  double *dot = nullptr);
// the sweep xout = u + omega B^-1 (r - A u) in one pass (cellwise or
// constant kappa; xout must not alias u, whose neighbours the pass reads)
void dg_smooth_f64(const DGArgs<double> &a, int k, const double *u, const DGSmooth<double> &sm, const int *skip,
                   hipStream_t s);
void dg_smooth_f32(const DGArgs<float> &a, int k, const float *u, const DGSmooth<float> &sm, const int *skip,
                   hipStream_t s);
// BF16 MFMA contraction variant (FP32 data and accumulation)
void dg_apply_bf16(const DGArgs<float> &a, int k, const float *u, float *w, const int *skip, hipStream_t s);
// workgroups of the DG apply launch for order k-1 on n[0] x n[1] x n[2]
// elements (its per-workgroup dot partials), and of the block-Jacobi launch
int64_t dg_blocks(int k, const int *n);
int64_t dg_bj_blocks(int k, int64_t E);
// operator diagonal
void dg_diagonal(const DGArgs<double> &a, int k, double *d, hipStream_t s);
// a DG function at the volume quadrature points [E][k^3] (x fastest):
// out = base + coef u^pw (base may be null)
template <typename T>
void dg_eval_qp(const DGArgs<T> &a, int k, const T *u, const T *base, T coef, int pw, T *out, hipStream_t s);
// load vector from f at volume quadrature points
void dg_rhs(const DGArgs<double> &a, int k, const double *fq, double *F, hipStream_t s, bool add = false);
// F += boundary-data part of the load vector (u = g on the boundary, g at
// the face quadrature points [faces_d][m^2] per normal direction)
void dg_rhs_bnd(const DGArgs<double> &a, int k, const double *const g[3], double *F, hipStream_t s);

// Exact element-block inverse for cellwise kappa (Kronecker-sum structure of
// the element block of the SIPG diffusion operator in the orthonormal
// Legendre basis): per element and direction an orthogonal K x K matrix Q_d
// and eigenvalues lam_d with A_KK = sum_d (I..A_d..I), A_d = Q_d diag(lam_d) Q_d^T.
// A pointwise reaction cvol enters by its element mean, or (bound) by its
// largest value in the element, which makes the block an upper bound of the
// true one (the form a smoother needs).  Computed in FP64, stored as TO
// (double or float).
template <typename TO>
void dg_block_setup(const DGArgs<double> &a, int k, TO *Q /*[E][3][K][K]*/, TO *lam /*[E][3][K]*/,
                    hipStream_t s, bool bound = false);
// `dot` (may be null): each workgroup writes its FP64 partials of <x, y> to
// dot[workgroup] and of <x, x> to dot[dg_bj_blocks(k, E) + workgroup]
template <typename T>
void dg_block_apply(int k, int64_t E, const T *Q, const T *lam, const T *x, T *y, const int *skip, hipStream_t s,
                    double *dot = nullptr);


}  // namespace femamd
