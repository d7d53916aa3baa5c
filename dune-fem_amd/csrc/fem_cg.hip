// Continuous Lagrange Q2 operator on structured quad meshes: matrix-free
// apply, diagonal and load vector, for one instance or a batch of
// independent instances in one launch.
//
// Reference semantics: GalerkinOperator::evaluate element loop with
// addLocalDofs scatter (dune/fem/schemes/galerkin.hh:811-917, 929-934) and
// DirichletWrapperOperator::operator() (dirichletwrapper.hh:101-105: w_D = u_D).
//
// Design: a workgroup owns a CG_TX x CG_TY tile of elements and the nodes at
// their lower-left 2x2 node block; one thread per element of the tile plus a
// lower halo row/column, so every contribution to an owned node is computed
// inside the workgroup.  Element results are added into an LDS node lattice
// in four colour phases (elements of one colour share no node), which makes
// the summation order fixed: results are deterministic and need no atomics.
#include "fem_cg.hpp"

#include <algorithm>

namespace femamd {

CGTab make_cg_tab()
{
    CGTab t;
    double x[3], w[3];
    gauss_legendre01(3, x, w);
    for (int q = 0; q < 3; ++q) {
        t.w[q] = w[q];
        for (int i = 0; i < 3; ++i) {
            t.L[q][i] = lagrange2(i, x[q]);
            t.Ld[q][i] = lagrange2_d(i, x[q]);
        }
    }
    return t;
}

template <int TX, int TY, typename T>
__global__ void __launch_bounds__((TX + 1) * (TY + 1))
    k_cg(int mode, const CGInst *__restrict__ insts, int ninst, const CGTab *__restrict__ tabp,
         const T *__restrict__ kvol, const T *__restrict__ in, T *__restrict__ out, const int *__restrict__ skip,
         const int *__restrict__ inst_done, const CGReaction<T> rx)
{
    constexpr int CG_LX = 2 * TX + 3, CG_LY = 2 * TY + 3;
    if (skip && *skip) return;
    __shared__ T lat[CG_LX * CG_LY];
    __shared__ int s_inst;
    const unsigned bid = blockIdx.x;
    if (threadIdx.x == 0) {
        int lo = 0, hi = ninst - 1;
        while (lo < hi) {  // last instance with tile_off <= bid
            int mid = (lo + hi + 1) >> 1;
            if ((unsigned)insts[mid].tile_off <= bid) lo = mid; else hi = mid - 1;
        }
        s_inst = lo;
    }
    for (int k = threadIdx.x; k < CG_LX * CG_LY; k += blockDim.x) lat[k] = 0;
    __syncthreads();
    if (inst_done && inst_done[s_inst]) return;   // uniform over the block
    const CGInst P = insts[s_inst];
    const CGTab &tb = *tabp;
    const int tile = (int)bid - P.tile_off;
    const int ex0 = (tile % P.tiles_x) * TX, ey0 = (tile / P.tiles_x) * TY;
    const int tx = threadIdx.x % (TX + 1), ty = threadIdx.x / (TX + 1);
    const int ex = ex0 - 1 + tx, ey = ey0 - 1 + ty;
    const bool valid = ex >= 0 && ey >= 0 && ex < P.nx && ey < P.ny;
    const int NX = 2 * P.nx + 1;
    const T *u = in + P.dof_off;
    T wl[3][3];
    if (valid) {
        const double hx = 1.0 / P.nx, hy = 1.0 / P.ny, detJ = hx * hy;
        const int64_t e = P.elem_off + (int64_t)ey * P.nx + ex;
        if (mode == CG_MODE_APPLY) {
            T ul[3][3];
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix) ul[iy][ix] = u[(int64_t)(2 * ey + iy) * NX + 2 * ex + ix];
            // sum factorisation: x then y
            T t0[3][3], t1[3][3];
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int qx = 0; qx < 3; ++qx) {
                    T a = 0, b = 0;
#pragma unroll
                    for (int ix = 0; ix < 3; ++ix) {
                        a += (T)tb.L[qx][ix] * ul[iy][ix];
                        b += (T)tb.Ld[qx][ix] * ul[iy][ix];
                    }
                    t0[iy][qx] = a;
                    t1[iy][qx] = b;
                }
            T fx[3][3], fy[3][3], fm[3][3];
#pragma unroll
            for (int qy = 0; qy < 3; ++qy)
#pragma unroll
                for (int qx = 0; qx < 3; ++qx) {
                    T uq = 0, gx = 0, gy = 0;
#pragma unroll
                    for (int iy = 0; iy < 3; ++iy) {
                        uq += (T)tb.L[qy][iy] * t0[iy][qx];
                        gx += (T)tb.L[qy][iy] * t1[iy][qx];
                        gy += (T)tb.Ld[qy][iy] * t0[iy][qx];
                    }
                    const T W = (T)(tb.w[qx] * tb.w[qy] * detJ);
                    const T kap = kvol ? kvol[e * 9 + qx + 3 * qy] : (T)P.kappa_const;
                    fx[qy][qx] = W * kap * gx * (T)(1.0 / (hx * hx));
                    fy[qy][qx] = W * kap * gy * (T)(1.0 / (hy * hy));
                    // reaction c + cvol + coef u^pw (the pointwise and
                    // nonlinear parts of a single instance)
                    T r = (T)P.c;
                    if (rx.cvol) r += rx.cvol[e * 9 + qx + 3 * qy];
                    if (rx.pw) {
                        T pw = rx.coef;
                        for (int t = 0; t < rx.pw; ++t) pw *= uq;
                        r += pw;
                    }
                    fm[qy][qx] = W * r * uq;
                }
            // transpose: y then x
#pragma unroll
            for (int qx = 0; qx < 3; ++qx)
#pragma unroll
                for (int iy = 0; iy < 3; ++iy) {
                    T a = 0, b = 0;
#pragma unroll
                    for (int qy = 0; qy < 3; ++qy) {
                        a += (T)tb.L[qy][iy] * (fx[qy][qx]);
                        b += (T)tb.Ld[qy][iy] * fy[qy][qx] + (T)tb.L[qy][iy] * fm[qy][qx];
                    }
                    t0[iy][qx] = a;   // to be contracted with Ld in x
                    t1[iy][qx] = b;   // with L in x
                }
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix) {
                    T a = 0;
#pragma unroll
                    for (int qx = 0; qx < 3; ++qx) a += (T)tb.Ld[qx][ix] * t0[iy][qx] + (T)tb.L[qx][ix] * t1[iy][qx];
                    wl[iy][ix] = a;
                }
        } else {
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix) {
                    double a = 0;
                    for (int qy = 0; qy < 3; ++qy)
                        for (int qx = 0; qx < 3; ++qx) {
                            const double W = tb.w[qx] * tb.w[qy] * detJ;
                            const double k = kvol ? (double)kvol[e * 9 + qx + 3 * qy] : P.kappa_const;
                            const double lx = tb.L[qx][ix], ly = tb.L[qy][iy];
                            if (mode == CG_MODE_DIAG) {
                                const double gx = tb.Ld[qx][ix] * ly / hx, gy = lx * tb.Ld[qy][iy] / hy;
                                const double r = P.c + (rx.cvol ? (double)rx.cvol[e * 9 + qx + 3 * qy] : 0.0);
                                a += W * (k * (gx * gx + gy * gy) + r * lx * lx * ly * ly);
                            } else {
                                a += W * k * lx * ly;  // kvol carries f for the load vector
                            }
                        }
                    wl[iy][ix] = (T)a;
                }
        }
    }
    // coloured accumulation into the LDS node lattice
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        if (valid && (tx & 1) == (col & 1) && (ty & 1) == (col >> 1)) {
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix) lat[(2 * ty + iy) * CG_LX + 2 * tx + ix] += wl[iy][ix];
        }
        __syncthreads();
    }
    // owned nodes
    const int I0 = 2 * ex0, J0 = 2 * ey0;
    const int I1 = (ex0 + TX >= P.nx) ? 2 * P.nx + 1 : 2 * ex0 + 2 * TX;
    const int J1 = (ey0 + TY >= P.ny) ? 2 * P.ny + 1 : 2 * ey0 + 2 * TY;
    const int wx = I1 - I0, wy = J1 - J0;
    T *o = out + P.dof_off;
    for (int k = threadIdx.x; k < wx * wy; k += blockDim.x) {
        const int I = I0 + k % wx, J = J0 + k / wx;
        const int64_t node = (int64_t)J * NX + I;
        const bool bnd = I == 0 || J == 0 || I == 2 * P.nx || J == 2 * P.ny;
        T val = lat[(J - J0 + 2) * CG_LX + (I - I0 + 2)];
        if (mode == CG_MODE_RHS_ADD) {
            // accumulate into the interior rows; the Dirichlet rows keep what
            // the apply wrote there
            if (!bnd) o[node] += val;
            continue;
        }
        if (bnd) val = mode == CG_MODE_APPLY ? u[node] : mode == CG_MODE_DIAG ? T(1) : T(0);
        o[node] = val;
    }
}

// out[e*9 + qx + 3 qy] = base + coef u(x_q)^pw (base may be null: 0): u at
// the volume quadrature points of a single instance, one thread per element
template <typename T>
__global__ void __launch_bounds__(256)
    k_cg_eval_qp(int nx, int ny, const CGTab *__restrict__ tabp, const T *__restrict__ u, const T *__restrict__ base,
                 T coef, int pw, T *__restrict__ out)
{
    const int64_t E = (int64_t)nx * ny;
    const CGTab &tb = *tabp;
    const int NX = 2 * nx + 1;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const int ex = (int)(e % nx), ey = (int)(e / nx);
        T ul[3][3];
        for (int iy = 0; iy < 3; ++iy)
            for (int ix = 0; ix < 3; ++ix) ul[iy][ix] = u[(int64_t)(2 * ey + iy) * NX + 2 * ex + ix];
        for (int qy = 0; qy < 3; ++qy)
            for (int qx = 0; qx < 3; ++qx) {
                T uq = 0;
                for (int iy = 0; iy < 3; ++iy)
                    for (int ix = 0; ix < 3; ++ix) uq += (T)tb.L[qy][iy] * (T)tb.L[qx][ix] * ul[iy][ix];
                T v = coef;
                for (int t = 0; t < pw; ++t) v *= uq;
                out[e * 9 + qx + 3 * qy] = (base ? base[e * 9 + qx + 3 * qy] : T(0)) + v;
            }
    }
}

template <typename T>
void cg_eval_qp(int nx, int ny, const CGTab *d_tab, const T *u, const T *base, T coef, int pw, T *out, hipStream_t s)
{
    const int64_t E = (int64_t)nx * ny;
    const unsigned G = (unsigned)std::min<int64_t>((E + 255) / 256, 4096);
    hipLaunchKernelGGL(k_cg_eval_qp<T>, dim3(G), dim3(256), 0, s, nx, ny, d_tab, u, base, coef, pw, out);
    FEM_HIP(hipGetLastError());
}
template void cg_eval_qp<double>(int, int, const CGTab *, const double *, const double *, double, int, double *,
                                 hipStream_t);
template void cg_eval_qp<float>(int, int, const CGTab *, const float *, const float *, float, int, float *,
                                hipStream_t);

template <typename T>
void cg_launch_shape(int shape, int mode, const CGInst *d_inst, int ninst, int total_tiles, const CGTab *d_tab,
                     const T *kvol, const T *in, T *out, const int *skip, const int *inst_done, hipStream_t s,
                     const CGReaction<T> &rx)
{
    FEM_CHECK(ninst == 1 || (!rx.cvol && !rx.pw), FEM_ERR_ARG, "pointwise/nonlinear reaction: a single instance");
#define CGS(ID, TX, TY)                                                                                       \
    case ID:                                                                                                 \
        hipLaunchKernelGGL((k_cg<TX, TY, T>), dim3(total_tiles), dim3((TX + 1) * (TY + 1)), 0, s, mode, d_inst, \
                           ninst, d_tab, kvol, in, out, skip, inst_done, rx);                                \
        break;
    switch (shape) {
        CGS(0, 15, 15) CGS(1, 7, 7) CGS(2, 15, 7) CGS(3, 31, 7) CGS(4, 7, 31)
    default: throw Error(FEM_ERR_ARG, "tile shape");
    }
#undef CGS
    FEM_HIP(hipGetLastError());
}

template <typename T>
void cg_launch(int mode, const CGInst *d_inst, int ninst, int total_tiles, const CGTab *d_tab, const T *kvol,
               const T *in, T *out, const int *skip, hipStream_t s, const CGReaction<T> &rx)
{
    cg_launch_shape<T>(0, mode, d_inst, ninst, total_tiles, d_tab, kvol, in, out, skip, nullptr, s, rx);
}

template void cg_launch<double>(int, const CGInst *, int, int, const CGTab *, const double *, const double *,
                                double *, const int *, hipStream_t, const CGReaction<double> &);
template void cg_launch<float>(int, const CGInst *, int, int, const CGTab *, const float *, const float *,
                               float *, const int *, hipStream_t, const CGReaction<float> &);
template void cg_launch_shape<double>(int, int, const CGInst *, int, int, const CGTab *, const double *,
                                      const double *, double *, const int *, const int *, hipStream_t,
                                      const CGReaction<double> &);
template void cg_launch_shape<float>(int, int, const CGInst *, int, int, const CGTab *, const float *,
                                     const float *, float *, const int *, const int *, hipStream_t,
                                     const CGReaction<float> &);

}  // namespace femamd
