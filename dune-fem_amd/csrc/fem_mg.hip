// Multigrid preconditioner for the SIPG DG operator (cellwise or constant
// kappa, no advection; a reaction c and a pointwise reaction cvol at the
// volume quadrature points, as the Jacobians of the nonlinear operators
// have, are taken into the levels: c as it is, cvol by its element means on
// the coarse DG levels, and from there by the Galerkin products into the
// vertex levels).  No reference counterpart: dune-fem
// reaches multigrid only through external packages (the "amg-jacobi"
// preconditioning method of dune/fem/solver/parameter.hh:48-59, an assembled-
// matrix AMG of dune-istl or PETSc).  This one is matrix-free and geometric.
//
// Levels (one symmetric V-cycle per application, so it is a fixed SPD
// operator and may precondition CG):
//   * DG levels Q_p -> Q_{p/2} -> ... -> Q_1 on the fine mesh.  The
//     orthonormal Legendre basis is hierarchical, so the prolongation embeds
//     the low modes and the restriction of a residual keeps them; with the
//     penalty factor of the fine operator the rediscretised Q_{p'} operator
//     equals the Galerkin product P^T A P exactly (the (p'+1)-point rules
//     integrate the Q_{p'} products exactly).  Smoother: damped exact
//     element-block Jacobi, nu sweeps before and after.
//   * The conforming trilinear space on the (n+1)^3 vertices, embedded in
//     DG Q1 (per element the 8 Legendre modes of the trilinear interpolant).
//     Its Galerkin operator P_c^T A_Q1 P_c is a 27-point stencil (interior
//     face terms vanish on continuous functions); it is computed on the
//     device in closed form from the Q1 level's coefficients.  Smoother:
//     damped Jacobi.
//   * Geometric vertex levels: every second vertex (and the last one) is
//     kept per direction, trilinear interpolation, Galerkin stencils
//     P^T A P formed directly from the finer stencil, down to at most 512
//     vertices, which are solved with a dense
//     inverse (Gauss-Jordan on the device; on the host, with pivoting, for
//     an indefinite operator).
// The conforming coarse space is what makes the method robust for the
// high-contrast coefficients of the configurations: element means (Q_0) as
// the coarse space of DG Q1 leave far more iterations (measured on the CPU
// on a 12^3 channel problem: 69 PCG iterations with an exact Q_0 coarse
// solve against 27 with the exact conforming one).
#include <cmath>
#include <type_traits>

#include "fem_blas.hpp"
#include "fem_dg_bj.hpp"
#include "fem_mg.hpp"

namespace femamd {

namespace {

constexpr int MGT = 256;   // threads of the grid-stride level kernels
constexpr int MG_DENSE_MAX = 512;

inline unsigned mg_grid(int64_t n) { return grid_for(n, MGT, 16384); }

#define MG_SKIP \
    if (skip && *skip) return

// ---- DG levels ----------------------------------------------------------------------
// mode 0: x = omega B^-1 r;  mode 1: x += omega B^-1 (r - t), t = A x.  B^-1 is
// the exact element-block inverse (fem_dg_bj.hpp); EPB elements of K*K lanes
// per wavefront, extra lanes shadow element 0 without writing.
template <int K, typename T>
__global__ void __launch_bounds__(64)
    k_mg_bj(int64_t E, const T *__restrict__ Q, const T *__restrict__ lam, const T *__restrict__ r,
            const T *__restrict__ t, T *__restrict__ x, T omega, int mode, const int *__restrict__ skip)
{
    using S = BJShape<K>;
    constexpr int K2 = S::K2, K3 = S::K3, EPB = S::EPB;
    MG_SKIP;
    __shared__ T smem[EPB * K3];
    __shared__ T qsm[EPB * (3 * K2 + 3 * K)];
    const int tid = threadIdx.x, el0 = tid / K2, tl = tid % K2, i = tl / K, j = tl % K;
    const bool wr = el0 < EPB;
    const int el = wr ? el0 : 0;
    const int64_t e_raw = (int64_t)blockIdx.x * EPB + el;
    const bool active = wr && e_raw < E;
    const int64_t e = e_raw < E ? e_raw : 0;
    const int64_t base = e * K3 + i * K2 + j * K;
    T v[K];
#pragma unroll
    for (int c = 0; c < K; ++c) v[c] = mode ? r[base + c] - t[base + c] : r[base + c];
    bj_element<K>(Q + e * 3 * K2, lam + e * 3 * K, smem + el * K3, qsm + el * (3 * K2 + 3 * K), i, j, v);
    if (active)
#pragma unroll
        for (int c = 0; c < K; ++c) x[base + c] = mode ? x[base + c] + omega * v[c] : omega * v[c];
}

// rc = (r - t) on the modes a, b, c < Kc (local index a Kf^2 + b Kf + c)
template <typename T>
__global__ void k_mg_restrict_modes(int64_t E, int Kf, int Kc, const T *__restrict__ r, const T *__restrict__ t,
                                    T *__restrict__ rc, const int *__restrict__ skip)
{
    MG_SKIP;
    const int Kc3 = Kc * Kc * Kc, Kf3 = Kf * Kf * Kf;
    const int64_t n = E * Kc3;
    const bool narrow = n < (int64_t(1) << 31);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = narrow ? (int64_t)((unsigned)q / (unsigned)Kc3) : q / Kc3;
        const int l = (int)(q - e * Kc3);
        const int a = l / (Kc * Kc), b = (l / Kc) % Kc, c = l % Kc;
        const int64_t f = e * Kf3 + a * Kf * Kf + b * Kf + c;
        rc[q] = r[f] - t[f];
    }
}

// xf += P xc (embedding of the low modes)
template <typename T>
__global__ void k_mg_prolong_modes(int64_t E, int Kf, int Kc, const T *__restrict__ xc, T *__restrict__ xf,
                                   const int *__restrict__ skip)
{
    MG_SKIP;
    const int Kc3 = Kc * Kc * Kc, Kf3 = Kf * Kf * Kf;
    const int64_t n = E * Kc3;
    const bool narrow = n < (int64_t(1) << 31);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = narrow ? (int64_t)((unsigned)q / (unsigned)Kc3) : q / Kc3;
        const int l = (int)(q - e * Kc3);
        const int a = l / (Kc * Kc), b = (l / Kc) % Kc, c = l % Kc;
        xf[e * Kf3 + a * Kf * Kf + b * Kf + c] += xc[q];
    }
}

// ---- conforming trilinear <-> DG Q1 ------------------------------------------------------
// The 1D nodal functions 1 - s and s in the orthonormal Legendre modes
// phi_0 = 1, phi_1 = sqrt(3) (2 s - 1): coefficient of mode a for vertex i
__device__ __forceinline__ double tq1(int a, int i)
{
    return a == 0 ? 0.5 : (i ? 0.28867513459481287 : -0.28867513459481287);   // 1 / (2 sqrt 3)
}

// rv = P_c^T (r - t) (t may be null), in two passes.  First, per element,
// the trilinear functional of its Q1 modes at each of its 8 vertices,
// c[e][q] with q = i + 2 j + 4 k for the vertex (cx + i, cy + j, cz + k),
// kept in FP64; then per vertex the sum over the (up to) 8 elements around
// it.  A single per-vertex pass would gather 8 x 8 mode values per vertex.
template <typename T>
__global__ void k_mg_dg1_corners(int64_t E, const T *__restrict__ r, const T *__restrict__ t,
                                 double *__restrict__ c, const int *__restrict__ skip)
{
    MG_SKIP;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < E * 8; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = q >> 3;
        const int cq = (int)(q & 7), i = cq & 1, j = (cq >> 1) & 1, k = cq >> 2;
        double acc = 0;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const int a = l >> 2, b = (l >> 1) & 1, cc = l & 1;
            const double val = (double)r[e * 8 + l] - (t ? (double)t[e * 8 + l] : 0.0);
            acc = fma(tq1(a, i) * tq1(b, j) * tq1(cc, k), val, acc);
        }
        c[q] = acc;
    }
}

template <typename T>
__global__ void k_mg_corners_to_vertex(int nx, int ny, int nz, const double *__restrict__ c, T *__restrict__ rv,
                                       const int *__restrict__ skip)
{
    MG_SKIP;
    const int nvx = nx + 1, nvy = ny + 1;
    const int64_t V = (int64_t)nvx * nvy * (nz + 1);
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int vx = (int)(v % nvx), vy = (int)((v / nvx) % nvy), vz = (int)(v / ((int64_t)nvx * nvy));
        double acc = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = q & 1, j = (q >> 1) & 1, k = q >> 2;
            const int cx = vx - i, cy = vy - j, cz = vz - k;
            const bool in = cx >= 0 && cx < nx && cy >= 0 && cy < ny && cz >= 0 && cz < nz;
            const int64_t e = in ? cx + (int64_t)nx * (cy + (int64_t)ny * cz) : 0;
            const double val = c[e * 8 + q];
            acc += in ? val : 0.0;
        }
        rv[v] = (T)acc;
    }
}

template <typename T>
void dg1_to_vertex(const int *n, const T *r, const T *t, double *corners, T *rv, const int *skip, hipStream_t s)
{
    const int64_t E = (int64_t)n[0] * n[1] * n[2];
    const int64_t V = (int64_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    hipLaunchKernelGGL(k_mg_dg1_corners<T>, dim3(mg_grid(E * 8)), dim3(MGT), 0, s, E, r, t, corners, skip);
    hipLaunchKernelGGL(k_mg_corners_to_vertex<T>, dim3(mg_grid(V)), dim3(MGT), 0, s, n[0], n[1], n[2],
                       (const double *)corners, rv, skip);
    FEM_HIP(hipGetLastError());
}

// x += P_c xv: the Q1 modes of the trilinear interpolant of the vertex values
template <typename T>
__global__ void k_mg_vertex_to_dg1(int nx, int ny, int nz, const T *__restrict__ xv, T *__restrict__ x,
                                   const int *__restrict__ skip)
{
    MG_SKIP;
    const int nvx = nx + 1, nvy = ny + 1;
    const int64_t n = (int64_t)nx * ny * nz * 8;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = q >> 3;
        const int l = (int)(q & 7), a = l >> 2, b = (l >> 1) & 1, c = l & 1;
        const int cx = (int)(e % nx), cy = (int)((e / nx) % ny), cz = (int)(e / ((int64_t)nx * ny));
        double acc = 0;
        for (int k = 0; k < 2; ++k)
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i)
                    acc += tq1(a, i) * tq1(b, j) * tq1(c, k) *
                           (double)xv[(cx + i) + (int64_t)nvx * ((cy + j) + (int64_t)nvy * (cz + k))];
        x[q] = (T)((double)x[q] + acc);
    }
}

// ---- vertex levels -------------------------------------------------------------------
struct VDims {
    int n[3];
};

// t = sg * (A x) + (r ? r : 0) with sg = -1 (residual r - A x) or +1 (A x,
// r null).  The stencil is stored by offset, A[s * V + v] with s = (dx+1) +
// 3 (dy+1) + 9 (dz+1), so each of the 27 coefficient loads is coalesced; the
// 27 terms are unrolled, and a neighbour outside the lattice reads x[v] with
// weight 0, so all loads are issued together.
template <typename T>
__global__ void k_mg_v_apply(VDims d, const T *__restrict__ A, const T *__restrict__ x, const T *__restrict__ r,
                             T *__restrict__ t, int sg, const int *__restrict__ skip)
{
    MG_SKIP;
    const int64_t V = (int64_t)d.n[0] * d.n[1] * d.n[2];
    const int64_t sy = d.n[0], sz = (int64_t)d.n[0] * d.n[1];
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int vx = (int)(v % d.n[0]), vy = (int)((v / d.n[0]) % d.n[1]), vz = (int)(v / sz);
        T acc = 0;
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
            const bool in = vx + dx >= 0 && vx + dx < d.n[0] && vy + dy >= 0 && vy + dy < d.n[1] && vz + dz >= 0 &&
                            vz + dz < d.n[2];
            const T a = A[s * V + v];
            const T xv = x[in ? v + dx + dy * sy + dz * sz : v];
            acc = fma(in ? a : T(0), xv, acc);
        }
        t[v] = (r ? r[v] : T(0)) + (sg < 0 ? -acc : acc);
    }
}

// the stencil from one row per vertex (A[v * 27 + s]) to one row per offset
__global__ void k_mg_by_offset(int64_t V, const double *__restrict__ Av, double *__restrict__ As)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < V * 27; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = q / 27;
        const int s = (int)(q - v * 27);
        As[s * V + v] = Av[q];
    }
}

// mode 0: x = omega dinv r;  mode 1: x += omega dinv t (t = r - A x)
template <typename T>
__global__ void k_mg_v_jacobi(int64_t V, const T *__restrict__ dinv, const T *__restrict__ r,
                              const T *__restrict__ t, T *__restrict__ x, T omega, int mode,
                              const int *__restrict__ skip)
{
    MG_SKIP;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        x[v] = mode ? x[v] + omega * dinv[v] * t[v] : omega * dinv[v] * r[v];
}

// 1D trilinear interpolation from the coarse vertices (fine positions 0, 2,
// 4, ..., and the last fine vertex) to fine vertex i: weight of coarse J.
// The last coarse vertex, J = ceil((nf - 1) / 2), is (i >> 1) or (i >> 1) + 1
// for i = nf - 1, so the two candidates of a fine vertex always cover it.
__device__ __forceinline__ double w1(int i, int J, int nf, int nc)
{
    if (i == nf - 1) return J == nc - 1 ? 1.0 : 0.0;
    if ((i & 1) == 0) return J == (i >> 1) ? 1.0 : 0.0;
    return (J == (i >> 1) || J == (i >> 1) + 1) ? 0.5 : 0.0;
}

// rc = P^T tf
template <typename T>
__global__ void k_mg_v_restrict(VDims f, VDims cdim, const T *__restrict__ tf, T *__restrict__ rc,
                                const int *__restrict__ skip)
{
    MG_SKIP;
    const int64_t Vc = (int64_t)cdim.n[0] * cdim.n[1] * cdim.n[2];
    for (int64_t J = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; J < Vc; J += (int64_t)gridDim.x * blockDim.x) {
        const int Jx = (int)(J % cdim.n[0]), Jy = (int)((J / cdim.n[0]) % cdim.n[1]),
                  Jz = (int)(J / ((int64_t)cdim.n[0] * cdim.n[1]));
        // per direction the fine vertices 2J-1, 2J, 2J+1 (in range) and
        // their weights; the last fine vertex, the only other candidate,
        // has weight 0 unless it is one of them (coarse1 puts the last
        // coarse vertex at or beyond it)
        int ix[3], iy[3], iz[3];
        double wx[3], wy[3], wz[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int fx = 2 * Jx - 1 + m, fy = 2 * Jy - 1 + m, fz = 2 * Jz - 1 + m;
            const bool okx = fx >= 0 && fx < f.n[0], oky = fy >= 0 && fy < f.n[1], okz = fz >= 0 && fz < f.n[2];
            ix[m] = okx ? fx : 0;
            iy[m] = oky ? fy : 0;
            iz[m] = okz ? fz : 0;
            wx[m] = okx ? w1(fx, Jx, f.n[0], cdim.n[0]) : 0.0;
            wy[m] = oky ? w1(fy, Jy, f.n[1], cdim.n[1]) : 0.0;
            wz[m] = okz ? w1(fz, Jz, f.n[2], cdim.n[2]) : 0.0;
        }
        double acc = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    acc = fma(wx[a] * wy[b] * wz[c],
                              (double)tf[ix[a] + (int64_t)f.n[0] * (iy[b] + (int64_t)f.n[1] * iz[c])], acc);
        rc[J] = (T)acc;
    }
}

// xf += P xc
template <typename T>
__global__ void k_mg_v_prolong(VDims f, VDims cdim, const T *__restrict__ xc, T *__restrict__ xf,
                               const int *__restrict__ skip)
{
    MG_SKIP;
    const int64_t Vf = (int64_t)f.n[0] * f.n[1] * f.n[2];
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < Vf; v += (int64_t)gridDim.x * blockDim.x) {
        const int ix = (int)(v % f.n[0]), iy = (int)((v / f.n[0]) % f.n[1]), iz = (int)(v / ((int64_t)f.n[0] * f.n[1]));
        double acc = 0;
        for (int c = 0; c < 2; ++c) {
            const int Jz = (iz >> 1) + c;
            if (Jz >= cdim.n[2]) continue;
            const double wz = w1(iz, Jz, f.n[2], cdim.n[2]);
            if (wz == 0.0) continue;
            for (int b = 0; b < 2; ++b) {
                const int Jy = (iy >> 1) + b;
                if (Jy >= cdim.n[1]) continue;
                const double wy = w1(iy, Jy, f.n[1], cdim.n[1]);
                if (wy == 0.0) continue;
                for (int a = 0; a < 2; ++a) {
                    const int Jx = (ix >> 1) + a;
                    if (Jx >= cdim.n[0]) continue;
                    const double wx = w1(ix, Jx, f.n[0], cdim.n[0]);
                    if (wx == 0.0) continue;
                    acc += wx * wy * wz * (double)xc[Jx + (int64_t)cdim.n[0] * (Jy + (int64_t)cdim.n[1] * Jz)];
                }
            }
        }
        xf[v] = (T)((double)xf[v] + acc);
    }
}

// x = Ainv r with Ainv row-major: one wavefront per row (four per block),
// the lanes striding its columns (coalesced), then a butterfly sum.  The
// row index is the same on all lanes of a wavefront, so a wavefront past
// the last row leaves as a whole.
template <typename T>
__global__ void __launch_bounds__(256)
    k_mg_dense(int V, const T *__restrict__ Ainv, const T *__restrict__ r, T *__restrict__ x,
               const int *__restrict__ skip)
{
    MG_SKIP;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= V) return;
    const T *a = Ainv + (int64_t)row * V;
    double acc = 0;
    for (int w = lane; w < V; w += 64) acc = fma((double)a[w], (double)r[w], acc);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) x[row] = (T)acc;
}

// ---- set-up (FP64) -----------------------------------------------------------------------

// the Gauss-weighted element means of a pointwise coefficient [E][K^3]
// (grid-stride: the level kernels' grid is capped)
__global__ void k_mg_cvol_mean(int64_t E, int K, const Tab1D<double> *__restrict__ tab,
                               const double *__restrict__ cvol, double *__restrict__ mean)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const double *c = cvol + e * K * K * K;
        double s = 0;
        for (int qz = 0; qz < K; ++qz)
            for (int qy = 0; qy < K; ++qy)
                for (int qx = 0; qx < K; ++qx)
                    s += tab->w[qx] * tab->w[qy] * tab->w[qz] * c[qx + K * (qy + K * qz)];
        mean[e] = s;
    }
}

// a coarse level's pointwise coefficient: the element mean at each of its points
__global__ void k_mg_cvol_fill(int64_t E, int K3, const double *__restrict__ mean, double *__restrict__ cvol)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < E * K3; i += (int64_t)gridDim.x * blockDim.x)
        cvol[i] = mean[i / K3];
}


// row v of the stencil: the probe of colour col answers for the one stencil
// neighbour of that colour (distinct colours within distance 1)

// the conforming level's stencil in closed form: a(phi_W, phi_v) of the DG Q1
// operator `a` on the trilinear hat functions of the (n+1)^3 vertices.  The
// interior face terms of SIPG vanish on continuous functions (no jumps; the
// flux terms against [v] cancel between the two sides), which leaves per
// element the stiffness kappa sum_d (h_t1 h_t2 / h_d) k(a_d, b_d) m(a_t1,
// b_t1) m(a_t2, b_t2) (k = +-1, m = 1/3, 1/6: the 1D linear stiffness and mass
// on [0, 1]), the reaction (c + cvol) at the 2^3 Gauss points, and on the
// domain boundary the Nitsche terms kappa (beta/h_d phi_a phi_b - n_d
// (d_d phi_a phi_b + phi_a d_d phi_b)) over the face.  One thread per vertex
// v and stencil offset s (W = v + d(s)); a, b are the local corners of W and
// v in each element that holds both.
__device__ __forceinline__ double mg_m1(int a, int b) { return a == b ? (1.0 / 3.0) : (1.0 / 6.0); }

__global__ void k_mg_conforming(const DGArgs<double> a, VDims vd, double *__restrict__ A)
{
    const int64_t V = (int64_t)vd.n[0] * vd.n[1] * vd.n[2];
    const Tab1D<double> &tb = *a.tab;
    const double detJ = a.h[0] * a.h[1] * a.h[2];
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < V * 27; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = g / 27;
        const int so = (int)(g - v * 27);
        const int P[3] = { (int)(v % vd.n[0]), (int)((v / vd.n[0]) % vd.n[1]), (int)(v / ((int64_t)vd.n[0] * vd.n[1])) };
        const int W[3] = { P[0] + so % 3 - 1, P[1] + (so / 3) % 3 - 1, P[2] + so / 9 - 1 };
        double acc = 0;
        if (W[0] >= 0 && W[1] >= 0 && W[2] >= 0 && W[0] < vd.n[0] && W[1] < vd.n[1] && W[2] < vd.n[2]) {
            for (int ez = P[2] - 1; ez <= P[2]; ++ez)
                for (int ey = P[1] - 1; ey <= P[1]; ++ey)
                    for (int ex = P[0] - 1; ex <= P[0]; ++ex) {
                        const int ec[3] = { ex, ey, ez };
                        int ca[3], cb[3];
                        bool in = true;
                        for (int d = 0; d < 3; ++d) {
                            cb[d] = P[d] - ec[d];
                            ca[d] = W[d] - ec[d];
                            in = in && ec[d] >= 0 && ec[d] < a.n[d] && ca[d] >= 0 && ca[d] <= 1;
                        }
                        if (!in) continue;
                        const int64_t e = ex + (int64_t)a.n[0] * (ey + (int64_t)a.n[1] * ez);
                        const double kap = a.kappa_mode == FEM_KAPPA_CELL ? a.kcell[e] : a.kappa_const;
                        for (int d = 0; d < 3; ++d) {
                            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                            const double mt = mg_m1(ca[t1], cb[t1]) * mg_m1(ca[t2], cb[t2]);
                            const double area = a.h[t1] * a.h[t2];
                            acc += kap * area / a.h[d] * (ca[d] == cb[d] ? 1.0 : -1.0) * mt;
                            for (int sd = 0; sd < 2; ++sd) {
                                if (ec[d] != (sd ? a.n[d] - 1 : 0)) continue;   // boundary faces only
                                const double nrm = sd ? 1.0 : -1.0;
                                const double va = ca[d] == sd ? 1.0 : 0.0, vb = cb[d] == sd ? 1.0 : 0.0;
                                const double da = ca[d] ? 1.0 : -1.0, db = cb[d] ? 1.0 : -1.0;
                                acc += area * mt * kap * (a.beta / a.h[d] * va * vb - nrm / a.h[d] * (da * vb + va * db));
                            }
                        }
                        for (int qz = 0; qz < 2; ++qz)
                            for (int qy = 0; qy < 2; ++qy)
                                for (int qx = 0; qx < 2; ++qx) {
                                    const int q[3] = { qx, qy, qz };
                                    double pa = 1, pb = 1;
                                    for (int d = 0; d < 3; ++d) {
                                        const double xq = tb.x[q[d]];
                                        pa *= ca[d] ? xq : 1.0 - xq;
                                        pb *= cb[d] ? xq : 1.0 - xq;
                                    }
                                    const double r = a.c + (a.cvol ? a.cvol[e * 8 + qx + 2 * qy + 4 * qz] : 0.0);
                                    acc += tb.w[qx] * tb.w[qy] * tb.w[qz] * detJ * r * pa * pb;
                                }
                    }
        }
        A[g] = acc;
    }
}

// the Galerkin product P^T A_f P of a vertex level, directly: one thread per
// coarse vertex V and stencil offset s (W = V + d(s)) sums
// P(i, V) A_f(i, j) P(j, W) over the fine vertices i in the support of V
// (per direction 2V-1 .. 2V+1, which holds the last fine vertex when it
// belongs to V) and the 27 stencil neighbours j of i.  Offsets that leave
// the coarse lattice hold 0.
__global__ void k_mg_galerkin(VDims f, VDims cd, const double *__restrict__ Af, double *__restrict__ Ac)
{
    const int64_t Vc = (int64_t)cd.n[0] * cd.n[1] * cd.n[2];
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Vc * 27; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = g / 27;
        const int so = (int)(g - v * 27);
        const int V[3] = { (int)(v % cd.n[0]), (int)((v / cd.n[0]) % cd.n[1]), (int)(v / ((int64_t)cd.n[0] * cd.n[1])) };
        const int W[3] = { V[0] + so % 3 - 1, V[1] + (so / 3) % 3 - 1, V[2] + so / 9 - 1 };
        double acc = 0;
        if (W[0] >= 0 && W[1] >= 0 && W[2] >= 0 && W[0] < cd.n[0] && W[1] < cd.n[1] && W[2] < cd.n[2]) {
            // the 1D weights of V and W at the fine vertices 2V-1 .. 2V+1 and
            // 2W-1 .. 2W+1 (the supports of both; 0 outside the lattice)
            double wv[3][3], ww[3][3];
            for (int d = 0; d < 3; ++d)
                for (int m = 0; m < 3; ++m) {
                    const int fv = 2 * V[d] - 1 + m, fw = 2 * W[d] - 1 + m;
                    wv[d][m] = (fv >= 0 && fv < f.n[d]) ? w1(fv, V[d], f.n[d], cd.n[d]) : 0.0;
                    ww[d][m] = (fw >= 0 && fw < f.n[d]) ? w1(fw, W[d], f.n[d], cd.n[d]) : 0.0;
                }
            const int sh[3] = { 2 * (V[0] - W[0]), 2 * (V[1] - W[1]), 2 * (V[2] - W[2]) };
            for (int az = 0; az < 3; ++az) {
                if (wv[2][az] == 0.0) continue;
                const int iz = 2 * V[2] - 1 + az;
                for (int ay = 0; ay < 3; ++ay) {
                    if (wv[1][ay] == 0.0) continue;
                    const int iy = 2 * V[1] - 1 + ay;
                    for (int ax = 0; ax < 3; ++ax) {
                        if (wv[0][ax] == 0.0) continue;
                        const int ix = 2 * V[0] - 1 + ax;
                        const double *a = Af + (ix + (int64_t)f.n[0] * (iy + (int64_t)f.n[1] * iz)) * 27;
                        double row = 0;
                        // neighbour j = i + t - 1 per direction sits at index
                        // b = j - (2W - 1) = a + sh + t - 1 of W's support
                        for (int tz = 0; tz < 3; ++tz) {
                            const int bz = az + sh[2] + tz - 1;
                            if (bz < 0 || bz > 2 || ww[2][bz] == 0.0) continue;
                            for (int ty = 0; ty < 3; ++ty) {
                                const int by = ay + sh[1] + ty - 1;
                                if (by < 0 || by > 2 || ww[1][by] == 0.0) continue;
                                const double pzy = ww[2][bz] * ww[1][by];
                                for (int tx = 0; tx < 3; ++tx) {
                                    const int bx = ax + sh[0] + tx - 1;
                                    if (bx < 0 || bx > 2) continue;
                                    row = fma(a[tx + 3 * ty + 9 * tz], pzy * ww[0][bx], row);
                                }
                            }
                        }
                        acc = fma(wv[2][az] * wv[1][ay] * wv[0][ax], row, acc);
                    }
                }
            }
        }
        Ac[g] = acc;
    }
}

// the Jacobi scaling of a vertex level: 1 / max(a_vv, lim sum_w |a_vw|).  By
// Gershgorin the eigenvalues of diag(d)^-1 A are then at most 1 / lim, so a
// sweep damped by omega = 1.9 lim stays convergent.  A diffusion stencil
// (off-diagonal sum -a_vv) keeps its diagonal; a reaction-dominated one (the
// Q1 mass stencil has row sum 3.375 a_vv) is scaled down instead of
// diverging.
__global__ void k_mg_dinv(int64_t V, const double *__restrict__ A, double lim, double *__restrict__ dinv)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        double rs = 0;
        for (int j = 0; j < 27; ++j) rs += fabs(A[v * 27 + j]);
        dinv[v] = 1.0 / fmax(A[v * 27 + 13], lim * rs);
    }
}

// the coarsest level's operator as a dense row-major V x V matrix from its
// 27-point stencil (offset s = sx + 3 sy + 9 sz, 0..2 each, about the vertex)
__global__ void k_mg_dense_fill(VDims d, const double *__restrict__ A, double *__restrict__ M)
{
    const int64_t V = (int64_t)d.n[0] * d.n[1] * d.n[2];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V * V; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i / V, w = i - v * V;
        const int dx = (int)(w % d.n[0]) - (int)(v % d.n[0]);
        const int dy = (int)((w / d.n[0]) % d.n[1]) - (int)((v / d.n[0]) % d.n[1]);
        const int dz = (int)(w / ((int64_t)d.n[0] * d.n[1])) - (int)(v / ((int64_t)d.n[0] * d.n[1]));
        const bool in = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
        M[i] = in ? A[v * 27 + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] : 0.0;
    }
}

// one pivot step of the in-place Gauss-Jordan inversion without pivoting
// (X -> Y, pivot c): Y_cc = 1/p, Y_ck = X_ck / p, Y_rc = -X_rc / p, Y_rk =
// X_rk - X_rc X_ck / p, p = X_cc.  After the steps c = 0..V-1 the matrix is
// the inverse.  The pivots are the leading Schur complements' diagonals:
// all positive exactly when the matrix is symmetric positive definite, for
// which no pivoting is needed; they are recorded for the host to check.
__device__ __forceinline__ double gj_entry(const double *__restrict__ X, int V, int c, double ip, int r, int k)
{
    if (r == c) return k == c ? ip : X[(int64_t)c * V + k] * ip;
    if (k == c) return -X[(int64_t)r * V + c] * ip;
    return X[(int64_t)r * V + k] - X[(int64_t)r * V + c] * X[(int64_t)c * V + k] * ip;
}

// `steps` (1 or 2) pivot steps c, c + 1 in one launch: the entries of the
// intermediate matrix that step c + 1 reads (its own, row c + 1, column
// c + 1) are formed from X by the same operations as in a launch of their
// own, so the result equals two single steps bit for bit
__global__ void k_mg_gj_step(int V, int c, int steps, const double *__restrict__ X, double *__restrict__ Y,
                             double *__restrict__ piv)
{
    const int64_t V2 = (int64_t)V * V;
    const double pc = X[(int64_t)c * V + c];
    const double ip = 1.0 / pc;
    const int c1 = c + 1;
    const double pc1 = steps == 2 ? gj_entry(X, V, c, ip, c1, c1) : 1.0;
    const double ip1 = 1.0 / pc1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V2; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / V), k = (int)(i - (int64_t)r * V);
        double y;
        if (steps == 1) {
            y = gj_entry(X, V, c, ip, r, k);
        } else if (r == c1) {
            y = k == c1 ? ip1 : gj_entry(X, V, c, ip, c1, k) * ip1;
        } else if (k == c1) {
            y = -gj_entry(X, V, c, ip, r, c1) * ip1;
        } else {
            y = gj_entry(X, V, c, ip, r, k) - gj_entry(X, V, c, ip, r, c1) * gj_entry(X, V, c, ip, c1, k) * ip1;
        }
        if (i == 0) {
            piv[c] = pc;
            if (steps == 2) piv[c1] = pc1;
        }
        Y[i] = y;
    }
}

int coarse1(int nf) { return (nf - 1 + 1) / 2 + 1; }   // ceil((nf - 1) / 2) + 1

// the level's values in precision T
void store_as(int prec, DevBuf &dst, const DevBuf &src64, int64_t n, Context *ctx)
{
    dst.ensure(n * (prec == FEM_F64 ? 8 : 4));
    Vec a, b;
    a.ctx = b.ctx = ctx;
    a.n = b.n = n;
    a.prec = prec;
    b.prec = FEM_F64;
    a.data = dst.p;
    b.data = src64.p;
    a.owner = b.owner = false;
    vec_copy(a, b);
}

}  // namespace

void Multigrid::init(Operator *fine)
{
    op = fine;
    FEM_CHECK(fine->is_dg(), FEM_ERR_UNSUPPORTED, "multigrid: DG operators only");
    const fem_op_desc &d0 = fine->desc;
    FEM_CHECK(d0.b[0] == 0 && d0.b[1] == 0 && d0.b[2] == 0, FEM_ERR_UNSUPPORTED,
              "multigrid: operators without advection (b = 0)");
    FEM_CHECK(!fine->nonlinear(), FEM_ERR_STATE, "multigrid: the operator is nonlinear (build it on the Jacobian)");
    FEM_CHECK(d0.kappa_mode == FEM_KAPPA_CELL || d0.kappa_mode == FEM_KAPPA_CONST, FEM_ERR_UNSUPPORTED,
              "multigrid: cellwise or constant kappa");
    FEM_CHECK(fine->a64.kappa_mode >= 0, FEM_ERR_STATE, "coefficient data not set");
    Context *ctx = fine->ctx;
    desc = d0;
    reaction = fine->cvol64.p != nullptr;
    prec = fine->vec_prec;
    const int lp = prec == FEM_F64 ? FEM_F64 : FEM_F32;
    const int n[3] = { d0.n[0], d0.n[1], d0.n[2] };
    const int64_t E = fine->E;
    // a coarse operator: the finest one's description at another order and
    // precision, with buffers for the coefficients update() fills (cellwise
    // kappa; cvol as the element means)
    auto make_op = [&](int order, int precision) {
        std::unique_ptr<Operator> o(new Operator());
        fem_op_desc d = d0;
        d.order = order;
        d.precision = precision;
        o->init(ctx, d);
        if (d0.kappa_mode == FEM_KAPPA_CELL) {
            o->kcell64.alloc(sizeof(double) * E);
            o->kcell32.alloc(sizeof(float) * E);
            o->a64.kcell = o->kcell64.as<double>();
            o->a32.kcell = o->kcell32.as<float>();
            o->a64.kappa_mode = o->a32.kappa_mode = FEM_KAPPA_CELL;
        }
        if (reaction) {
            const int64_t nq = E * o->k * o->k * o->k;
            o->cvol64.alloc(sizeof(double) * nq);
            o->cvol32.alloc(sizeof(float) * nq);
            o->a64.cvol = o->cvol64.as<double>();
            o->a32.cvol = o->cvol32.as<float>();
        }
        return o;
    };
    const size_t es = lp == FEM_F64 ? 8 : 4;
    // DG levels: p, p/2, ..., 1
    for (int p = d0.order;; p = std::max(1, p / 2)) {
        std::unique_ptr<MGDGLevel> L(new MGDGLevel());
        if (dg.empty()) {
            L->op = fine;
            // BF16 rounds the quadrature values to 2^-8 of their size, which
            // is noise of order 4e-3 * lambda_max |x| in A x: far above the
            // smooth part of a smoothed residual, which is what the coarse
            // levels must see. The smoothers tolerate it; the restricted
            // residual is formed in FP32.
            if (d0.precision == FEM_BF16MFMA) L->resop = make_op(d0.order, FEM_F32);
        } else {
            L->own = make_op(p, lp);
            L->op = L->own.get();
            L->r.alloc(L->op->N * es);
            L->x.alloc(L->op->N * es);
        }
        L->k = p + 1;
        L->t.alloc(L->op->N * es);
        dg.push_back(std::move(L));
        if (p == 1) break;
    }
    corners.alloc(E * 8 * sizeof(double));
    // vertex levels: the conforming one, then every second vertex per
    // direction down to at most MG_DENSE_MAX vertices
    {
        std::unique_ptr<MGVertexLevel> V0(new MGVertexLevel());
        for (int d = 0; d < 3; ++d) V0->nv[d] = n[d] + 1;
        V0->V = (int64_t)V0->nv[0] * V0->nv[1] * V0->nv[2];
        vx.push_back(std::move(V0));
    }
    while (vx.back()->V > MG_DENSE_MAX) {
        const MGVertexLevel &F = *vx.back();
        std::unique_ptr<MGVertexLevel> C(new MGVertexLevel());
        for (int d = 0; d < 3; ++d) C->nv[d] = coarse1(F.nv[d]);
        C->V = (int64_t)C->nv[0] * C->nv[1] * C->nv[2];
        if (C->V == F.V) break;   // nothing left to coarsen
        vx.push_back(std::move(C));
    }
    FEM_CHECK(vx.back()->V <= 4096, FEM_ERR_UNSUPPORTED, "multigrid: coarsest vertex level too large");
    for (auto &L : vx) {
        A64.emplace_back(new DevBuf());
        A64.back()->alloc(L->V * 27 * 8);
        L->x.alloc(L->V * es);
        L->r.alloc(L->V * es);
        L->t.alloc(L->V * es);
    }
    d64.alloc(vx[0]->V * 8);
    As64.alloc(vx[0]->V * 27 * 8);
    const int64_t VC = vx.back()->V;
    gjX.alloc((size_t)VC * VC * 8);
    gjY.alloc((size_t)VC * VC * 8);
    gjp.alloc((size_t)VC * 8);
    if (reaction) cmean.alloc(sizeof(double) * E);
    update();
}

bool Multigrid::updatable_for(const Operator *o) const
{
    return o == op && (o->cvol64.p != nullptr) == reaction && !o->nonlinear();
}

void Multigrid::update()
{
    Operator *fine = op;
    Context *ctx = fine->ctx;
    hipStream_t s = ctx->stream;
    const int lp = prec == FEM_F64 ? FEM_F64 : FEM_F32;
    const int64_t E = fine->E;
    FEM_CHECK(fine->a64.kappa_mode >= 0, FEM_ERR_STATE, "coefficient data not set");
    FEM_CHECK((fine->cvol64.p != nullptr) == reaction, FEM_ERR_STATE,
              "multigrid: the pointwise reaction was added or removed since the set-up");
    auto round_into = [&](void *lo, const void *hi, int64_t nv) {
        Vec a, b;
        a.ctx = b.ctx = ctx;
        a.n = b.n = nv;
        a.prec = FEM_F32;
        b.prec = FEM_F64;
        a.data = lo;
        b.data = const_cast<void *>(hi);
        a.owner = b.owner = false;
        vec_copy(a, b);
    };
    // coarse coefficients: the cellwise kappa as it is, cvol by its
    // Gauss-weighted element means
    if (reaction) {
        hipLaunchKernelGGL(k_mg_cvol_mean, dim3(mg_grid(E)), dim3(MGT), 0, s, E, fine->k, fine->a64.tab,
                           fine->cvol64.as<double>(), cmean.as<double>());
        FEM_HIP(hipGetLastError());
    }
    auto coefficients = [&](Operator *o) {
        if (desc.kappa_mode == FEM_KAPPA_CELL) {
            FEM_HIP(hipMemcpyAsync(o->kcell64.p, fine->kcell64.p, sizeof(double) * E, hipMemcpyDeviceToDevice, s));
            round_into(o->kcell32.p, o->kcell64.p, E);
        }
        if (reaction) {
            const int64_t nq = E * o->k * o->k * o->k;
            hipLaunchKernelGGL(k_mg_cvol_fill, dim3(mg_grid(nq)), dim3(MGT), 0, s, E, o->k * o->k * o->k,
                               cmean.as<double>(), o->cvol64.as<double>());
            FEM_HIP(hipGetLastError());
            round_into(o->cvol32.p, o->cvol64.p, nq);
        }
    };
    for (auto &L : dg) {
        if (L->own) coefficients(L->own.get());
        if (L->resop) coefficients(L->resop.get());
    }
    // smoother blocks; a varying pointwise reaction enters them by its
    // largest value per element: the damped sweep stays convergent (a block
    // below the true one can make it diverge, and the V-cycle indefinite)
    for (auto &L : dg) {
        if (L->bj.op) L->bj.update();
        else L->bj.init(L->op, FEM_PRE_BLOCK_JACOBI, true);
    }
    // vertex stencils: the conforming level from the Q1 level's FP64
    // coefficient data (every operator precision keeps it), then the
    // Galerkin products
    {
        const MGVertexLevel &V0 = *vx[0];
        VDims dv{ { V0.nv[0], V0.nv[1], V0.nv[2] } };
        hipLaunchKernelGGL(k_mg_conforming, dim3(mg_grid(V0.V * 27)), dim3(MGT), 0, s, dg.back()->op->a64, dv,
                           A64[0]->as<double>());
        FEM_HIP(hipGetLastError());
    }
    for (size_t l = 1; l < vx.size(); ++l) {
        const MGVertexLevel &F = *vx[l - 1], &C = *vx[l];
        VDims df{ { F.nv[0], F.nv[1], F.nv[2] } }, dc{ { C.nv[0], C.nv[1], C.nv[2] } };
        hipLaunchKernelGGL(k_mg_galerkin, dim3(mg_grid(C.V * 27)), dim3(MGT), 0, s, df, dc, A64[l - 1]->as<double>(),
                           A64[l]->as<double>());
        FEM_HIP(hipGetLastError());
    }
    // level data in the preconditioner's precision; a reaction-free level
    // keeps the plain diagonal (its stencils have row sums of at most twice
    // the diagonal away from the boundary)
    const double lim = (desc.c != 0 || reaction) ? omega_vertex / 1.9 : 0.0;
    for (size_t l = 0; l < vx.size(); ++l) {
        MGVertexLevel &L = *vx[l];
        hipLaunchKernelGGL(k_mg_dinv, dim3(mg_grid(L.V)), dim3(MGT), 0, s, L.V, A64[l]->as<double>(), lim,
                           d64.as<double>());
        hipLaunchKernelGGL(k_mg_by_offset, dim3(mg_grid(L.V * 27)), dim3(MGT), 0, s, L.V, A64[l]->as<double>(),
                           As64.as<double>());
        FEM_HIP(hipGetLastError());
        store_as(lp, L.A, As64, L.V * 27, ctx);
        store_as(lp, L.dinv, d64, L.V, ctx);
    }
    // dense inverse of the coarsest level, stored row-major: Gauss-Jordan on
    // the device (two pivots per launch, no pivoting) for the symmetric
    // positive definite operators of diffusion and non-negative reaction; a
    // non-positive pivot (an indefinite operator) sends the inversion to the
    // host, with partial pivoting
    const MGVertexLevel &LC = *vx.back();
    const int VC = (int)LC.V;
    bool spd = true;
    {
        VDims dc{ { LC.nv[0], LC.nv[1], LC.nv[2] } };
        const unsigned G = mg_grid((int64_t)VC * VC);
        hipLaunchKernelGGL(k_mg_dense_fill, dim3(G), dim3(MGT), 0, s, dc, A64.back()->as<double>(), gjX.as<double>());
        int launches = 0;
        for (int c = 0; c < VC; c += 2, ++launches) {
            const DevBuf &src = (launches & 1) ? gjY : gjX;
            DevBuf &dst = (launches & 1) ? gjX : gjY;
            hipLaunchKernelGGL(k_mg_gj_step, dim3(G), dim3(MGT), 0, s, VC, c, std::min(2, VC - c), src.as<double>(),
                               dst.as<double>(), gjp.as<double>());
        }
        FEM_HIP(hipGetLastError());
        std::vector<double> piv(VC);
        FEM_HIP(hipMemcpyAsync(piv.data(), gjp.p, sizeof(double) * VC, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        for (int c = 0; c < VC; ++c) spd = spd && piv[c] > 0 && std::isfinite(piv[c]);
        if (spd) store_as(lp, coarse_inv, (launches & 1) ? gjY : gjX, (int64_t)VC * VC, ctx);
    }
    if (!spd) {
        const MGVertexLevel &L = LC;
        const int V = VC;
        std::vector<double> st(L.V * 27);
        FEM_HIP(hipMemcpyAsync(st.data(), A64.back()->p, sizeof(double) * st.size(), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        std::vector<double> M((size_t)V * V, 0.0), I((size_t)V * V, 0.0);
        for (int v = 0; v < V; ++v) {
            const int x0 = v % L.nv[0], y0 = (v / L.nv[0]) % L.nv[1], z0 = v / (L.nv[0] * L.nv[1]);
            for (int sidx = 0; sidx < 27; ++sidx) {
                const int wx = x0 + sidx % 3 - 1, wy = y0 + (sidx / 3) % 3 - 1, wz = z0 + sidx / 9 - 1;
                if (wx < 0 || wy < 0 || wz < 0 || wx >= L.nv[0] || wy >= L.nv[1] || wz >= L.nv[2]) continue;
                M[(size_t)v * V + wx + L.nv[0] * (wy + L.nv[1] * wz)] = st[(size_t)v * 27 + sidx];
            }
            I[(size_t)v * V + v] = 1.0;
        }
        for (int c = 0; c < V; ++c) {
            int piv = c;
            for (int r = c + 1; r < V; ++r)
                if (std::fabs(M[(size_t)r * V + c]) > std::fabs(M[(size_t)piv * V + c])) piv = r;
            FEM_CHECK(M[(size_t)piv * V + c] != 0.0, FEM_ERR_ARG, "multigrid: singular coarsest operator");
            if (piv != c)
                for (int k = 0; k < V; ++k) {
                    std::swap(M[(size_t)piv * V + k], M[(size_t)c * V + k]);
                    std::swap(I[(size_t)piv * V + k], I[(size_t)c * V + k]);
                }
            const double ip = 1.0 / M[(size_t)c * V + c];
            for (int k = 0; k < V; ++k) {
                M[(size_t)c * V + k] *= ip;
                I[(size_t)c * V + k] *= ip;
            }
            for (int r = 0; r < V; ++r) {
                if (r == c) continue;
                const double f = M[(size_t)r * V + c];
                if (f == 0.0) continue;
                for (int k = 0; k < V; ++k) {
                    M[(size_t)r * V + k] -= f * M[(size_t)c * V + k];
                    I[(size_t)r * V + k] -= f * I[(size_t)c * V + k];
                }
            }
        }
        // row-major: element (row v, column w) at v * V + w
        const std::vector<double> &cm = I;
        DevBuf c64;
        c64.alloc(cm.size() * 8);
        FEM_HIP(hipMemcpy(c64.p, cm.data(), cm.size() * 8, hipMemcpyHostToDevice));
        store_as(lp, coarse_inv, c64, (int64_t)cm.size(), ctx);
    }
    FEM_HIP(hipStreamSynchronize(s));
}

template <typename T>
void Multigrid::cycle_vertex(size_t l, const T *r, T *x, const int *skip)
{
    MGVertexLevel &L = *vx[l];
    hipStream_t s = op->ctx->stream;
    const unsigned G = mg_grid(L.V);
    if (l + 1 == vx.size()) {
        hipLaunchKernelGGL(k_mg_dense<T>, dim3((unsigned)((L.V + 3) / 4)), dim3(256), 0, s, (int)L.V,
                           coarse_inv.as<T>(), r, x, skip);
        return;
    }
    VDims d{ { L.nv[0], L.nv[1], L.nv[2] } };
    T *t = L.t.as<T>();
    const T *A = L.A.as<T>(), *dv = L.dinv.as<T>();
    const T om = (T)omega_vertex;
    auto sweep = [&]() {
        hipLaunchKernelGGL(k_mg_v_apply<T>, dim3(G), dim3(MGT), 0, s, d, A, (const T *)x, r, t, -1, skip);
        hipLaunchKernelGGL(k_mg_v_jacobi<T>, dim3(G), dim3(MGT), 0, s, L.V, dv, r, (const T *)t, x, om, 1, skip);
    };
    hipLaunchKernelGGL(k_mg_v_jacobi<T>, dim3(G), dim3(MGT), 0, s, L.V, dv, r, (const T *)nullptr, x, om, 0, skip);
    for (int it = 1; it < nu_vertex; ++it) sweep();
    hipLaunchKernelGGL(k_mg_v_apply<T>, dim3(G), dim3(MGT), 0, s, d, A, (const T *)x, r, t, -1, skip);
    MGVertexLevel &C = *vx[l + 1];
    VDims dc{ { C.nv[0], C.nv[1], C.nv[2] } };
    hipLaunchKernelGGL(k_mg_v_restrict<T>, dim3(mg_grid(C.V)), dim3(MGT), 0, s, d, dc, (const T *)t, C.r.as<T>(), skip);
    cycle_vertex<T>(l + 1, C.r.as<T>(), C.x.as<T>(), skip);
    hipLaunchKernelGGL(k_mg_v_prolong<T>, dim3(G), dim3(MGT), 0, s, d, dc, (const T *)C.x.as<T>(), x, skip);
    for (int it = 0; it < nu_vertex; ++it) sweep();
    FEM_HIP(hipGetLastError());
}

template <typename T>
static void mg_bj(const MGDGLevel &L, const T *r, const T *t, T *x, T omega, int mode, const int *skip, hipStream_t s)
{
    const int64_t E = L.op->E;
    const T *Q = L.bj.Q.as<T>(), *lam = L.bj.lam.as<T>();
#define MGBJ_CASE(KK)                                                                                        \
    case KK:                                                                                                 \
        hipLaunchKernelGGL((k_mg_bj<KK, T>), dim3((unsigned)((E + BJShape<KK>::EPB - 1) / BJShape<KK>::EPB)), \
                           dim3(64), 0, s, E, Q, lam, r, t, x, omega, mode, skip);                           \
        break;
    switch (L.k) {
        MGBJ_CASE(2) MGBJ_CASE(3) MGBJ_CASE(4) MGBJ_CASE(5)
    default: throw Error(FEM_ERR_UNSUPPORTED, "multigrid: order must be 1..4");
    }
#undef MGBJ_CASE
}

// one damped block-Jacobi sweep on a level, out = in + omega B^-1 (r - A in),
// in one pass of the apply kernel (FP64 and FP32 operators) or as an apply
// followed by the block-Jacobi kernel (BF16-MFMA operators, whose sweep
// keeps the BF16 apply; then out == in and t is the scratch for A in)
template <typename T>
static void mg_sweep(const MGDGLevel &L, const T *r, const T *in, T *out, T *t, T omega, const int *skip, hipStream_t s)
{
    const Operator &o = *L.op;
    // the fused sweep has no pointwise reaction: apply, then the sweep
    if (o.desc.precision == FEM_BF16MFMA || o.cvol64.p) {
        L.op->apply_raw(in, t, skip);
        mg_bj<T>(L, r, t, out, omega, 1, skip, s);
        return;
    }
    DGSmooth<T> sm;
    sm.r = r;
    sm.Q = L.bj.Q.as<T>();
    sm.lam = L.bj.lam.as<T>();
    sm.omega = omega;
    sm.xout = out;
    if constexpr (std::is_same<T, double>::value) dg_smooth_f64(o.a64, o.k, in, sm, skip, s);
    else dg_smooth_f32(o.a32, o.k, in, sm, skip, s);
}

template <typename T>
void Multigrid::cycle_dg(size_t l, const T *r, T *x, const int *skip)
{
    MGDGLevel &L = *dg[l];
    hipStream_t s = op->ctx->stream;
    const int *n = op->desc.n;
    const int64_t E = op->E;
    T *t = L.t.as<T>();
    const T om = (T)omega_dg;
    // the fused sweeps write a second buffer (a sweep reads its input's
    // neighbours): cur is the buffer holding the current iterate, and the
    // first sweep is arranged so that the last one ends in x
    const bool fused = L.op->desc.precision != FEM_BF16MFMA && !L.op->cvol64.p;
    T *cur = x;
    auto sweep = [&]() {
        if (!fused) {
            mg_sweep<T>(L, r, cur, cur, t, om, skip, s);
            return;
        }
        T *nxt = cur == x ? t : x;
        mg_sweep<T>(L, r, cur, nxt, t, om, skip, s);
        cur = nxt;
    };
    // the pre-smoothing leaves its result in t when the remaining sweeps
    // (nu - 1 before, nu after) are odd in number, so that they end in x
    if (fused && ((2 * nu_dg - 1) & 1)) cur = t;
    mg_bj<T>(L, r, nullptr, cur, om, 0, skip, s);
    for (int it = 1; it < nu_dg; ++it) sweep();
    // A cur for the restriction, in whichever buffer cur is not
    T *ax = fused ? (cur == x ? t : x) : t;
    (L.resop ? L.resop.get() : L.op)->apply_raw(cur, ax, skip);
    if (l + 1 < dg.size()) {
        MGDGLevel &C = *dg[l + 1];
        hipLaunchKernelGGL(k_mg_restrict_modes<T>, dim3(mg_grid(E * C.k * C.k * C.k)), dim3(MGT), 0, s, E, L.k, C.k, r,
                           (const T *)ax, C.r.as<T>(), skip);
        cycle_dg<T>(l + 1, C.r.as<T>(), C.x.as<T>(), skip);
        hipLaunchKernelGGL(k_mg_prolong_modes<T>, dim3(mg_grid(E * C.k * C.k * C.k)), dim3(MGT), 0, s, E, L.k, C.k,
                           (const T *)C.x.as<T>(), cur, skip);
    } else {
        MGVertexLevel &V0 = *vx[0];
        dg1_to_vertex<T>(n, r, (const T *)ax, corners.as<double>(), V0.r.as<T>(), skip, s);
        cycle_vertex<T>(0, V0.r.as<T>(), V0.x.as<T>(), skip);
        hipLaunchKernelGGL(k_mg_vertex_to_dg1<T>, dim3(mg_grid(E * 8)), dim3(MGT), 0, s, n[0], n[1], n[2],
                           (const T *)V0.x.as<T>(), cur, skip);
    }
    for (int it = 0; it < nu_dg; ++it) sweep();
    FEM_CHECK(cur == x, FEM_ERR_STATE, "multigrid: the sweeps end outside the output buffer");
    FEM_HIP(hipGetLastError());
}

void Multigrid::smooth_fine(const void *r, const void *t, void *x, const int *skip)
{
    hipStream_t s = op->ctx->stream;
    if (prec == FEM_F64)
        mg_bj<double>(*dg[0], (const double *)r, (const double *)t, (double *)x, (double)omega_dg, 1, skip, s);
    else
        mg_bj<float>(*dg[0], (const float *)r, (const float *)t, (float *)x, (float)omega_dg, 1, skip, s);
    FEM_HIP(hipGetLastError());
}

void Multigrid::apply_raw(const void *r, void *z, const int *skip)
{
    if (prec == FEM_F64) cycle_dg<double>(0, (const double *)r, (double *)z, skip);
    else cycle_dg<float>(0, (const float *)r, (float *)z, skip);
}

}  // namespace femamd
