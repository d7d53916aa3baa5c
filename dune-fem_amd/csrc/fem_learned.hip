// Learned coarse-space preconditioner for SIPG DG operators.
//
// Coarse space: element means (the phi_0 coefficient of the orthonormal
// Legendre basis).  Restriction R v = v[e*nb] of a dual vector, prolongation
// P c = c_e in the phi_0 slot; A_c = P^T A P is a 7-point stencil on the
// element grid (only the penalty terms act on piecewise constants).
//
// The exact two-level preconditioner would use the coarse solve A_c^-1 R v.
// Here the network gives the initial guess of that coarse solve and a few
// Jacobi-preconditioned CG steps on A_c (device-resident scalars) finish it:
//   c   = CG_k(A_c, R v; x0 = s0 gamma0 CNN(R v / s0, R v / (kappa s1)))
//   z1  = alpha P c,   alpha = <v, Pc> / <Pc, A Pc>   (energy-optimal step)
//   z   = z1 + BJ(v - A z1)
// BJ is the exact element-block inverse.  The map is nonlinear in v, so it is
// used inside FGMRES.  Mode FEM_LEARNED_COARSE_ONLY starts the coarse CG from
// zero (the same method without the network), FEM_LEARNED_OFF is block Jacobi.
//
// Training pairs (r_k, e_k) are the residuals and errors of the iterates of
// the plain FP64 CG + block-Jacobi solve.  The objective is the energy-norm
// distance between the coarse correction and the solver's error,
//   ||e_k - P c||_A^2 = ||e_k - P c*||_A^2 + (c - c*)^T A_c (c - c*),
//   c* = A_c^-1 P^T A e_k = A_c^-1 R r_k,
// so only R r_k enters; per pair it is normalised by the best achievable
// reduction w = (R r_k)^T c* (loss 1 for c = 0, 0 for c = c*).
//
// No reference counterpart (new preconditioner); the solve is checked by its
// FP64 true residual.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "fem_blas.hpp"
#include "fem_cnn.hpp"
#include "fem_learned.hpp"
#include "fem_mg.hpp"
#include "fem_solver.hpp"

namespace femamd {

constexpr int LR_BLOCKS = 256;

__device__ __forceinline__ unsigned short lf2bf(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float lround_bf(float x) { return __uint_as_float((unsigned)lf2bf(x) << 16); }

struct CGrid {
    int nx, ny, nz, V;
};

// sum over the six neighbours of w_q u(q) (zero outside the grid)
template <typename F>
__device__ __forceinline__ float offdiag(const CGrid &g, const float *w, int v, F u)
{
    const int x = v % g.nx, y = (v / g.nx) % g.ny, z = v / (g.nx * g.ny);
    const int sy = g.nx, sz = g.nx * g.ny;
    float a = 0.f;
    if (x > 0) a += w[1] * u(v - 1);
    if (x < g.nx - 1) a += w[2] * u(v + 1);
    if (y > 0) a += w[3] * u(v - sy);
    if (y < g.ny - 1) a += w[4] * u(v + sy);
    if (z > 0) a += w[5] * u(v - sz);
    if (z < g.nz - 1) a += w[6] * u(v + sz);
    return a;
}

// ---- coarse CG (Jacobi-preconditioned, on the 7-point stencil) --------------------
// Chronopoulos-Gear form of PCG (the two inner products of a step are reduced
// together, so one kernel per step).  With u = D^-1 r and w = A_c u:
//   p_i = u_i + beta_i p_{i-1},  s_i = w_i + beta_i s_{i-1}   (s_i = A_c p_i)
//   x_{i+1} = x_i + alpha_i p_i,  r_{i+1} = r_i - alpha_i s_i
//   gamma_{i+1} = (r_{i+1}, u_{i+1}),  delta_{i+1} = (w_{i+1}, u_{i+1})
//   beta_{i+1} = gamma_{i+1} / gamma_i
//   alpha_{i+1} = gamma_{i+1} / (delta_{i+1} - beta_{i+1} gamma_{i+1} / alpha_i)
// In exact arithmetic the iterates are those of the two-reduction PCG.  The
// step kernel forms u_{i+1} of the six neighbours on the fly from the old r,
// w, s, so r, w, s are double-buffered.
// No atomics: a kernel leaves one (gamma, delta) partial per block, and every
// block of the next step sums all of them in the same fixed order, so all
// blocks derive bitwise the same alpha and beta.  Block 0 records
// (gamma_i, alpha_i) for the step after (ping-pong scalar slots).
// sc[0..1], sc[2..3]: (gamma, alpha) of the previous step; sc[4] = gamma_0

// x0 = k0 y (or 0), r0 = b - A_c x0, p = s = 0; no previous step
template <typename T>
__global__ void __launch_bounds__(256) k_cc_start(CGrid g, const float *__restrict__ st, const T *__restrict__ src,
                                                  int stride, const float *__restrict__ y,
                                                  const float *__restrict__ red, float gamma0,
                                                  float *__restrict__ x, float *__restrict__ r,
                                                  float *__restrict__ p, float *__restrict__ s, double *sc0,
                                                  const int *__restrict__ skip)
{
    if (skip && *skip) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc0[0] = 0.0;
        sc0[1] = 0.0;
    }
    const float k0 = y ? red[0] * gamma0 : 0.f;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.V; v += gridDim.x * blockDim.x) {
        const float *w = st + (size_t)v * 7;
        const float b = (float)src[(size_t)v * stride];
        float rv = b;
        if (y) {
            const float xv = k0 * y[v];
            rv = b - (w[0] * xv - offdiag(g, w, v, [&](int q) { return k0 * y[q]; }));
            x[v] = xv;
        } else {
            x[v] = 0.f;
        }
        r[v] = rv;
        p[v] = 0.f;
        s[v] = 0.f;
    }
}

// block partials of (gamma, delta): part[b], part[gridDim + b]
__device__ __forceinline__ void cc_partials(double ag, double ad, double *lds, double *part)
{
    const double tg = block_sum(ag, lds);
    __syncthreads();
    const double td = block_sum(ad, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tg;
        part[gridDim.x + blockIdx.x] = td;
    }
}

// w0 = A_c u0 and the partials of gamma_0, delta_0
__global__ void __launch_bounds__(256) k_cc_init(CGrid g, const float *__restrict__ st, const float *__restrict__ r,
                                                 float *__restrict__ wo, double *__restrict__ pout,
                                                 const int *__restrict__ skip)
{
    if (skip && *skip) return;
    double ag = 0, ad = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.V; v += gridDim.x * blockDim.x) {
        const float *w = st + (size_t)v * 7;
        const float uv = r[v] / w[0];
        const float wv = w[0] * uv - offdiag(g, w, v, [&](int q) { return r[q] / st[(size_t)q * 7]; });
        wo[v] = wv;
        ag += (double)r[v] * uv;
        ad += (double)wv * uv;
    }
    __shared__ double lds[4];
    cc_partials(ag, ad, lds, pout);
}

// The cell data one step reads: the cell's old r, w, s, p and its seven
// stencil weights, and the old r, w, s and diagonal of the six neighbours
// (zero weight outside the grid), in the stencil's neighbour order.
struct CCCell {
    float r, w, s, p, wt[7];
    float nr[6], nw[6], ns[6], nd[6];
};

__device__ __forceinline__ void cc_load(const CGrid &g, int v, const float *__restrict__ st,
                                        const float *__restrict__ ro, const float *__restrict__ wo,
                                        const float *__restrict__ so, const float *__restrict__ p, CCCell &c)
{
    c.r = ro[v];
    c.w = wo[v];
    c.s = so[v];
    c.p = p[v];
#pragma unroll
    for (int k = 0; k < 7; ++k) c.wt[k] = st[(size_t)v * 7 + k];
    const int x = v % g.nx, y = (v / g.nx) % g.ny, z = v / (g.nx * g.ny);
    const int sy = g.nx, sz = g.nx * g.ny;
    const bool in[6] = { x > 0, x < g.nx - 1, y > 0, y < g.ny - 1, z > 0, z < g.nz - 1 };
    const int off[6] = { -1, 1, -sy, sy, -sz, sz };
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int q = in[k] ? v + off[k] : v;
        c.nr[k] = ro[q];
        c.nw[k] = wo[q];
        c.ns[k] = so[q];
        c.nd[k] = st[(size_t)q * 7];
        if (!in[k]) c.wt[k + 1] = 0.f;
    }
}

// r_{i+1} of a cell from its old r, w, s (the same expression for a cell and
// for the neighbours that form it on the fly)
__device__ __forceinline__ float cc_rnext(float r, float w, float s, float al, float be)
{
    return r - al * (w + be * s);
}

// one CG step: (x, p in place; r, w, s from the `o` buffers into the `n`
// buffers); pin holds the gridDim partials of gamma_i and of delta_i.  The
// first cell of each thread is loaded before the scalars are formed, so the
// loads overlap the partial sums.
__global__ void __launch_bounds__(256) k_cc_step(CGrid g, const float *__restrict__ st, float *__restrict__ x,
                                                 float *__restrict__ p, const float *__restrict__ ro,
                                                 const float *__restrict__ wo, const float *__restrict__ so,
                                                 float *__restrict__ rn, float *__restrict__ wn,
                                                 float *__restrict__ sn, const double *__restrict__ pin,
                                                 double *__restrict__ pout, const double *__restrict__ scin,
                                                 double *__restrict__ scout, double *__restrict__ g0, int first,
                                                 const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double lds[4];
    __shared__ float scal[2];
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    CCCell c;
    if (v0 < g.V) cc_load(g, v0, st, ro, wo, so, p, c);
    double pg = 0, pd = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        pg += pin[b];
        pd += pin[gridDim.x + b];
    }
    const double ga = block_sum(pg, lds);
    __syncthreads();
    const double de = block_sum(pd, lds);
    if (threadIdx.x == 0) {
        const double gold = scin[0], aold = scin[1];
        const double beta = (gold > 0.0 && aold > 0.0) ? ga / gold : 0.0;
        const double den = de - (beta != 0.0 ? beta * ga / aold : 0.0);
        const double alpha = (ga > 0.0 && den > 0.0) ? ga / den : 0.0;
        scal[0] = (float)alpha;
        scal[1] = (float)beta;
        if (blockIdx.x == 0) {
            scout[0] = ga;
            scout[1] = alpha;
            if (first) *g0 = ga;
        }
    }
    __syncthreads();
    const float al = scal[0], be = scal[1];
    double ag = 0, ad = 0;
    for (int v = v0; v < g.V; v += gridDim.x * blockDim.x) {
        if (v != v0) cc_load(g, v, st, ro, wo, so, p, c);
        const float d = c.wt[0];
        const float pv = c.r / d + be * c.p;
        const float sv = c.w + be * c.s;
        x[v] += al * pv;
        const float rv = cc_rnext(c.r, c.w, c.s, al, be);
        const float uv = rv / d;
        float off = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) off += c.wt[k + 1] * (cc_rnext(c.nr[k], c.nw[k], c.ns[k], al, be) / c.nd[k]);
        const float wv = d * uv - off;
        p[v] = pv;
        sn[v] = sv;
        rn[v] = rv;
        wn[v] = wv;
        ag += (double)rv * uv;
        ad += (double)wv * uv;
    }
    __syncthreads();
    cc_partials(ag, ad, lds, pout);
}

// coarse CG launches: one cell per thread up to CC_MAX_BLOCKS workgroups
constexpr int CC_MAX_BLOCKS = 1024;
static unsigned cc_blocks(int V)
{
    return (unsigned)std::max(1, std::min(CC_MAX_BLOCKS, (V + 255) / 256));
}

void CoarseCG::init(const CNN *c)
{
    cnn = c;
    V = (int)c->V;
    for (DevBuf *b : { &x, &p, &r[0], &r[1], &w[0], &w[1], &s[0], &s[1] }) b->alloc(sizeof(float) * V);
    sc.alloc(sizeof(double) * 8);
    for (DevBuf *b : { &part[0], &part[1] }) b->alloc(sizeof(double) * 2 * CC_MAX_BLOCKS);
    FEM_HIP(hipMemset(sc.p, 0, sc.bytes));
}

template <typename T>
void CoarseCG::start(const T *src, int stride, const float *y, const float *red, float gamma0, hipStream_t st,
                     const int *skip)
{
    CGrid g{ cnn->n[0], cnn->n[1], cnn->n[2], V };
    cur = 0;
    nstep = 0;
    hipLaunchKernelGGL(k_cc_start<T>, dim3(LR_BLOCKS), dim3(256), 0, st, g, cnn->stencil.as<float>(), src, stride, y,
                       red, gamma0, x.as<float>(), r[0].as<float>(), p.as<float>(), s[0].as<float>(), sc.as<double>(),
                       skip);
    hipLaunchKernelGGL(k_cc_init, dim3(cc_blocks(V)), dim3(256), 0, st, g, cnn->stencil.as<float>(), r[0].as<float>(),
                       w[0].as<float>(), part[0].as<double>(), skip);
    FEM_HIP(hipGetLastError());
}

void CoarseCG::step(hipStream_t st, const int *skip)
{
    CGrid g{ cnn->n[0], cnn->n[1], cnn->n[2], V };
    const int o = cur, n = cur ^ 1;
    hipLaunchKernelGGL(k_cc_step, dim3(cc_blocks(V)), dim3(256), 0, st, g, cnn->stencil.as<float>(), x.as<float>(),
                       p.as<float>(), r[o].as<float>(), w[o].as<float>(), s[o].as<float>(), r[n].as<float>(),
                       w[n].as<float>(), s[n].as<float>(), part[o].as<double>(), part[n].as<double>(),
                       sc.as<double>() + 2 * o, sc.as<double>() + 2 * n, sc.as<double>() + 4, nstep == 0 ? 1 : 0,
                       skip);
    FEM_HIP(hipGetLastError());
    cur = n;
    ++nstep;
}

void CoarseCG::gammas(hipStream_t st, double &now, double &first) const
{
    const int nb = (int)cc_blocks(V);
    double h[CC_MAX_BLOCKS + 1];
    FEM_HIP(hipMemcpyAsync(h, part[cur].p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    FEM_HIP(hipMemcpyAsync(h + nb, sc.as<double>() + 4, sizeof(double), hipMemcpyDeviceToHost, st));
    FEM_HIP(hipStreamSynchronize(st));
    now = 0;
    for (int b = 0; b < nb; ++b) now += h[b];
    first = nstep ? h[nb] : now;
}

// ---- network input ---------------------------------------------------------------------
// red[0] = RMS of rc, red[1] = RMS of rc * kinv; rc = src[e * stride]
template <typename T>
__global__ void __launch_bounds__(256) k_coarse_rms(const T *__restrict__ src, int stride, const float *__restrict__ kinv,
                                                    int E, double *partials, unsigned *counter, float *red,
                                                    const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double lds[4];
    double a = 0, b = 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const double r = (double)src[(size_t)e * stride];
        const double rk = r * kinv[e];
        a += r * r;
        b += rk * rk;
    }
    double own[2];
    own[0] = block_sum(a, lds);
    __syncthreads();
    own[1] = block_sum(b, lds);
    __shared__ double res[2];
    if (finish_reduction(0.0, partials, counter, res, 2, own)) {
        const float s0 = (float)sqrt(res[0] / E), s1 = (float)sqrt(res[1] / E);
        red[0] = s0 > 0.f ? s0 : 1.f;
        red[1] = s1 > 0.f ? s1 : 1.f;
    }
}

// features x0[e] = (bf16(rc / s0), bf16(rc * kinv / s1))
template <typename T>
__global__ void k_coarse_feat(const T *__restrict__ src, int stride, const float *__restrict__ kinv, int E,
                              const float *__restrict__ red, float *__restrict__ x0, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float r = (float)src[(size_t)e * stride];
    x0[2 * e] = lround_bf(r / red[0]);
    x0[2 * e + 1] = lround_bf(r * kinv[e] / red[1]);
}

// ---- fine-level pieces of the two-level map ------------------------------------------------
// z = 0 except z[e*nb] = c_e
template <typename T>
__global__ void k_prolong(const float *__restrict__ c, int nb, int64_t N, T *__restrict__ z,
                          const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = (i % nb == 0) ? (T)c[i / nb] : T(0);
}

template <typename T>
__global__ void k_addv(T *__restrict__ y, const T *__restrict__ z, int64_t N, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += z[i];
}

// alpha = <v, z1> / <z1, w> (0 when the direction has no positive energy) into red[2]
template <typename T>
__global__ void __launch_bounds__(256) k_alpha(const T *__restrict__ v, const T *__restrict__ z1,
                                               const T *__restrict__ w, int64_t N, double *partials, unsigned *counter,
                                               float *red, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double lds[4];
    double a = 0, b = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const double zi = (double)z1[i];
        a += (double)v[i] * zi;
        b += zi * (double)w[i];
    }
    double own[2];
    own[0] = block_sum(a, lds);
    __syncthreads();
    own[1] = block_sum(b, lds);
    __shared__ double res[2];
    if (finish_reduction(0.0, partials, counter, res, 2, own))
        red[2] = (res[1] > 0 && res[0] == res[0]) ? (float)(res[0] / res[1]) : 0.f;
}

// t = v - alpha w ; z1 *= alpha
template <typename T>
__global__ void k_scaled_residual(const T *__restrict__ v, const T *__restrict__ w, T *__restrict__ t,
                                  T *__restrict__ z1, int64_t N, const float *__restrict__ red,
                                  const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const T al = (T)red[2];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        t[i] = v[i] - al * w[i];
        z1[i] *= al;
    }
}

// The same three steps when the apply hands over per-workgroup partials of
// <z1, A z1>: z1 = P c is zero outside the phi_0 slots, so <v, z1> is a sum
// over the elements, t = v - alpha w1 needs no pass over z1, and y += alpha z1
// touches the phi_0 slots only.
template <typename T>
__global__ void __launch_bounds__(256) k_alpha_coarse(const T *__restrict__ v, int nb, const float *__restrict__ c,
                                                      int E, const double *__restrict__ parts, int64_t np,
                                                      double *partials, unsigned *counter, float *red,
                                                      const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double lds[4];
    double a = 0, b = 0;
    const int64_t g0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, gs = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = g0; e < E; e += gs) a += (double)v[e * nb] * (double)(T)c[e];
    for (int64_t k = g0; k < np; k += gs) b += parts[k];
    double own[2];
    own[0] = block_sum(a, lds);
    __syncthreads();
    own[1] = block_sum(b, lds);
    __shared__ double res[2];
    if (finish_reduction(0.0, partials, counter, res, 2, own))
        red[2] = (res[1] > 0 && res[0] == res[0]) ? (float)(res[0] / res[1]) : 0.f;
}

// t = v - alpha w
template <typename T>
__global__ void k_residual_t(const T *__restrict__ v, const T *__restrict__ w, T *__restrict__ t, int64_t N,
                             const float *__restrict__ red, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const T al = (T)red[2];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        t[i] = v[i] - al * w[i];
}

// y[e nb] += alpha c_e (the same rounding as scaling z1 and adding it)
template <typename T>
__global__ void k_add_coarse(T *__restrict__ y, const float *__restrict__ c, int nb, int E,
                             const float *__restrict__ red, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const T al = (T)red[2];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) y[(int64_t)e * nb] += (T)c[e] * al;
}

// ---------------------------------------------------------------------------------
// t = v - w
template <typename T>
__global__ void k_sub(const T *__restrict__ v, const T *__restrict__ w, T *__restrict__ t, int64_t N,
                      const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        t[i] = v[i] - w[i];
}

// energy-optimal step of the coarse correction P c of the residual t:
// alpha = <t, P c> / (c^T A_c c) with A_c c given (acc), into red[2]
template <typename T>
__global__ void __launch_bounds__(256) k_alpha_galerkin(const T *__restrict__ t, int nb, const float *__restrict__ c,
                                                        const float *__restrict__ acc, int E, double *partials,
                                                        unsigned *counter, float *red, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double lds[4];
    double a = 0, b = 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        a += (double)t[e * nb] * (double)c[e];
        b += (double)c[e] * (double)acc[e];
    }
    double own[2];
    own[0] = block_sum(a, lds);
    __syncthreads();
    own[1] = block_sum(b, lds);
    __shared__ double res[2];
    if (finish_reduction(0.0, partials, counter, res, 2, own))
        red[2] = (res[1] > 0 && res[0] == res[0]) ? (float)(res[0] / res[1]) : 0.f;
}

void Learned::init(Operator *o, CNN *c, int steps, int smoother_kind)
{
    op = o;
    FEM_CHECK(smoother_kind == FEM_PRE_BLOCK_JACOBI || smoother_kind == FEM_PRE_MULTIGRID, FEM_ERR_ARG,
              "learned preconditioner: smoother is FEM_PRE_BLOCK_JACOBI or FEM_PRE_MULTIGRID");
    smoother = smoother_kind;
    cnn = c;
    FEM_CHECK(steps >= 0 && steps <= 1000, FEM_ERR_ARG, "learned preconditioner: coarse_steps in [0, 1000]");
    coarse_steps = steps;
    FEM_CHECK(o->is_dg(), FEM_ERR_UNSUPPORTED, "learned preconditioner needs a DG operator");
    FEM_CHECK(!o->has_reaction(), FEM_ERR_UNSUPPORTED, "learned preconditioner: no pointwise or nonlinear reaction");
    FEM_CHECK(c->V == o->E, FEM_ERR_ARG, "learned preconditioner: network grid must match the element grid");
    FEM_CHECK((int64_t)c->n[0] == o->desc.n[0] && c->n[1] == o->desc.n[1] && c->n[2] == o->desc.n[2], FEM_ERR_ARG,
              "learned preconditioner: network grid must match the element grid");
    FEM_CHECK(c->kinv.p && c->stencil.p, FEM_ERR_STATE,
              "learned preconditioner: set the network's coefficient first");
    bj.init(o, smoother);
    ccg.init(c);
    acc.alloc(sizeof(float) * o->E);
    const size_t vb = (size_t)o->N * (o->vec_prec == FEM_F64 ? 8 : 4);
    z1.alloc(vb);
    t1.alloc(vb);
    w1.alloc(vb);
    x0.alloc(sizeof(float) * 2 * o->E);
    yc.alloc(sizeof(float) * o->E);
    red.alloc(sizeof(float) * 4);
    partials.alloc(sizeof(double) * 2 * LR_BLOCKS);
    counter.alloc(sizeof(unsigned));
    FEM_HIP(hipMemset(counter.p, 0, sizeof(unsigned)));
    // the FP64 and FP32 DG applies can write per-workgroup <z1, A z1> partials
    if (o->desc.precision != FEM_BF16MFMA) dparts.alloc(sizeof(double) * o->dot_blocks());
}

template <typename T>
static void learned_apply_t(Learned &L, const T *x, T *y, const int *skip)
{
    Operator &o = *L.op;
    hipStream_t s = o.ctx->stream;
    const int nb = o.k * o.k * o.k;
    const int E = (int)o.E;
    const int64_t N = o.N;
    const unsigned G = grid_for(N, 256, 8192);
    const bool net = L.mode == FEM_LEARNED_NET;
    if (L.smoother == FEM_PRE_MULTIGRID) {
        // y = MG(x); t = x - A y; c from the network on R t (and coarse CG
        // steps); y += alpha P c with the energy-optimal alpha; then one
        // damped block-Jacobi sweep of the fine level on x - A y, which takes
        // out the jumps a piecewise-constant correction leaves at element
        // faces.  Without the network (FEM_LEARNED_OFF, or no coarse steps)
        // the V-cycle and the same sweep, so the two differ by the network
        // alone.
        T *t1 = L.t1.as<T>(), *w1 = L.w1.as<T>();
        L.bj.apply_raw(x, y, skip);
        auto post_sweep = [&]() {
            o.apply_raw(y, w1, skip);
            L.bj.mg->smooth_fine(x, w1, y, skip);
        };
        if (!(L.active() && (net || L.coarse_steps > 0))) {
            post_sweep();
            return;
        }
        o.apply_raw(y, w1, skip);
        hipLaunchKernelGGL(k_sub<T>, dim3(G), dim3(256), 0, s, x, (const T *)w1, t1, N, skip);
        if (net) {
            hipLaunchKernelGGL(k_coarse_rms<T>, dim3(LR_BLOCKS), dim3(256), 0, s, (const T *)t1, nb,
                               L.cnn->kinv.as<float>(), E, L.partials.as<double>(), L.counter.as<unsigned>(),
                               L.red.as<float>(), skip);
            hipLaunchKernelGGL(k_coarse_feat<T>, dim3((E + 255) / 256), dim3(256), 0, s, (const T *)t1, nb,
                               L.cnn->kinv.as<float>(), E, L.red.as<float>(), L.x0.as<float>(), skip);
            L.cnn->forward(L.x0.as<float>(), L.yc.as<float>(), s, skip);
        }
        L.ccg.start<T>(t1, nb, net ? L.yc.as<float>() : nullptr, L.red.as<float>(), L.cnn->gamma0, s, skip);
        for (int k = 0; k < L.coarse_steps; ++k) L.ccg.step(s, skip);
        L.cnn->coarse_apply(L.ccg.x.as<float>(), L.acc.as<float>(), s);
        hipLaunchKernelGGL(k_alpha_galerkin<T>, dim3(LR_BLOCKS), dim3(256), 0, s, (const T *)t1, nb,
                           (const float *)L.ccg.x.as<float>(), (const float *)L.acc.as<float>(), E,
                           L.partials.as<double>(), L.counter.as<unsigned>(), L.red.as<float>(), skip);
        hipLaunchKernelGGL(k_add_coarse<T>, dim3((E + 255) / 256), dim3(256), 0, s, y, L.ccg.x.as<float>(), nb, E,
                           L.red.as<float>(), skip);
        FEM_HIP(hipGetLastError());
        post_sweep();
        return;
    }
    if (!L.active()) {
        L.bj.apply_raw(x, y, skip);
        return;
    }
    if (net) {
        hipLaunchKernelGGL(k_coarse_rms<T>, dim3(LR_BLOCKS), dim3(256), 0, s, x, nb, L.cnn->kinv.as<float>(), E,
                           L.partials.as<double>(), L.counter.as<unsigned>(), L.red.as<float>(), skip);
        hipLaunchKernelGGL(k_coarse_feat<T>, dim3((E + 255) / 256), dim3(256), 0, s, x, nb, L.cnn->kinv.as<float>(), E,
                           L.red.as<float>(), L.x0.as<float>(), skip);
        L.cnn->forward(L.x0.as<float>(), L.yc.as<float>(), s, skip);
    }
    L.ccg.start<T>(x, nb, net ? L.yc.as<float>() : nullptr, L.red.as<float>(), L.cnn->gamma0, s, skip);
    for (int k = 0; k < L.coarse_steps; ++k) L.ccg.step(s, skip);
    // z1 = alpha P c; y = z1 + BJ(x - A z1)
    T *z1 = L.z1.as<T>(), *t1 = L.t1.as<T>(), *w1 = L.w1.as<T>();
    hipLaunchKernelGGL(k_prolong<T>, dim3(G), dim3(256), 0, s, L.ccg.x.as<float>(), nb, N, z1, skip);
    if (L.dparts.p) {
        o.apply_raw(z1, w1, skip, L.dparts.as<double>());
        hipLaunchKernelGGL(k_alpha_coarse<T>, dim3(LR_BLOCKS), dim3(256), 0, s, x, nb, L.ccg.x.as<float>(), E,
                           L.dparts.as<double>(), o.dot_blocks(), L.partials.as<double>(), L.counter.as<unsigned>(),
                           L.red.as<float>(), skip);
        hipLaunchKernelGGL(k_residual_t<T>, dim3(G), dim3(256), 0, s, x, w1, t1, N, L.red.as<float>(), skip);
        L.bj.apply_raw(t1, y, skip);
        hipLaunchKernelGGL(k_add_coarse<T>, dim3((E + 255) / 256), dim3(256), 0, s, y, L.ccg.x.as<float>(), nb, E,
                           L.red.as<float>(), skip);
        FEM_HIP(hipGetLastError());
        return;
    }
    o.apply_raw(z1, w1, skip);
    hipLaunchKernelGGL(k_alpha<T>, dim3(LR_BLOCKS), dim3(256), 0, s, x, z1, w1, N, L.partials.as<double>(),
                       L.counter.as<unsigned>(), L.red.as<float>(), skip);
    hipLaunchKernelGGL(k_scaled_residual<T>, dim3(G), dim3(256), 0, s, x, w1, t1, z1, N, L.red.as<float>(), skip);
    L.bj.apply_raw(t1, y, skip);
    hipLaunchKernelGGL(k_addv<T>, dim3(G), dim3(256), 0, s, y, z1, N, skip);
    FEM_HIP(hipGetLastError());
}

void Learned::apply_raw(const void *x, void *y, const int *skip)
{
    if (op->vec_prec == FEM_F64) learned_apply_t<double>(*this, (const double *)x, (double *)y, skip);
    else learned_apply_t<float>(*this, (const float *)x, (float *)y, skip);
}

// ---- training data ---------------------------------------------------------------------
__global__ void k_gather_coarse(const double *__restrict__ x, int nb, int E, float *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) out[e] = (float)x[(size_t)e * nb];
}

// e_k = x*_c - x_k,c  (in place over the stored snapshots)
__global__ void k_err_from_snap(const double *__restrict__ xstar, int nb, int E, float *__restrict__ snaps, int nsnap)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float xs = (float)xstar[(size_t)e * nb];
    for (int k = 0; k < nsnap; ++k) snaps[(size_t)k * E + e] = xs - snaps[(size_t)k * E + e];
}

void cnn_set_coefficient(CNN &c, const double *kcell, double beta)
{
    c.set_problem(kcell, beta);
}

void cnn_coarse_solve(const CNN &c, const float *b, int steps, float *x, double rz[2])
{
    FEM_CHECK(steps >= 0 && steps <= 100000, FEM_ERR_ARG, "coarse solve: steps in [0, 100000]");
    FEM_CHECK(c.stencil.p, FEM_ERR_STATE, "coarse solve: set the network's coefficient first");
    hipStream_t s = c.ctx->stream;
    CoarseCG cg;
    cg.init(&c);
    cg.start<float>(b, 1, nullptr, nullptr, 0.f, s, nullptr);
    for (int k = 0; k < steps; ++k) cg.step(s, nullptr);
    FEM_HIP(hipMemcpyAsync(x, cg.x.p, sizeof(float) * c.V, hipMemcpyDeviceToDevice, s));
    cg.gammas(s, rz[0], rz[1]);
}

int cnn_collect_pairs(CNN &c, Operator *op64, Precond *bj64, const Vec &b, int max_pairs, int every)
{
    FEM_CHECK(bj64 && (bj64->kind == FEM_PRE_BLOCK_JACOBI || bj64->kind == FEM_PRE_MULTIGRID) && bj64->op->N == op64->N &&
                  bj64->prec == FEM_F64,
              FEM_ERR_ARG, "collect: an FP64 block-Jacobi or multigrid preconditioner of the operator");
    const bool post_mg = bj64->kind == FEM_PRE_MULTIGRID;
    FEM_CHECK(op64->vec_prec == FEM_F64 && op64->is_dg() && op64->E == c.V, FEM_ERR_ARG,
              "collect: FP64 DG operator on the network grid");
    FEM_CHECK(max_pairs > 0 && every > 0, FEM_ERR_ARG, "collect: max_pairs, every");
    const int E = (int)c.V;
    const int nb = op64->k * op64->k * op64->k;
    const int need = c.ds_n + max_pairs;
    if (need > c.ds_cap) {
        // grow the device dataset (keeps the existing pairs)
        DevBuf nr, ne;
        nr.alloc(sizeof(float) * (size_t)need * E);
        ne.alloc(sizeof(float) * (size_t)need * E);
        if (c.ds_n) {
            FEM_HIP(hipMemcpy(nr.p, c.ds_r.p, sizeof(float) * (size_t)c.ds_n * E, hipMemcpyDeviceToDevice));
            FEM_HIP(hipMemcpy(ne.p, c.ds_e.p, sizeof(float) * (size_t)c.ds_n * E, hipMemcpyDeviceToDevice));
        }
        std::swap(c.ds_r.p, nr.p); std::swap(c.ds_r.bytes, nr.bytes);
        std::swap(c.ds_e.p, ne.p); std::swap(c.ds_e.bytes, ne.bytes);
        c.ds_cap = need;
    }
    Solver S;
    fem_solver_params prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.method = FEM_SOLVER_CG;
    prm.tolerance = 1e-10;
    prm.errormeasure = FEM_ERRMEASURE_RELATIVE;
    prm.maxiterations = 200000;
    prm.restart = 20;
    prm.residual_check = 1;
    S.init(op64->ctx, prm);
    S.bind(op64, bj64, nullptr, nullptr);
    const int first = c.ds_n;
    int taken = 0;
    S.snap_chunk = every;
    DevBuf mz, mw, mt;
    if (post_mg) {
        mz.alloc(b.bytes());
        mw.alloc(b.bytes());
        mt.alloc(b.bytes());
    }
    S.snapshot = [&](const void *x, const void *r) {
        if (taken >= max_pairs) return;
        const int k = c.ds_n + taken;
        hipStream_t st = op64->ctx->stream;
        hipLaunchKernelGGL(k_gather_coarse, dim3((E + 255) / 256), dim3(256), 0, st, (const double *)x, nb, E,
                           c.ds_e.as<float>() + (size_t)k * E);
        const double *rr = (const double *)r;
        if (post_mg) {
            // what one V-cycle leaves of the residual: r - A MG(r)
            bj64->apply_raw(r, mz.p, nullptr);
            op64->apply_raw(mz.p, mw.p, nullptr);
            hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(b.n, 256, 8192)), dim3(256), 0, st, rr,
                               (const double *)mw.p, mt.as<double>(), b.n, (const int *)nullptr);
            rr = mt.as<double>();
        }
        hipLaunchKernelGGL(k_gather_coarse, dim3((E + 255) / 256), dim3(256), 0, st, rr, nb, E,
                           c.ds_r.as<float>() + (size_t)k * E);
        ++taken;
    };
    DevBuf xb;
    xb.alloc(b.bytes());
    FEM_HIP(hipMemsetAsync(xb.p, 0, b.bytes(), op64->ctx->stream));
    Vec x;
    x.ctx = op64->ctx; x.n = b.n; x.prec = FEM_F64; x.data = xb.p; x.owner = false;
    fem_solver_info info;
    S.solve(b, x, info);
    FEM_CHECK(info.converged, FEM_ERR_STATE, "collect: the plain solve did not converge");
    hipLaunchKernelGGL(k_err_from_snap, dim3((E + 255) / 256), dim3(256), 0, op64->ctx->stream, (const double *)x.data, nb,
                       E, c.ds_e.as<float>() + (size_t)first * E, taken);
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(op64->ctx->stream));
    c.ds_n += taken;
    return taken;
}

// ---- training --------------------------------------------------------------------------------
// per pair p: out[3p] = |r|^2, out[3p+1] = |t|^2, out[3p+2] = r.t
__global__ void __launch_bounds__(256) k_pair_stats(const float *__restrict__ r, const float *__restrict__ t, int E,
                                                    double *out)
{
    __shared__ double lds[4];
    const int p = blockIdx.x;
    double a = 0, b = 0, c = 0;
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
        const double rv = r[(size_t)p * E + i], tv = t[(size_t)p * E + i];
        a += rv * rv;
        b += tv * tv;
        c += rv * tv;
    }
    const double A = block_sum(a, lds);
    __syncthreads();
    const double B = block_sum(b, lds);
    __syncthreads();
    const double C = block_sum(c, lds);
    if (threadIdx.x == 0) {
        out[3 * p] = A;
        out[3 * p + 1] = B;
        out[3 * p + 2] = C;
    }
}

// energy objective of the correction c = k y (k = s0 gamma0, Ay = A_c y):
//   L = (c^T A_c c - 2 c^T r) / w + 1,  g4 = dL/dy = 2 k (k Ay - r) / w
// per-block partial sums of L - 1 into part
__global__ void __launch_bounds__(256) k_energy_grad(int V, const float *__restrict__ y, const float *__restrict__ Ay,
                                                     const float *__restrict__ r, const float *__restrict__ red,
                                                     float gamma0, double w, float *__restrict__ g4,
                                                     double *__restrict__ part)
{
    const float k = red[0] * gamma0;
    const float sw = (float)(2.0 / w) * k;
    double acc = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const float ay = k * Ay[v];
        g4[v] = sw * (ay - r[v]);
        acc += (double)k * y[v] * ((double)ay - 2.0 * r[v]);
    }
    __shared__ double lds[4];
    const double tot = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = tot / w;
}

// e = s0 gamma0 y - t
__global__ void k_corr_err(int V, const float *__restrict__ y, const float *__restrict__ t,
                           const float *__restrict__ red, float gamma0, float *__restrict__ e)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) e[v] = red[0] * gamma0 * y[v] - t[v];
}

__global__ void __launch_bounds__(256) k_dot_parts(int V, const float *__restrict__ a, const float *__restrict__ b,
                                                   double *__restrict__ part)
{
    double acc = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) acc += (double)a[v] * b[v];
    __shared__ double lds[4];
    const double tot = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

static double sum_parts(const DevBuf &part, int n, hipStream_t s)
{
    std::vector<double> h(n);
    FEM_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    double t = 0;
    for (double v : h) t += v;
    return t;
}

void cnn_train(CNN &c, int epochs, int batch, float lr, float val_frac, unsigned seed, double &train_loss,
               double &val_rel)
{
    FEM_CHECK(c.ds_n >= 2, FEM_ERR_STATE, "train: collect pairs first");
    FEM_CHECK(c.kinv.p && c.stencil.p, FEM_ERR_STATE, "train: coefficient not set");
    FEM_CHECK(epochs >= 0 && batch >= 1 && val_frac >= 0.f && val_frac < 1.f, FEM_ERR_ARG,
              "train: epochs >= 0, batch >= 1, 0 <= val_frac < 1");
    hipStream_t s = c.ctx->stream;
    const int E = (int)c.V, P = c.ds_n;
    // targets c*_p = A_c^-1 R r_p by coarse CG to a relative (r, z) of 1e-12.
    // The pairs' solves are independent chains of small latency-bound
    // launches: NS of them run at a time, each on a stream of its own
    // (default priority), each checked every 32 steps as one solve alone.
    DevBuf tg;
    tg.alloc(sizeof(float) * (size_t)P * E);
    constexpr int NS = 4;
    struct Streams {
        hipStream_t h[NS] = {};
        ~Streams()
        {
            for (hipStream_t x : h)
                if (x) {
                    (void)hipStreamSynchronize(x);
                    (void)hipStreamDestroy(x);
                }
        }
    } ss;
    CoarseCG cgs[NS];
    for (int i = 0; i < NS; ++i) {
        FEM_HIP(hipStreamCreate(&ss.h[i]));
        cgs[i].init(&c);
    }
    FEM_HIP(hipStreamSynchronize(s));   // the pairs' residuals are in place
    for (int p0 = 0; p0 < P; p0 += NS) {
        const int nb = std::min(NS, P - p0);
        bool done[NS] = {};
        for (int i = 0; i < nb; ++i)
            cgs[i].start<float>(c.ds_r.as<float>() + (size_t)(p0 + i) * E, 1, nullptr, nullptr, 0.f, ss.h[i], nullptr);
        for (int it = 0; it < 100000; it += 32) {
            for (int i = 0; i < nb; ++i)
                if (!done[i])
                    for (int k = 0; k < 32; ++k) cgs[i].step(ss.h[i], nullptr);
            bool all = true;
            for (int i = 0; i < nb; ++i) {
                if (done[i]) continue;
                double gnow, gfirst;
                cgs[i].gammas(ss.h[i], gnow, gfirst);
                if (!(gnow > 1e-12 * gfirst)) done[i] = true;
                else all = false;
            }
            if (all) break;
        }
        for (int i = 0; i < nb; ++i)
            FEM_HIP(hipMemcpyAsync(tg.as<float>() + (size_t)(p0 + i) * E, cgs[i].x.p, sizeof(float) * E,
                                   hipMemcpyDeviceToDevice, ss.h[i]));
    }
    for (int i = 0; i < NS; ++i) FEM_HIP(hipStreamSynchronize(ss.h[i]));
    DevBuf st;
    st.alloc(sizeof(double) * 3 * P);
    hipLaunchKernelGGL(k_pair_stats, dim3(P), dim3(256), 0, s, c.ds_r.as<float>(), tg.as<float>(), E, st.as<double>());
    std::vector<double> hs(3 * P);
    FEM_HIP(hipMemcpyAsync(hs.data(), st.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    // output scale gamma0 = median ||c*|| / ||R r||; pairs with a zero
    // residual are dropped
    std::vector<double> ratio;
    std::vector<int> idx;
    for (int p = 0; p < P; ++p)
        if (hs[3 * p] > 0 && hs[3 * p + 2] > 0) {
            ratio.push_back(std::sqrt(hs[3 * p + 1] / hs[3 * p]));
            idx.push_back(p);
        }
    FEM_CHECK(!ratio.empty(), FEM_ERR_STATE, "train: all residual pairs are zero");
    std::sort(ratio.begin(), ratio.end());
    c.gamma0 = (float)ratio[ratio.size() / 2];
    // deterministic shuffle and split
    uint64_t rs = 0x243F6A8885A308D3ull ^ seed;
    auto rnd = [&]() { rs = rs * 6364136223846793005ull + 1442695040888963407ull; return (unsigned)(rs >> 33); };
    for (int i = (int)idx.size() - 1; i > 0; --i) std::swap(idx[i], idx[rnd() % (i + 1)]);
    const int PP = (int)idx.size();
    const int nval = val_frac > 0.f ? std::max(1, (int)(PP * val_frac)) : 0;
    const int ntr = PP - nval;
    FEM_CHECK(ntr >= 1, FEM_ERR_ARG, "train: no training pairs left after the validation split");
    // a fresh network starts from a zero output layer (correction 0, loss 1)
    if (c.adam_t == 0) {
        std::vector<float> w(c.nparams);
        c.get_weights(w.data());
        std::fill(w.begin() + c.woff[CNN_LAYERS - 1], w.end(), 0.f);
        c.set_weights(w.data());
    }
    c.feat_red.alloc(sizeof(float) * 2);
    DevBuf part, cnt, ay;
    part.alloc(sizeof(double) * 2 * LR_BLOCKS);
    cnt.alloc(sizeof(unsigned));
    ay.alloc(sizeof(float) * E);
    FEM_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned), s));
    auto prep = [&](int p) {
        const float *r = c.ds_r.as<float>() + (size_t)p * E;
        hipLaunchKernelGGL(k_coarse_rms<float>, dim3(LR_BLOCKS), dim3(256), 0, s, r, 1, c.kinv.as<float>(), E,
                           part.as<double>(), cnt.as<unsigned>(), c.feat_red.as<float>(), (const int *)nullptr);
        hipLaunchKernelGGL(k_coarse_feat<float>, dim3((E + 255) / 256), dim3(256), 0, s, r, 1, c.kinv.as<float>(), E,
                           c.feat_red.as<float>(), c.x0.as<float>(), (const int *)nullptr);
        c.forward(c.x0.as<float>(), c.y.as<float>(), s);
        c.coarse_apply(c.y.as<float>(), ay.as<float>(), s);
    };
    train_loss = 1.0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = ntr - 1; i > 0; --i) std::swap(idx[i], idx[rnd() % (i + 1)]);
        double tl = 0;
        for (int b0 = 0; b0 < ntr; b0 += batch) {
            const int nbat = std::min(batch, ntr - b0);
            c.zero_grad(s);
            for (int k = 0; k < nbat; ++k) {
                const int p = idx[b0 + k];
                prep(p);
                hipLaunchKernelGGL(k_energy_grad, dim3(256), dim3(256), 0, s, E, c.y.as<float>(), ay.as<float>(),
                                   c.ds_r.as<float>() + (size_t)p * E, c.feat_red.as<float>(), c.gamma0,
                                   hs[3 * p + 2], c.g4.as<float>(), c.lossp.as<double>());
                c.backward_from_g4(s);
                if (ep == epochs - 1) tl += 1.0 + sum_parts(c.lossp, 256, s);
            }
            c.adam_step(lr, 1.0f / nbat, s);
        }
        if (ep == epochs - 1) train_loss = tl / ntr;
    }
    // validation: relative coarse energy error (c - c*)^T A_c (c - c*) / (r^T c*)
    double vr = 0;
    DevBuf e;
    e.alloc(sizeof(float) * E);
    for (int k = ntr; k < PP; ++k) {
        const int p = idx[k];
        prep(p);
        hipLaunchKernelGGL(k_corr_err, dim3((E + 255) / 256), dim3(256), 0, s, E, c.y.as<float>(),
                           tg.as<float>() + (size_t)p * E, c.feat_red.as<float>(), c.gamma0, e.as<float>());
        c.coarse_apply(e.as<float>(), ay.as<float>(), s);
        hipLaunchKernelGGL(k_dot_parts, dim3(256), dim3(256), 0, s, E, e.as<float>(), ay.as<float>(), part.as<double>());
        vr += sum_parts(part, 256, s) / hs[3 * p + 2];
    }
    val_rel = nval ? vr / nval : -1.0;
    FEM_HIP(hipGetLastError());
}

}  // namespace femamd
