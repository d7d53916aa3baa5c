// Device-driven Krylov solvers.
//
//   CG     LinearSolver::cg (dune/fem/solver/linear/cg.hh:19-117): the same
//          recurrence (r = Ax - b, p = b - Ax, q = M^-1 p, ...), the same
//          stopping rule on <p, M^-1 p> against eps^2 * (1 | ||b||^2 | r0)
//          and the signed iteration count.
//   GMRES  LinearSolver::gmres (dune/fem/solver/linear/gmres.hh:117-299):
//          restarted, right preconditioned, classical Gram-Schmidt, Givens
//          rotations, M^-1 applied once to V y at the end of a cycle.
//   FGMRES flexible variant storing z_j = M^-1 v_j (no in-tree counterpart;
//          dune-istl's RestartedFlexibleGMResSolver is reached through
//          dune/fem/solver/istlinverseoperators.hh:176-178).
//
// All scalars (alpha, beta, Hessenberg matrix, rotations) live in device
// memory and are updated by the last block of each reduction kernel, so a
// Krylov step needs no host round trip.  Every kernel reads a device `done`
// flag and returns at once when the solve has finished; the host launches
// steps in batches and reads the state back once per batch (CG) or once per
// restart cycle (GMRES).
#include <chrono>
#include <cmath>
#include <cstdio>

#include "fem_blas.hpp"
#include "fem_learned.hpp"
#include "fem_solver.hpp"

namespace femamd {

constexpr int GRP = 8;   // vectors per register group in the multi-dot
// the CG reductions use RED_DOT_BLOCKS partial slots per quantity (two)
static_assert(2 * RED_DOT_BLOCKS <= RED_BLOCKS * (GM_MAX + 1), "solver partials buffer");

__device__ __forceinline__ void fin_cg_alpha(KState *st, double qh)
{
    st->qh = qh;
    st->alpha = st->res / qh;
}

__device__ __forceinline__ void fin_cg_res(KState *st, double r, double rr)
{
    st->prev = st->res;
    st->res = r;
    st->rr = rr;
    st->beta = r / st->prev;
    st->it += 1;
    st->done = !((st->crit_res ? rr : r) > st->tol) || st->it >= st->maxit;
}

// ---- CG kernels ---------------------------------------------------------------
template <typename T>
__global__ void k_cg_setup(const T *__restrict__ b, const T *__restrict__ h, T *__restrict__ r, T *__restrict__ p,
                           int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T hb = h[i], bb = b[i];
        r[i] = hb - bb;
        p[i] = bb - hb;
    }
}

__global__ void k_cg_init(KState *st, double eps, int criteria, int maxit, int crit_res)
{
    const double r0 = crit_res ? st->rr : st->res;
    const double ref = criteria == FEM_ERRMEASURE_RELATIVE ? st->bnorm2 : criteria == FEM_ERRMEASURE_REDUCTION ? r0 : 1.0;
    st->tol = eps * eps * ref;
    st->it = 0;
    st->maxit = maxit;
    st->crit_res = crit_res;
    st->done = !(r0 > st->tol) || maxit <= 0;
}

template <typename T>
__global__ void k_cg_qupd(T *__restrict__ q, const T *__restrict__ s, T *__restrict__ p, const T *__restrict__ r,
                          int64_t n, const KState *__restrict__ st, int haspre)
{
    if (st->done || st->it == 0) return;
    const double beta = st->beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (haspre) q[i] = (T)(beta * (double)q[i] + (double)s[i]);
        else p[i] = (T)(beta * (double)p[i] - (double)r[i]);
    }
}

// reduction kernels share one pattern: grid-stride update + partial dot,
// block sum, deterministic finish by the last block
template <typename T, int MODE>
__global__ void __launch_bounds__(RED_THREADS)
    k_cg_step(T *__restrict__ x, T *__restrict__ p, const T *__restrict__ q, const T *__restrict__ h,
              T *__restrict__ s, T *__restrict__ r, const T *__restrict__ dinv, int64_t n, KState *st,
              double *partials, unsigned *counter)
{
    if (st->done) return;
    __shared__ double lds[RED_THREADS / 64];
    double acc = 0, acc2 = 0;
    const double alpha = st->alpha;
    int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (MODE == 0 || MODE == 3) {
        // pure reductions: four independent loads of each operand in flight
        // per thread before they are used
        const T *A = MODE == 0 ? q : p;
        const T *Bv = MODE == 0 ? h : s;
        constexpr int U = 4;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (; i0 + (U - 1) * stride < n; i0 += U * stride) {
            T va[U], vb[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                va[k] = A[i0 + k * stride];
                vb[k] = Bv[i0 + k * stride];
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                acc += (double)va[k] * (double)vb[k];
                if (MODE == 3) acc2 += (double)va[k] * (double)va[k];
            }
        }
    }
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (MODE == 0) {          // dot(q, h) -> alpha
            acc += (double)q[i] * (double)h[i];
        } else if (MODE == 1) {   // pointwise Jacobi: x += a q, p -= a h, s = d p, <p, s>
            x[i] = (T)((double)x[i] + alpha * (double)q[i]);
            const double pn = (double)p[i] - alpha * (double)h[i];
            const T pt = (T)pn;
            p[i] = pt;
            const T sv = dinv[i] * pt;
            s[i] = sv;
            acc += (double)pt * (double)sv;
            acc2 += (double)pt * (double)pt;
        } else if (MODE == 2) {   // general M: x += a q, p -= a h (no reduction)
            x[i] = (T)((double)x[i] + alpha * (double)q[i]);
            p[i] = (T)((double)p[i] - alpha * (double)h[i]);
        } else if (MODE == 3) {   // <p, s> after a general M
            acc += (double)p[i] * (double)s[i];
            acc2 += (double)p[i] * (double)p[i];
        } else {                  // no preconditioner: x += a p, r += a h, <r, r>
            x[i] = (T)((double)x[i] + alpha * (double)q[i]);
            const T rn = (T)((double)r[i] + alpha * (double)h[i]);
            r[i] = rn;
            acc += (double)rn * (double)rn;
            acc2 = acc;
        }
    }
    if (MODE == 2) return;
    double b = block_sum(acc, lds);
    __syncthreads();
    double b2 = block_sum(acc2, lds);
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        publish_partial(&partials[blockIdx.x], b);
        publish_partial(&partials[gridDim.x + blockIdx.x], b2);
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    double v = 0, v2 = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
        v += read_partial(&partials[k]);
        v2 += read_partial(&partials[gridDim.x + k]);
    }
    double tot = block_sum(v, lds);
    __syncthreads();
    double tot2 = block_sum(v2, lds);
    if (threadIdx.x == 0) {
        if (MODE == 0) fin_cg_alpha(st, tot);
        else fin_cg_res(st, tot, tot2);
        reset_counter(counter);
    }
}

// Finish of a dot whose per-workgroup partials were written by the DG apply
// (MODE 0: <q, h> -> alpha) or by block Jacobi (MODE 3: <p, s> and <p, p> ->
// beta, residual, convergence).  The partials are summed in a fixed order, so
// the result is reproducible run to run.
template <int MODE>
__global__ void __launch_bounds__(RED_THREADS)
    k_cg_fin_parts(const double *__restrict__ ext, int64_t nb, KState *st, double *partials, unsigned *counter)
{
    if (st->done) return;
    __shared__ double lds[RED_THREADS / 64];
    double acc = 0, acc2 = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
        acc += ext[i];
        if (MODE == 3) acc2 += ext[nb + i];
    }
    double b = block_sum(acc, lds);
    __syncthreads();
    double b2 = block_sum(acc2, lds);
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        publish_partial(&partials[blockIdx.x], b);
        publish_partial(&partials[gridDim.x + blockIdx.x], b2);
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    double v = 0, v2 = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
        v += read_partial(&partials[k]);
        v2 += read_partial(&partials[gridDim.x + k]);
    }
    double tot = block_sum(v, lds);
    __syncthreads();
    double tot2 = block_sum(v2, lds);
    if (threadIdx.x == 0) {
        if (MODE == 0) fin_cg_alpha(st, tot);
        else fin_cg_res(st, tot, tot2);
        reset_counter(counter);
    }
}

// ---- GMRES kernels --------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    k_gm_start(T *__restrict__ v0, const T *__restrict__ b, int64_t n, KState *st, double *partials,
               unsigned *counter, int first)
{
    __shared__ double lds[RED_THREADS / 64];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T v = (T)((double)v0[i] - (double)b[i]);
        v0[i] = v;
        acc += (double)v * (double)v;
    }
    double bs = block_sum(acc, lds);
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        publish_partial(&partials[blockIdx.x], bs);
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    double v = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) v += read_partial(&partials[k]);
    double tot = block_sum(v, lds);
    if (threadIdx.x == 0) {
        const double res = sqrt(tot);
        if (first && st->criteria == FEM_ERRMEASURE_REDUCTION) st->gtol *= res;
        st->res = res;
        st->jlast = 0;
        if (res <= st->gtol * (1 + 1e-15)) {
            st->converged = 1;
            st->cycle_done = 1;
        } else {
            st->converged = 0;
            st->cycle_done = st->it >= st->maxit ? 1 : 0;
            st->g[0] = -res;
            for (int i = 1; i <= GM_MAX; ++i) st->g[i] = 0.0;
            st->hinv = 1.0 / res;
        }
        reset_counter(counter);
    }
}

template <typename T>
__global__ void k_gm_scale(T *__restrict__ v, int64_t n, const KState *__restrict__ st)
{
    if (st->cycle_done) return;
    const double a = st->hinv;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (T)((double)v[i] * a);
}

// H[l][j] = <v_l, w>, l = 0..j  (j+1 dots, w read once per group of GRP)
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    k_gm_multidot(const T *__restrict__ V, const T *__restrict__ w, int64_t n, int j, KState *st, double *partials,
                  unsigned *counter, int m)
{
    if (st->cycle_done) return;
    __shared__ double lds[RED_THREADS / 64];
    const int cnt = j + 1;
    // contiguous slice per block
    const int64_t lo = n * blockIdx.x / gridDim.x, hi = n * (blockIdx.x + 1) / gridDim.x;
    for (int g0 = 0; g0 < cnt; g0 += GRP) {
        double acc[GRP];
#pragma unroll
        for (int l = 0; l < GRP; ++l) acc[l] = 0;
        // four elements per thread and step, so that 4 (GRP + 1) loads are in
        // flight per thread (the grid has only two waves per SIMD)
        int64_t i = lo + threadIdx.x;
        for (; i + 3 * (int64_t)blockDim.x < hi; i += 4 * (int64_t)blockDim.x) {
            double wi[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) wi[u] = (double)w[i + u * blockDim.x];
#pragma unroll
            for (int l = 0; l < GRP; ++l)
                if (g0 + l < cnt) {
                    const T *vl = V + (int64_t)(g0 + l) * n + i;
                    double t = 0;
#pragma unroll
                    for (int u = 0; u < 4; ++u) t = fma((double)vl[u * blockDim.x], wi[u], t);
                    acc[l] += t;
                }
        }
        for (; i < hi; i += blockDim.x) {
            const double wi = (double)w[i];
#pragma unroll
            for (int l = 0; l < GRP; ++l)
                if (g0 + l < cnt) acc[l] += (double)V[(int64_t)(g0 + l) * n + i] * wi;
        }
#pragma unroll
        for (int l = 0; l < GRP; ++l) {
            if (g0 + l < cnt) {
                double b = block_sum(acc[l], lds);
                if (threadIdx.x == 0) publish_partial(&partials[(size_t)(g0 + l) * gridDim.x + blockIdx.x], b);
                __syncthreads();
            }
        }
    }
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    // one wavefront per dot: lane k sums blocks k, k + 64, ... (independent
    // reads in flight), then a butterfly over the wavefront; the dots are
    // spread over the block's wavefronts, so they finish concurrently
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int l = wave; l < cnt; l += nw) {
        double v = 0;
        for (unsigned k = lane; k < gridDim.x; k += 64) v += read_partial(&partials[(size_t)l * gridDim.x + k]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) st->H[l * m + j] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) reset_counter(counter);
}

__device__ __forceinline__ void rot(double &x, double &y, double c, double s)
{
    const double a = x, b = y;
    x = c * a + s * b;
    y = c * b - s * a;
}

// w -= sum_l H[l][j] v_l ; ||w||^2 -> H[j+1][j], Givens rotation, convergence
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    k_gm_orth(const T *__restrict__ V, T *__restrict__ w, int64_t n, int j, KState *st, double *partials,
              unsigned *counter, int m)
{
    if (st->cycle_done) return;
    __shared__ double lds[RED_THREADS / 64];
    __shared__ double h[GM_MAX + 1];
    for (int l = threadIdx.x; l <= j; l += blockDim.x) h[l] = st->H[l * m + j];
    __syncthreads();
    double acc = 0;
    const int64_t lo = n * blockIdx.x / gridDim.x, hi = n * (blockIdx.x + 1) / gridDim.x;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        double v = (double)w[i];
        // unrolled so that several basis-vector loads are in flight
#pragma unroll 8
        for (int l = 0; l <= j; ++l) v -= h[l] * (double)V[(int64_t)l * n + i];
        const T vt = (T)v;
        w[i] = vt;
        acc += (double)vt * (double)vt;
    }
    double bs = block_sum(acc, lds);
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        publish_partial(&partials[blockIdx.x], bs);
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    double v = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) v += read_partial(&partials[k]);
    double tot = block_sum(v, lds);
    if (threadIdx.x == 0) {
        double *H = st->H;
        const double hn = sqrt(tot);
        H[(j + 1) * m + j] = hn;
        st->hinv = 1.0 / hn;
        for (int i = 0; i < j; ++i) rot(H[(i + 1) * m + j], H[i * m + j], st->cs[i], st->sn[i]);
        const double a = H[j * m + j], b = H[(j + 1) * m + j];
        const double nrm = sqrt(a * a + b * b);
        st->cs[j] = a / nrm;
        st->sn[j] = -b / nrm;
        rot(H[(j + 1) * m + j], H[j * m + j], st->cs[j], st->sn[j]);
        rot(st->g[j + 1], st->g[j], st->cs[j], st->sn[j]);
        st->it += 1;
        st->jlast = j + 1;
        if (fabs(st->g[j + 1]) < st->gtol) {
            st->converged = 1;
            st->cycle_done = 1;
        } else if (st->it >= st->maxit || j + 1 == m) {
            st->cycle_done = 1;
        }
        reset_counter(counter);
    }
}

__global__ void k_gm_backsolve(KState *st, int m)
{
    const int last = st->jlast;
    const double *H = st->H;
    for (int i = last - 1; i >= 0; --i) {
        double s = 0;
        for (int k = i + 1; k < last; ++k) s += H[i * m + k] * st->y[k];
        st->y[i] = (st->g[i] - s) / H[i * m + i];
    }
}

// u = beta_u * u + sum_{i<last} y_i Z_i
template <typename T, typename TU>
__global__ void k_gm_update(TU *__restrict__ u, const T *__restrict__ Z, int64_t n, const KState *__restrict__ st,
                            int zero_first)
{
    __shared__ double y[GM_MAX];
    const int last = st->jlast;
    for (int l = threadIdx.x; l < last; l += blockDim.x) y[l] = st->y[l];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = zero_first ? 0.0 : (double)u[i];
        for (int l = 0; l < last; ++l) v += y[l] * (double)Z[(int64_t)l * n + i];
        u[i] = (TU)v;
    }
}

template <typename T>
__global__ void k_add(T *__restrict__ u, const T *__restrict__ z, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        u[i] = u[i] + z[i];
}

// ---- BiCGStab kernels --------------------------------------------------------------
// LinearSolver::bicgstab (dune/fem/solver/linear/bicgstab.hh:56-222), right
// preconditioned; the scalars stay on the device.
// nu = <r, r*>, tolerance by criterion (relative: eps ||b||, reduction:
// eps sqrt(nu), absolute: eps)
__global__ void k_bicg_init(KState *st, double eps, int criteria, int maxit)
{
    const double nu = st->nu;
    st->tol = criteria == FEM_ERRMEASURE_RELATIVE ? eps * sqrt(st->bnorm2)
            : criteria == FEM_ERRMEASURE_REDUCTION ? eps * sqrt(nu) : eps;
    st->it = 0;
    st->maxit = maxit;
    st->done = maxit <= 0;
    st->fin = 0;
}

// alpha = nu / <tmp, r*>
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    k_bicg_alpha(const T *__restrict__ tmp, const T *__restrict__ rs, int64_t n, KState *st, double *partials,
                 unsigned *counter)
{
    if (st->done) return;
    __shared__ double lds[RED_THREADS / 64];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += (double)tmp[i] * (double)rs[i];
    double own = block_sum(acc, lds);
    __shared__ double res[1];
    if (finish_reduction(0.0, partials, counter, res, 1, &own) && threadIdx.x == 0) st->alpha = st->nu / res[0];
}

// s = r - alpha tmp
template <typename T>
__global__ void k_bicg_s(const T *__restrict__ r, const T *__restrict__ tmp, T *__restrict__ s, int64_t n,
                         const KState *__restrict__ st)
{
    if (st->done) return;
    const double a = st->alpha;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s[i] = (T)((double)r[i] - a * (double)tmp[i]);
}

// the five inner products of scalarProductVecs (bicgstab.hh:19-54) and the
// scalars of the step: omega, the residual estimate, beta, nu; fin marks the
// last iteration (res < tol or it >= maxit)
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    k_bicg_dots(const T *__restrict__ r, const T *__restrict__ s, const T *__restrict__ rs, int64_t n, KState *st,
                double *partials, unsigned *counter)
{
    if (st->done) return;
    __shared__ double lds[RED_THREADS / 64];
    double a[5] = { 0, 0, 0, 0, 0 };
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = (double)r[i], si = (double)s[i], ti = (double)rs[i];
        a[0] += ri * si;
        a[1] += ri * ri;
        a[2] += si * si;
        a[3] += si * ti;
        a[4] += ri * ti;
    }
    double own[5];
    for (int k = 0; k < 5; ++k) {
        own[k] = block_sum(a[k], lds);
        __syncthreads();
    }
    __shared__ double g[5];
    if (finish_reduction(0.0, partials, counter, g, 5, own) && threadIdx.x == 0) {
        const double omega = g[0] / g[1];
        const double res = sqrt(g[2] - omega * (2.0 * g[0] - omega * g[1]));
        const double beta = (g[3] - omega * g[4]) * st->alpha / (omega * st->nu);
        st->nu = g[3] - omega * g[4];
        st->omega = omega;
        st->beta = beta;
        st->bres = res;
        st->it += 1;
        st->fin = !(res >= st->tol) || st->it >= st->maxit;
    }
}

// x += alpha p + omega s; unless this was the last iteration:
// r = s - omega r, p = beta (p - omega tmp) + r
template <typename T>
__global__ void k_bicg_update(T *__restrict__ x, T *__restrict__ p, T *__restrict__ r, const T *__restrict__ s,
                              const T *__restrict__ tmp, int64_t n, const KState *__restrict__ st)
{
    if (st->done) return;
    const double a = st->alpha, w = st->omega, b = st->beta;
    const bool last = st->fin;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double pi = (double)p[i], si = (double)s[i];
        x[i] = (T)((double)x[i] + a * pi + w * si);
        if (!last) {
            const double rn = si - w * (double)r[i];
            r[i] = (T)rn;
            p[i] = (T)(b * (pi - w * (double)tmp[i]) + rn);
        }
    }
}

// r = b - r
template <typename T>
__global__ void k_bicg_res(const T *__restrict__ b, T *__restrict__ r, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = (T)((double)b[i] - (double)r[i]);
}

__global__ void k_bicg_commit(KState *st)
{
    if (!st->done && st->fin) st->done = 1;
}

// ---- refinement helpers -----------------------------------------------------------
// r64 = b - h (h = A x in FP64), partial ||r||^2 -> out
__global__ void __launch_bounds__(RED_THREADS)
    k_residual64(const double *__restrict__ b, const double *__restrict__ h, double *__restrict__ r, int64_t n,
                 double *partials, unsigned *counter, double *out)
{
    __shared__ double lds[RED_THREADS / 64];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = b[i] - h[i];
        r[i] = v;
        acc += v * v;
    }
    double bs = block_sum(acc, lds);
    finish_reduction(0.0, partials, counter, out, 1, &bs);
}

__global__ void k_scale_to32(const double *__restrict__ r, float *__restrict__ r32, int64_t n, double a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r32[i] = (float)(r[i] * a);
}

__global__ void k_update64(double *__restrict__ x, const float *__restrict__ d, int64_t n, double a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += a * (double)d[i];
}

// ---------------------------------------------------------------------------------
void Solver::init(Context *c, const fem_solver_params &p)
{
    ctx = c;
    prm = p;
    FEM_CHECK(p.method >= FEM_SOLVER_CG && p.method <= FEM_SOLVER_BICGSTAB, FEM_ERR_ARG, "solver method");
    FEM_CHECK(p.restart >= 1 && p.restart <= GM_MAX, FEM_ERR_ARG, "gmres restart must be in 1..64");
    state.alloc(sizeof(KState));
    partials.alloc(sizeof(double) * RED_BLOCKS * (GM_MAX + 2));
    counters.alloc(sizeof(unsigned) * 8);
    FEM_HIP(hipMemset(counters.p, 0, counters.bytes));
    FEM_HIP(hipHostMalloc((void **)&hstate, sizeof(KState), hipHostMallocDefault));
}

Solver::~Solver()
{
    if (hstate) (void)hipHostFree(hstate);
    for (DevBuf *b : work) delete b;
}

void Solver::bind(Operator *o, Precond *p, Operator *io, Precond *ip)
{
    op = iop = nullptr;
    pre = ipre = nullptr;
    FEM_CHECK(o, FEM_ERR_ARG, "bind: null operator");
    FEM_CHECK(!o->nonlinear() && (!io || !io->nonlinear()), FEM_ERR_STATE,
              "bind: the operator is nonlinear (bind its Jacobian; the Newton solve does)");
    FEM_CHECK(!p || p->op == o || p->op->N == o->N, FEM_ERR_ARG, "preconditioner does not match the operator");
    // an FP32 preconditioner may serve an FP64 solve (rounded in, widened
    // out); otherwise the precisions agree
    FEM_CHECK(!p || p->prec == o->vec_prec || (p->prec == FEM_F32 && o->vec_prec == FEM_F64 && !prm.refinement),
              FEM_ERR_ARG, "preconditioner precision does not match the operator");
    // the learned preconditioner is nonlinear in its argument (the coarse
    // scaling depends on it): only the flexible GMRES, which stores
    // z_j = M(v_j), stays a valid Krylov method with it
    auto flexible_only = [&](const Precond *q) {
        FEM_CHECK(!q || !q->learned_part() || prm.method == FEM_SOLVER_FGMRES, FEM_ERR_UNSUPPORTED,
                  "the learned preconditioner is nonlinear: it needs fem.solver.method = fgmres");
    };
    flexible_only(p);
    if (prm.refinement) {
        FEM_CHECK(io, FEM_ERR_ARG, "refinement needs an inner operator");
        FEM_CHECK(io->N == o->N, FEM_ERR_ARG, "inner operator size");
        FEM_CHECK(o->vec_prec == FEM_F64, FEM_ERR_ARG, "refinement: the outer operator must be FP64");
        // the inner Krylov solve runs in FP32 on FP32 work vectors
        FEM_CHECK(io->vec_prec == FEM_F32, FEM_ERR_ARG, "refinement: the inner operator must be FP32 or BF16-MFMA");
        FEM_CHECK(!ip || ip->prec == io->vec_prec, FEM_ERR_ARG, "inner preconditioner precision");
        FEM_CHECK(!ip || ip->op->N == io->N, FEM_ERR_ARG, "inner preconditioner size");
        flexible_only(ip);
    }
    // fem.solver.preconditioning.method: built here when no preconditioner
    // object is passed (KrylovInverseOperator::bind creates its own the same
    // way, krylovinverseoperators.hh:86-97).  Binding the same operator again
    // (a Newton solve binds each Jacobian S'(w) to the one solver) updates
    // the one built before in place: the same result as a new one, without
    // building the hierarchy and its buffers again.
    widened.reset();
    if (p && p->prec == FEM_F32 && o->vec_prec == FEM_F64) {
        widened.reset(new Precond());
        widened->init_widening(o, p);
        p = widened.get();
    }
    auto own = [&](std::unique_ptr<Precond> &mine, Operator *target) {
        if (mine && mine->updatable_for(target, prm.preconditioning)) {
            mine->update();
        } else {
            mine.reset();
            mine.reset(new Precond());
            mine->init(target, prm.preconditioning);
        }
        return mine.get();
    };
    if (!p && prm.preconditioning != FEM_PRE_NONE && !prm.refinement) p = own(own_pre, o);
    else own_pre.reset();
    if (prm.refinement && !ip && prm.preconditioning != FEM_PRE_NONE) ip = own(own_ipre, io);
    else own_ipre.reset();
    op = o;
    pre = p;
    iop = io;
    ipre = ip;
}

void Solver::unbind()
{
    op = iop = nullptr;
    pre = ipre = nullptr;
    own_pre.reset();
    own_ipre.reset();
    widened.reset();
    for (DevBuf *b : work) delete b;
    work.clear();
}

void *Solver::workbuf(int i, size_t bytes)
{
    while ((int)work.size() <= i) work.push_back(new DevBuf());
    if (work[i]->bytes < bytes) work[i]->alloc(bytes);
    return work[i]->p;
}

void Solver::fetch_state()
{
    FEM_HIP(hipMemcpyAsync(hstate, state.p, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    FEM_HIP(hipStreamSynchronize(ctx->stream));
}

template <typename T>
static int run_krylov(Solver &S, int method, Operator *o, Precond *p, const T *b, T *x, double eps, int criteria,
                      int maxit, int m, int64_t N, int verbose)
{
    hipStream_t s = S.ctx->stream;
    KState *st = S.state.as<KState>();
    double *part = S.partials.as<double>();
    unsigned *cnt = S.counters.as<unsigned>();
    const unsigned G = grid_for(N, 256, 4096);
    const size_t vb = sizeof(T) * N;
    const bool haspre = p && p->kind != FEM_PRE_NONE;
    FEM_HIP(hipMemsetAsync(st, 0, sizeof(KState), s));
    if (method == FEM_SOLVER_CG) {
        T *h = (T *)S.workbuf(0, vb), *r = (T *)S.workbuf(1, vb), *pp = (T *)S.workbuf(2, vb);
        T *ss = haspre ? (T *)S.workbuf(3, vb) : pp, *q = haspre ? (T *)S.workbuf(4, vb) : pp;
        o->apply_raw(x, h, nullptr);
        hipLaunchKernelGGL(k_cg_setup<T>, dim3(G), dim3(256), 0, s, b, h, r, pp, N);
        if (haspre) {
            p->apply_raw(pp, q, nullptr);
            FEM_HIP(hipMemcpyAsync(ss, q, vb, hipMemcpyDeviceToDevice, s));
        }
        dot_async(*S.ctx, N, pp, q, &st->res, cnt + 1, part);
        dot_async(*S.ctx, N, pp, pp, &st->rr, cnt + 1, part);
        dot_async(*S.ctx, N, b, b, &st->bnorm2, cnt + 1, part);
        hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, s, st, eps, criteria, maxit, S.prm.residual_check);
        FEM_HIP(hipGetLastError());
        const T *dinv = (haspre && p->pointwise()) ? (p->prec == FEM_F64 ? (const T *)p->dinv.p : (const T *)p->dinv.p) : nullptr;
        // DG operator with block Jacobi: the apply and the preconditioner
        // write per-workgroup partials of the two CG dots, which a small
        // kernel finishes, instead of two more passes over the vectors
        const bool fuse = o->is_dg() && haspre && p->fused_dot() && o->desc.precision != FEM_BF16MFMA;
        const int64_t nbd = fuse ? o->dot_blocks() : 0;        // apply workgroups
        const int64_t nbj = fuse ? dg_bj_blocks(p->op->k, p->op->E) : 0;   // block-Jacobi workgroups
        // about 4 partials per thread: a small launch finishes quickly
        auto fin_grid = [](int64_t n) {
            return (unsigned)std::min<int64_t>(RED_DOT_BLOCKS, std::max<int64_t>(1, (n + 4 * RED_THREADS - 1) / (4 * RED_THREADS)));
        };
        const unsigned GF = fin_grid(nbd), GJ = fin_grid(nbj);
        double *dq = fuse ? (double *)S.workbuf(5, sizeof(double) * nbd) : nullptr;
        double *dp = fuse ? (double *)S.workbuf(6, 2 * sizeof(double) * nbj) : nullptr;
        int chunk = S.snapshot ? S.snap_chunk : 4;
        // one CG step as a HIP graph: every kernel of the step reads the
        // device done flag, so replays after convergence do nothing; the step
        // holds no memset node.  Captured per solve (the vectors differ).
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        struct GraphGuard {
            hipGraph_t &g;
            hipGraphExec_t &e;
            ~GraphGuard()
            {
                if (e) (void)hipGraphExecDestroy(e);
                if (g) (void)hipGraphDestroy(g);
            }
        } gguard{ graph, gexec };
        auto cg_step = [&]() {
                hipLaunchKernelGGL(k_cg_qupd<T>, dim3(G), dim3(256), 0, s, q, ss, pp, r, N, st, haspre ? 1 : 0);
                if (fuse) {
                    o->apply_raw(q, h, &st->done, dq);
                    hipLaunchKernelGGL(k_cg_fin_parts<0>, dim3(GF), dim3(RED_THREADS), 0, s, dq, nbd, st, part, cnt);
                    hipLaunchKernelGGL((k_cg_step<T, 2>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, x, pp, q, h, ss, r, dinv, N, st, part, cnt);
                    p->apply_raw(pp, ss, &st->done, dp);
                    hipLaunchKernelGGL(k_cg_fin_parts<3>, dim3(GJ), dim3(RED_THREADS), 0, s, dp, nbj, st, part, cnt);
                    return;
                }
                o->apply_raw(q, h, &st->done);
                hipLaunchKernelGGL((k_cg_step<T, 0>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, x, pp, q, h, ss, r, dinv, N, st, part, cnt);
                if (!haspre) {
                    hipLaunchKernelGGL((k_cg_step<T, 4>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, x, pp, q, h, ss, r, dinv, N, st, part, cnt);
                } else if (dinv) {
                    hipLaunchKernelGGL((k_cg_step<T, 1>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, x, pp, q, h, ss, r, dinv, N, st, part, cnt);
                } else {
                    hipLaunchKernelGGL((k_cg_step<T, 2>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, x, pp, q, h, ss, r, dinv, N, st, part, cnt);
                    p->apply_raw(pp, ss, &st->done);
                    hipLaunchKernelGGL((k_cg_step<T, 3>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, x, pp, q, h, ss, r, dinv, N, st, part, cnt);
                }
        };
        if (S.prm.graph && !S.snapshot) {
            FEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            cg_step();
            FEM_HIP(hipStreamEndCapture(s, &graph));
            FEM_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        while (true) {
            for (int c = 0; c < chunk; ++c) {
                if (gexec) FEM_HIP(hipGraphLaunch(gexec, s));
                else cg_step();
            }
            FEM_HIP(hipGetLastError());
            S.fetch_state();
            if (verbose) { std::printf("Fem::CG it: %d : residual %g\n", S.hstate->it, std::sqrt(S.hstate->res)); std::fflush(stdout); }
            if (S.hstate->done) break;
            if (S.snapshot) S.snapshot(x, pp);
            else if (chunk < 64) chunk *= 2;
        }
        const int it = S.hstate->it;
        return it < maxit ? it : -it;
    }
    if (method == FEM_SOLVER_BICGSTAB) {
        // bicgstab.hh:56-222: right preconditioning iterates on y with
        // x = M^-1 y; the initial guess enters as y (r = b - A M^-1 x) and
        // the result is mapped back at the end (x = M^-1 y)
        T *r = (T *)S.workbuf(0, vb), *rs = (T *)S.workbuf(1, vb), *pp = (T *)S.workbuf(2, vb);
        T *ss = (T *)S.workbuf(3, vb), *tmp = (T *)S.workbuf(4, vb);
        T *z = haspre ? (T *)S.workbuf(5, vb) : nullptr;
        if (haspre) {
            p->apply_raw(x, z, nullptr);
            o->apply_raw(z, r, nullptr);
        } else {
            o->apply_raw(x, r, nullptr);
        }
        hipLaunchKernelGGL(k_bicg_res<T>, dim3(G), dim3(256), 0, s, b, r, N);   // r = b - A x
        FEM_HIP(hipMemcpyAsync(pp, r, vb, hipMemcpyDeviceToDevice, s));
        FEM_HIP(hipMemcpyAsync(rs, r, vb, hipMemcpyDeviceToDevice, s));
        dot_async(*S.ctx, N, r, rs, &st->nu, cnt + 1, part);
        dot_async(*S.ctx, N, b, b, &st->bnorm2, cnt + 1, part);
        hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1), 0, s, st, eps, criteria, maxit);
        FEM_HIP(hipGetLastError());
        const int *skip = &st->done;
        int chunk = 4;
        while (true) {
            for (int c = 0; c < chunk; ++c) {
                if (haspre) {
                    p->apply_raw(pp, z, skip);
                    o->apply_raw(z, tmp, skip);
                } else {
                    o->apply_raw(pp, tmp, skip);
                }
                hipLaunchKernelGGL(k_bicg_alpha<T>, dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, tmp, rs, N, st, part, cnt);
                hipLaunchKernelGGL(k_bicg_s<T>, dim3(G), dim3(256), 0, s, r, tmp, ss, N, st);
                if (haspre) {
                    p->apply_raw(ss, z, skip);
                    o->apply_raw(z, r, skip);
                } else {
                    o->apply_raw(ss, r, skip);
                }
                hipLaunchKernelGGL(k_bicg_dots<T>, dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, s, r, ss, rs, N, st, part, cnt);
                hipLaunchKernelGGL(k_bicg_update<T>, dim3(G), dim3(256), 0, s, x, pp, r, ss, tmp, N, st);
                hipLaunchKernelGGL(k_bicg_commit, dim3(1), dim3(1), 0, s, st);
            }
            FEM_HIP(hipGetLastError());
            S.fetch_state();
            if (verbose) { std::printf("Fem::BiCGstab it: %d : %g\n", S.hstate->it, S.hstate->bres); std::fflush(stdout); }
            if (S.hstate->done) break;
            if (chunk < 64) chunk *= 2;
        }
        if (haspre) {
            p->apply_raw(x, z, nullptr);
            FEM_HIP(hipMemcpyAsync(x, z, vb, hipMemcpyDeviceToDevice, s));
        }
        const int it = S.hstate->it;
        return it < maxit ? it : -it;
    }
    // GMRES / FGMRES
    const bool flex = method == FEM_SOLVER_FGMRES;
    T *V = (T *)S.workbuf(0, vb * (m + 1));
    T *Z = (flex && haspre) ? (T *)S.workbuf(1, vb * m) : nullptr;
    T *z = (!flex && haspre) ? (T *)S.workbuf(2, vb) : nullptr;
    T *tmp = (!flex && haspre) ? (T *)S.workbuf(3, vb) : nullptr;
    // tolerance: eps * ||b|| for relative (gmres.hh:143-148)
    double bnorm = 0;
    if (criteria == FEM_ERRMEASURE_RELATIVE) {
        dot_async(*S.ctx, N, b, b, part + RED_BLOCKS * (GM_MAX + 1), cnt + 1, part);
        double hb;
        FEM_HIP(hipMemcpyAsync(&hb, part + RED_BLOCKS * (GM_MAX + 1), sizeof(double), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        bnorm = std::sqrt(hb);
    }
    {
        KState init;
        std::memset(&init, 0, sizeof(init));
        init.gtol = criteria == FEM_ERRMEASURE_RELATIVE ? eps * bnorm : eps;
        init.criteria = criteria;
        init.maxit = maxit;
        *S.hstate = init;
        FEM_HIP(hipMemcpyAsync(st, S.hstate, sizeof(KState), hipMemcpyHostToDevice, s));
    }
    bool first = true;
    Learned *lp = p ? p->learned_part() : nullptr;
    T *xsave = lp ? (T *)S.workbuf(4, vb) : nullptr;
    while (true) {
        if (lp && lp->active()) FEM_HIP(hipMemcpyAsync(xsave, x, vb, hipMemcpyDeviceToDevice, s));
        o->apply_raw(x, V, nullptr);
        hipLaunchKernelGGL(k_gm_start<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, V, b, N, st, part, cnt, first ? 1 : 0);
        hipLaunchKernelGGL(k_gm_scale<T>, dim3(G), dim3(256), 0, s, V, N, st);
        first = false;
        for (int j = 0; j < m; ++j) {
            T *vj = V + (int64_t)j * N, *w = V + (int64_t)(j + 1) * N;
            const int *skip = &st->cycle_done;
            if (haspre) {
                T *zj = flex ? Z + (int64_t)j * N : z;
                p->apply_raw(vj, zj, skip);
                o->apply_raw(zj, w, skip);
            } else {
                o->apply_raw(vj, w, skip);
            }
            hipLaunchKernelGGL(k_gm_multidot<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, V, w, N, j, st, part, cnt, m);
            hipLaunchKernelGGL(k_gm_orth<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, V, w, N, j, st, part, cnt, m);
            hipLaunchKernelGGL(k_gm_scale<T>, dim3(G), dim3(256), 0, s, w, N, st);
        }
        hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(1), 0, s, st, m);
        if (flex && haspre) {
            hipLaunchKernelGGL((k_gm_update<T, T>), dim3(G), dim3(256), 0, s, x, Z, N, st, 0);
        } else if (haspre) {
            hipLaunchKernelGGL((k_gm_update<T, T>), dim3(G), dim3(256), 0, s, tmp, V, N, st, 1);
            p->apply_raw(tmp, z, nullptr);
            hipLaunchKernelGGL(k_add<T>, dim3(G), dim3(256), 0, s, x, z, N);
        } else {
            hipLaunchKernelGGL((k_gm_update<T, T>), dim3(G), dim3(256), 0, s, x, V, N, st, 0);
        }
        FEM_HIP(hipGetLastError());
        S.fetch_state();
        const double gres = std::fabs(S.hstate->g[S.hstate->jlast]);
        if (verbose) { std::printf("Fem::%sGMRES it: %d : %g\n", flex ? "F" : "", S.hstate->it, gres); std::fflush(stdout); }
        if (lp && lp->active() && !S.hstate->converged && S.hstate->jlast > 0 && !(gres < 0.95 * S.hstate->res)) {
            // the learned correction stagnates or broke down: undo the cycle
            // and continue with block Jacobi only
            FEM_HIP(hipMemcpyAsync(x, xsave, vb, hipMemcpyDeviceToDevice, s));
            lp->mode = FEM_LEARNED_OFF;
            S.fallback_used = 1;
            continue;
        }
        if (S.hstate->converged || S.hstate->it >= maxit) break;
    }
    const int it = S.hstate->it;
    return it < maxit ? it : -it;
}

int Solver::krylov(int method, Operator *o, Precond *p, const void *b, void *x, double tol, int criteria, int maxit,
                   int restart, int64_t N, int prec)
{
    if (prec == FEM_F64)
        return run_krylov<double>(*this, method, o, p, (const double *)b, (double *)x, tol, criteria, maxit, restart, N, prm.verbose);
    return run_krylov<float>(*this, method, o, p, (const float *)b, (float *)x, tol, criteria, maxit, restart, N, prm.verbose);
}

double Solver::true_residual(const Vec &b, const Vec &x, Vec &r)
{
    // ||b - A x|| / ||b|| with FP64 accumulation
    op->apply(x, r);
    Vec &rr = r;
    vec_scal(rr, -1.0);
    vec_axpy(rr, 1.0, b);
    const double rn = vec_dot(rr, rr), bn = vec_dot(b, b);
    return bn > 0 ? std::sqrt(rn / bn) : std::sqrt(rn);
}

void Solver::solve(const Vec &b, Vec &x, fem_solver_info &info)
{
    FEM_CHECK(op, FEM_ERR_STATE, "solve before bind");
    FEM_CHECK(b.n == op->N && x.n == op->N, FEM_ERR_ARG, "solve: vector size");
    FEM_CHECK(b.prec == op->vec_prec && x.prec == op->vec_prec, FEM_ERR_ARG, "solve: vector precision");
    std::memset(&info, 0, sizeof(info));
    fallback_used = 0;
    // a stall fallback switches the learned preconditioner off for the rest
    // of this solve only: its mode is restored on return, so the next solve
    // runs the network again (and reports its own fallback)
    struct ModeGuard {
        Learned *l;
        int mode;
        ~ModeGuard() { if (l) l->mode = mode; }
    } g_outer{ pre ? pre->learned_part() : nullptr, pre && pre->learned_part() ? pre->learned_part()->mode : 0 },
        g_inner{ ipre ? ipre->learned_part() : nullptr, ipre && ipre->learned_part() ? ipre->learned_part()->mode : 0 };
    const int64_t N = op->N;
    FEM_HIP(hipStreamSynchronize(ctx->stream));
    auto t0 = std::chrono::steady_clock::now();
    Vec r;
    r.ctx = ctx;
    r.n = N;
    r.prec = op->vec_prec;
    r.data = workbuf(20, x.bytes());
    r.owner = false;
    if (!prm.refinement) {
        info.iterations = krylov(prm.method, op, pre, b.data, x.data, prm.tolerance, prm.errormeasure,
                                 prm.maxiterations, prm.restart, N, op->vec_prec);
        info.converged = info.iterations >= 0;
        // residualcheck holds CG to the true residual as well: when the
        // recursive residual has converged but b - A x has drifted above the
        // tolerance, CG restarts from x (at most twice)
        if (prm.residual_check && prm.method == FEM_SOLVER_CG && prm.errormeasure != FEM_ERRMEASURE_REDUCTION) {
            for (int rs = 0; rs < 2 && info.converged; ++rs) {
                const double rel = true_residual(b, x, r);
                const double bn = std::sqrt(vec_dot(b, b));
                const double res = prm.errormeasure == FEM_ERRMEASURE_RELATIVE ? rel : rel * (bn > 0 ? bn : 1.0);
                if (!(res > prm.tolerance)) break;
                const int left = prm.maxiterations - info.iterations;
                if (left <= 0) {
                    info.converged = 0;
                    info.iterations = -info.iterations;
                    break;
                }
                const int more = krylov(prm.method, op, pre, b.data, x.data, prm.tolerance, prm.errormeasure, left,
                                        prm.restart, N, op->vec_prec);
                info.iterations = more >= 0 ? info.iterations + more : -(info.iterations - more);
                info.converged = more >= 0;
            }
        }
    } else {
        // FP64 iterative refinement around a low-precision Krylov solve
        hipStream_t s = ctx->stream;
        double *h = (double *)workbuf(10, sizeof(double) * N);
        double *r = (double *)workbuf(11, sizeof(double) * N);
        float *r32 = (float *)workbuf(12, sizeof(float) * N);
        float *d32 = (float *)workbuf(13, sizeof(float) * N);
        double *part = partials.as<double>();
        unsigned *cnt = counters.as<unsigned>();
        double *dres = part + RED_BLOCKS * (GM_MAX + 1);
        const unsigned G = grid_for(N, 256, 4096);
        const double bn = std::sqrt(vec_dot(b, b));
        int total = 0, outer = 0;
        bool ok = false;
        double rn0 = 0;   // the first residual (residualreduction)
        while (true) {
            op->apply_raw(x.data, h, nullptr);
            hipLaunchKernelGGL(k_residual64, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, (const double *)b.data, h, r, N, part, cnt + 2, dres);
            double rn2;
            FEM_HIP(hipMemcpyAsync(&rn2, dres, sizeof(double), hipMemcpyDeviceToHost, s));
            FEM_HIP(hipStreamSynchronize(s));
            const double rn = std::sqrt(rn2);
            const double rel = bn > 0 ? rn / bn : rn;
            if (prm.verbose) { std::printf("IR outer %d: rel residual %g (inner its %d)\n", outer, rel, total); std::fflush(stdout); }
            if (outer == 0) rn0 = rn;
            const double target = prm.errormeasure == FEM_ERRMEASURE_RELATIVE    ? prm.tolerance * bn
                                  : prm.errormeasure == FEM_ERRMEASURE_REDUCTION ? prm.tolerance * rn0
                                                                                 : prm.tolerance;
            if (rn <= target || !(rn == rn)) { ok = rn == rn; break; }
            if (outer >= 100 || total >= prm.maxiterations) break;
            hipLaunchKernelGGL(k_scale_to32, dim3(G), dim3(256), 0, s, r, r32, N, 1.0 / rn);
            FEM_HIP(hipMemsetAsync(d32, 0, sizeof(float) * N, s));
            int its = krylov(prm.method, iop, ipre, r32, d32, prm.inner_tolerance, FEM_ERRMEASURE_RELATIVE,
                             prm.inner_maxiterations, prm.restart, N, FEM_F32);
            total += its < 0 ? -its : its;
            hipLaunchKernelGGL(k_update64, dim3(G), dim3(256), 0, s, (double *)x.data, d32, N, rn);
            FEM_HIP(hipGetLastError());
            ++outer;
        }
        info.iterations = ok ? total : -total;
        info.outer_iterations = outer;
        info.converged = ok;
    }
    FEM_HIP(hipStreamSynchronize(ctx->stream));
    auto t1 = std::chrono::steady_clock::now();
    info.time_s = std::chrono::duration<double>(t1 - t0).count();
    info.true_rel_residual = true_residual(b, x, r);
    info.fallback = fallback_used;
}

}  // namespace femamd
