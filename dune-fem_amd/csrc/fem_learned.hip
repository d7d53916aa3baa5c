// Learned coarse-space preconditioner for SIPG DG operators.
//
// Coarse space: element means (the phi_0 coefficient of the orthonormal
// Legendre basis).  Restriction R v = v[e*nb] of a dual vector, prolongation
// P c = c_e in the phi_0 slot.  The network maps the normalised coarse
// residual (channels r_c / s0 and (r_c / kappa_e) / s1, s = RMS over the
// grid) to the coarse error; its output is rescaled by s0 * gamma0, so the
// correction is positively homogeneous in the residual.
//
// Applied as a two-level method inside FGMRES (flexible, because the network
// is nonlinear):
//   multiplicative: z1 = P C(R v);  z = z1 + BJ(v - A z1)
//   additive:       z = BJ(v) + P C(R v)
// BJ is the exact element-block inverse.  Training pairs (r_k, e_k) are the
// coarse residuals and errors of the iterates of the plain CG + block-Jacobi
// solve (e_k = x* - x_k), i.e. produced by the solver itself.
//
// No reference counterpart (new preconditioner); the solve is judged by its
// FP64 true residual.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "fem_blas.hpp"
#include "fem_cnn.hpp"
#include "fem_learned.hpp"
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

// red[0] = mean(rc^2), red[1] = mean((rc * kinv)^2); rc = src[e * stride]
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
    const double A = block_sum(a, lds);
    __syncthreads();
    const double B = block_sum(b, lds);
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = A;
        partials[gridDim.x + blockIdx.x] = B;
        __threadfence();
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    __threadfence();
    double sa = 0, sb = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
        sa += __builtin_nontemporal_load(&partials[k]);
        sb += __builtin_nontemporal_load(&partials[gridDim.x + k]);
    }
    const double TA = block_sum(sa, lds);
    __syncthreads();
    const double TB = block_sum(sb, lds);
    if (threadIdx.x == 0) {
        const float s0 = (float)sqrt(TA / E), s1 = (float)sqrt(TB / E);
        red[0] = s0 > 0.f ? s0 : 1.f;
        red[1] = s1 > 0.f ? s1 : 1.f;
        *counter = 0;
    }
}

// features x0[e] = (bf16(rc / s0), bf16(rc * kinv / s1)); optional target
// t[e] = ec / (s0 * gamma0)
template <typename T>
__global__ void k_coarse_feat(const T *__restrict__ src, int stride, const float *__restrict__ kinv, int E,
                              const float *__restrict__ red, float *__restrict__ x0, const float *__restrict__ ec,
                              float gamma0, float *__restrict__ tgt, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float r = (float)src[(size_t)e * stride];
    x0[2 * e] = lround_bf(r / red[0]);
    x0[2 * e + 1] = lround_bf(r * kinv[e] / red[1]);
    if (tgt) tgt[e] = ec[e] / (red[0] * gamma0);
}

// z = 0 except z[e*nb] = c_e * s0 * gamma0
template <typename T>
__global__ void k_prolong(const float *__restrict__ c, const float *__restrict__ red, float gamma0, int nb, int64_t N,
                          T *__restrict__ z, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const float sc = red[0] * gamma0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = (i % nb == 0) ? (T)(c[i / nb] * sc) : T(0);
}

template <typename T>
__global__ void k_axmy(const T *__restrict__ v, T *__restrict__ t, int64_t N, const int *__restrict__ skip)
{
    // t = v - t
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        t[i] = v[i] - t[i];
}

template <typename T>
__global__ void k_addv(T *__restrict__ y, const T *__restrict__ z, int64_t N, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += z[i];
}

template <typename T>
__global__ void k_add_coarse(T *__restrict__ y, const float *__restrict__ c, const float *__restrict__ red,
                             float gamma0, int nb, int E, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) y[(size_t)e * nb] += (T)(c[e] * red[0] * gamma0);
}

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

// per-pair sums of squares of r and e
__global__ void k_pair_norms(const float *__restrict__ r, const float *__restrict__ e, int E, double *out)
{
    __shared__ double lds[4];
    const int p = blockIdx.x;
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
        const double rv = r[(size_t)p * E + i], ev = e[(size_t)p * E + i];
        a += rv * rv;
        b += ev * ev;
    }
    const double A = block_sum(a, lds);
    __syncthreads();
    const double B = block_sum(b, lds);
    if (threadIdx.x == 0) {
        out[2 * p] = A;
        out[2 * p + 1] = B;
    }
}

// Galerkin energy of the normalised coarse correction y (c = s0 gamma0 y):
//   L(y) = 1/2 gamma0 y^T A_c y - y^T b,  b = r_c / s0
// g4 = dL/dy / V; partial sums of L / V per block
__global__ void k_energy_grad(int V, const float *__restrict__ y, const float *__restrict__ Acy,
                              const float *__restrict__ rc, const float *__restrict__ red, float gamma0,
                              float *__restrict__ g4, double *__restrict__ part)
{
    __shared__ double sh[256];
    const float is0 = 1.0f / red[0];
    double acc = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const float b = rc[v] * is0;
        g4[v] = (gamma0 * Acy[v] - b) / (float)V;
        acc += 0.5 * (double)gamma0 * y[v] * Acy[v] - (double)y[v] * b;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// fused dots <v, z1> and <z1, w>; the last block stores alpha = <v,z1>/<z1,w>
// (0 when the direction has no positive energy) in red[2]
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
    const double A = block_sum(a, lds);
    __syncthreads();
    const double B = block_sum(b, lds);
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = A;
        partials[gridDim.x + blockIdx.x] = B;
        __threadfence();
        amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!amLast) return;
    __threadfence();
    double sa = 0, sb = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
        sa += __builtin_nontemporal_load(&partials[k]);
        sb += __builtin_nontemporal_load(&partials[gridDim.x + k]);
    }
    const double TA = block_sum(sa, lds);
    __syncthreads();
    const double TB = block_sum(sb, lds);
    if (threadIdx.x == 0) {
        red[2] = (TB > 0 && TA == TA) ? (float)(TA / TB) : 0.f;
        *counter = 0;
    }
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

// ---------------------------------------------------------------------------------
void Learned::init(Operator *o, CNN *c, int multiplicative_)
{
    op = o;
    cnn = c;
    multiplicative = multiplicative_;
    FEM_CHECK(o->is_dg(), FEM_ERR_UNSUPPORTED, "learned preconditioner needs a DG operator");
    FEM_CHECK(c->V == o->E, FEM_ERR_ARG, "learned preconditioner: network grid must match the element grid");
    FEM_CHECK((int64_t)c->n[0] == o->desc.n[0] && c->n[1] == o->desc.n[1] && c->n[2] == o->desc.n[2], FEM_ERR_ARG,
              "learned preconditioner: network grid must match the element grid");
    FEM_CHECK(c->kinv.p, FEM_ERR_STATE, "learned preconditioner: set the network's coefficient feature first");
    bj.init(o, FEM_PRE_BLOCK_JACOBI);
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
    if (!L.enabled || L.restart_mode()) {
        L.bj.apply_raw(x, y, skip);
        return;
    }
    hipLaunchKernelGGL(k_coarse_rms<T>, dim3(LR_BLOCKS), dim3(256), 0, s, x, nb, L.cnn->kinv.as<float>(), E,
                       L.partials.as<double>(), L.counter.as<unsigned>(), L.red.as<float>(), skip);
    hipLaunchKernelGGL(k_coarse_feat<T>, dim3((E + 255) / 256), dim3(256), 0, s, x, nb, L.cnn->kinv.as<float>(), E,
                       L.red.as<float>(), L.x0.as<float>(), (const float *)nullptr, 1.0f, (float *)nullptr, skip);
    L.cnn->forward(L.x0.as<float>(), L.yc.as<float>(), s, skip);
    if (L.multiplicative) {
        // z1 = alpha P C(R x) with the energy-optimal alpha; y = z1 + BJ(x - A z1)
        T *z1 = L.z1.as<T>(), *t1 = L.t1.as<T>(), *w1 = L.w1.as<T>();
        hipLaunchKernelGGL(k_prolong<T>, dim3(G), dim3(256), 0, s, L.yc.as<float>(), L.red.as<float>(), L.cnn->gamma0,
                           nb, N, z1, skip);
        o.apply_raw(z1, w1, skip);
        hipLaunchKernelGGL(k_alpha<T>, dim3(LR_BLOCKS), dim3(256), 0, s, x, z1, w1, N, L.partials.as<double>(),
                           L.counter.as<unsigned>(), L.red.as<float>(), skip);
        hipLaunchKernelGGL(k_scaled_residual<T>, dim3(G), dim3(256), 0, s, x, w1, t1, z1, N, L.red.as<float>(), skip);
        L.bj.apply_raw(t1, y, skip);
        hipLaunchKernelGGL(k_addv<T>, dim3(G), dim3(256), 0, s, y, z1, N, skip);
    } else {
        L.bj.apply_raw(x, y, skip);
        hipLaunchKernelGGL(k_add_coarse<T>, dim3((E + 255) / 256), dim3(256), 0, s, y, L.yc.as<float>(),
                           L.red.as<float>(), L.cnn->gamma0, nb, E, skip);
    }
    FEM_HIP(hipGetLastError());
}

template <typename T>
static void learned_correct_t(Learned &L, const T *r, T *x)
{
    Operator &o = *L.op;
    hipStream_t s = o.ctx->stream;
    const int nb = o.k * o.k * o.k;
    const int E = (int)o.E;
    hipLaunchKernelGGL(k_coarse_rms<T>, dim3(LR_BLOCKS), dim3(256), 0, s, r, nb, L.cnn->kinv.as<float>(), E,
                       L.partials.as<double>(), L.counter.as<unsigned>(), L.red.as<float>(), (const int *)nullptr);
    hipLaunchKernelGGL(k_coarse_feat<T>, dim3((E + 255) / 256), dim3(256), 0, s, r, nb, L.cnn->kinv.as<float>(), E,
                       L.red.as<float>(), L.x0.as<float>(), (const float *)nullptr, 1.0f, (float *)nullptr,
                       (const int *)nullptr);
    L.cnn->forward(L.x0.as<float>(), L.yc.as<float>(), s);
    const int64_t N = o.N;
    const unsigned G = grid_for(N, 256, 8192);
    T *z1 = L.z1.as<T>(), *w1 = L.w1.as<T>(), *t1 = L.t1.as<T>();
    hipLaunchKernelGGL(k_prolong<T>, dim3(G), dim3(256), 0, s, L.yc.as<float>(), L.red.as<float>(), L.cnn->gamma0, nb, N,
                       z1, (const int *)nullptr);
    o.apply_raw(z1, w1, nullptr);
    hipLaunchKernelGGL(k_alpha<T>, dim3(LR_BLOCKS), dim3(256), 0, s, r, z1, w1, N, L.partials.as<double>(),
                       L.counter.as<unsigned>(), L.red.as<float>(), (const int *)nullptr);
    hipLaunchKernelGGL(k_scaled_residual<T>, dim3(G), dim3(256), 0, s, r, w1, t1, z1, N, L.red.as<float>(),
                       (const int *)nullptr);
    hipLaunchKernelGGL(k_addv<T>, dim3(G), dim3(256), 0, s, x, z1, N, (const int *)nullptr);
    FEM_HIP(hipGetLastError());
}

void Learned::correct(const void *r, void *x)
{
    if (op->vec_prec == FEM_F64) learned_correct_t<double>(*this, (const double *)r, (double *)x);
    else learned_correct_t<float>(*this, (const float *)r, (float *)x);
}

void Learned::apply_raw(const void *x, void *y, const int *skip)
{
    if (op->vec_prec == FEM_F64) learned_apply_t<double>(*this, (const double *)x, (double *)y, skip);
    else learned_apply_t<float>(*this, (const float *)x, (float *)y, skip);
}

// ---- training data and training ----------------------------------------------------
void cnn_set_coefficient(CNN &c, const double *kcell, double beta)
{
    c.set_problem(kcell, beta);
}

int cnn_collect_pairs(CNN &c, Operator *op64, Precond *bj64, const Vec &b, int max_pairs, int every)
{
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
    S.snapshot = [&](const void *x, const void *r) {
        if (taken >= max_pairs) return;
        const int k = c.ds_n + taken;
        hipLaunchKernelGGL(k_gather_coarse, dim3((E + 255) / 256), dim3(256), 0, op64->ctx->stream, (const double *)x, nb, E,
                           c.ds_e.as<float>() + (size_t)k * E);
        hipLaunchKernelGGL(k_gather_coarse, dim3((E + 255) / 256), dim3(256), 0, op64->ctx->stream, (const double *)r, nb, E,
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

// host reference for the scale and the validation of the energy objective:
// y = A_c^-1 b by Jacobi-preconditioned CG on the 7-point stencil (double)
static void coarse_apply_host(const CNN &c, const std::vector<double> &x, std::vector<double> &y)
{
    const int nx = c.n[0], ny = c.n[1], nz = c.n[2];
    for (int64_t v = 0; v < c.V; ++v) {
        const int X = (int)(v % nx), Y = (int)((v / nx) % ny), Z = (int)(v / ((int64_t)nx * ny));
        const float *w = c.stencil_host.data() + v * 7;
        double a = w[0] * x[v];
        if (X > 0) a -= w[1] * x[v - 1];
        if (X < nx - 1) a -= w[2] * x[v + 1];
        if (Y > 0) a -= w[3] * x[v - nx];
        if (Y < ny - 1) a -= w[4] * x[v + nx];
        if (Z > 0) a -= w[5] * x[v - (int64_t)nx * ny];
        if (Z < nz - 1) a -= w[6] * x[v + (int64_t)nx * ny];
        y[v] = a;
    }
}

static std::vector<double> coarse_solve_host(const CNN &c, const std::vector<double> &b)
{
    const int64_t V = c.V;
    std::vector<double> x(V, 0.0), r = b, z(V), p(V), Ap(V);
    for (int64_t v = 0; v < V; ++v) z[v] = r[v] / c.stencil_host[v * 7];
    p = z;
    double rz = 0, bb = 0;
    for (int64_t v = 0; v < V; ++v) { rz += r[v] * z[v]; bb += b[v] * b[v]; }
    for (int it = 0; it < 20000; ++it) {
        coarse_apply_host(c, p, Ap);
        double pAp = 0;
        for (int64_t v = 0; v < V; ++v) pAp += p[v] * Ap[v];
        const double a = rz / pAp;
        double rr = 0;
        for (int64_t v = 0; v < V; ++v) { x[v] += a * p[v]; r[v] -= a * Ap[v]; rr += r[v] * r[v]; }
        if (rr <= 1e-20 * bb) break;
        double rz2 = 0;
        for (int64_t v = 0; v < V; ++v) { z[v] = r[v] / c.stencil_host[v * 7]; rz2 += r[v] * z[v]; }
        for (int64_t v = 0; v < V; ++v) p[v] = z[v] + (rz2 / rz) * p[v];
        rz = rz2;
    }
    return x;
}

void cnn_train(CNN &c, int epochs, int batch, float lr, float val_frac, unsigned seed, double &train_loss,
               double &val_rel)
{
    FEM_CHECK(c.ds_n >= 2, FEM_ERR_STATE, "train: collect pairs first");
    FEM_CHECK(c.kinv.p, FEM_ERR_STATE, "train: coefficient feature not set");
    hipStream_t s = c.ctx->stream;
    const int E = (int)c.V, P = c.ds_n;
    // gamma0 = median of ||e_c|| / ||r_c|| over the pairs
    DevBuf nrm;
    nrm.alloc(sizeof(double) * 2 * P);
    hipLaunchKernelGGL(k_pair_norms, dim3(P), dim3(256), 0, s, c.ds_r.as<float>(), c.ds_e.as<float>(), E, nrm.as<double>());
    std::vector<double> hn(2 * P);
    FEM_HIP(hipMemcpyAsync(hn.data(), nrm.p, sizeof(double) * 2 * P, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    std::vector<double> ratio;
    for (int p = 0; p < P; ++p)
        if (hn[2 * p] > 0) ratio.push_back(std::sqrt(hn[2 * p + 1] / hn[2 * p]));
    std::sort(ratio.begin(), ratio.end());
    c.gamma0 = ratio.empty() ? 1.0f : (float)ratio[ratio.size() / 2];
    if (c.objective == 1) {
        // scale of the Galerkin coarse solve: median ||A_c^-1 r_c|| / ||r_c|| over a few pairs
        std::vector<double> gr;
        std::vector<float> rh(E);
        for (int q = 0; q < std::min(P, 8); ++q) {
            const int p = (int)((int64_t)q * P / std::min(P, 8));
            FEM_HIP(hipMemcpy(rh.data(), c.ds_r.as<float>() + (size_t)p * E, sizeof(float) * E, hipMemcpyDeviceToHost));
            std::vector<double> b(rh.begin(), rh.end());
            double bn = 0;
            for (double v : b) bn += v * v;
            if (bn <= 0) continue;
            const std::vector<double> y = coarse_solve_host(c, b);
            double yn = 0;
            for (double v : y) yn += v * v;
            gr.push_back(std::sqrt(yn / bn));
        }
        std::sort(gr.begin(), gr.end());
        if (!gr.empty()) c.gamma0 = (float)gr[gr.size() / 2];
    }
    // deterministic shuffle and split
    std::vector<int> idx(P);
    std::iota(idx.begin(), idx.end(), 0);
    uint64_t st = 0x243F6A8885A308D3ull ^ seed;
    auto rnd = [&]() { st = st * 6364136223846793005ull + 1442695040888963407ull; return (unsigned)(st >> 33); };
    for (int i = P - 1; i > 0; --i) std::swap(idx[i], idx[rnd() % (i + 1)]);
    const int nval = std::max(1, (int)(P * val_frac));
    const int ntr = P - nval;
    FEM_CHECK(ntr >= 1, FEM_ERR_ARG, "train: no training pairs left after the validation split");
    c.feat_red.alloc(sizeof(float) * 2);
    c.tgt.alloc(sizeof(float) * E);
    DevBuf part, cnt;
    part.alloc(sizeof(double) * 2 * LR_BLOCKS);
    cnt.alloc(sizeof(unsigned));
    FEM_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned), s));
    auto prep = [&](int p) {
        const float *r = c.ds_r.as<float>() + (size_t)p * E;
        hipLaunchKernelGGL(k_coarse_rms<float>, dim3(LR_BLOCKS), dim3(256), 0, s, r, 1, c.kinv.as<float>(), E,
                           part.as<double>(), cnt.as<unsigned>(), c.feat_red.as<float>(), (const int *)nullptr);
        hipLaunchKernelGGL(k_coarse_feat<float>, dim3((E + 255) / 256), dim3(256), 0, s, r, 1, c.kinv.as<float>(), E,
                           c.feat_red.as<float>(), c.x0.as<float>(), c.ds_e.as<float>() + (size_t)p * E, c.gamma0,
                           c.tgt.as<float>(), (const int *)nullptr);
        c.forward(c.x0.as<float>(), c.y.as<float>(), s);
    };
    DevBuf acy;
    acy.alloc(sizeof(float) * E);
    // training objective: the Galerkin energy of the coarse correction (its
    // minimiser is the coarse solve A_c^-1 R r); the stored errors give the
    // scale gamma0 and the supervised validation error below
    auto energy_step = [&](int p) {
        prep(p);
        c.coarse_apply(c.y.as<float>(), acy.as<float>(), s);
        hipLaunchKernelGGL(k_energy_grad, dim3(256), dim3(256), 0, s, E, c.y.as<float>(), acy.as<float>(),
                           c.ds_r.as<float>() + (size_t)p * E, c.feat_red.as<float>(), c.gamma0, c.g4.as<float>(),
                           c.lossp.as<double>());
    };
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = ntr - 1; i > 0; --i) std::swap(idx[i], idx[rnd() % (i + 1)]);
        double tl = 0;
        for (int b0 = 0; b0 < ntr; b0 += batch) {
            const int nbat = std::min(batch, ntr - b0);
            c.zero_grad(s);
            for (int k = 0; k < nbat; ++k) {
                const int p = idx[b0 + k];
                if (c.objective == 1) {
                    energy_step(p);
                    c.backward_from_g4(s);
                } else {
                    // supervised: relative MSE per pair, mean(t^2) = ||e||^2 / (||r||^2 gamma0^2)
                    prep(p);
                    const double tt = hn[2 * p] > 0 ? hn[2 * p + 1] / (hn[2 * p] * (double)c.gamma0 * c.gamma0) : 1.0;
                    c.backward_accumulate(c.tgt.as<float>(), s, tt > 0 ? (float)(1.0 / tt) : 1.0f);
                }
                if (ep == epochs - 1) tl += c.loss_value(s);
            }
            c.adam_step(lr, 1.0f / nbat, s);
        }
        if (ep == epochs - 1) train_loss = tl / ntr;
    }
    if (c.objective == 1) {
        // validation: relative A_c-energy error of s0 gamma0 y against A_c^-1 r_c
        double vr = 0;
        std::vector<float> rh(E), yh(E), red(2);
        std::vector<double> e(E), Ae(E);
        for (int k = ntr; k < P; ++k) {
            const int p = idx[k];
            prep(p);
            FEM_HIP(hipMemcpyAsync(yh.data(), c.y.p, sizeof(float) * E, hipMemcpyDeviceToHost, s));
            FEM_HIP(hipMemcpyAsync(red.data(), c.feat_red.p, sizeof(float) * 2, hipMemcpyDeviceToHost, s));
            FEM_HIP(hipMemcpyAsync(rh.data(), c.ds_r.as<float>() + (size_t)p * E, sizeof(float) * E, hipMemcpyDeviceToHost, s));
            FEM_HIP(hipStreamSynchronize(s));
            const std::vector<double> b(rh.begin(), rh.end());
            const std::vector<double> ys = coarse_solve_host(c, b);
            for (int v = 0; v < E; ++v) e[v] = (double)yh[v] * red[0] * c.gamma0 - ys[v];
            coarse_apply_host(c, e, Ae);
            double num = 0, den = 0;
            for (int v = 0; v < E; ++v) { num += e[v] * Ae[v]; den += ys[v] * b[v]; }
            vr += den > 0 ? num / den : 0;
        }
        val_rel = vr / nval;
        return;
    }
    // validation: relative MSE = mean over pairs of ||y - t||^2 / ||t||^2
    double vr = 0;
    for (int k = ntr; k < P; ++k) {
        prep(idx[k]);
        std::vector<float> yh(E), th(E);
        FEM_HIP(hipMemcpyAsync(yh.data(), c.y.p, sizeof(float) * E, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipMemcpyAsync(th.data(), c.tgt.p, sizeof(float) * E, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        double num = 0, den = 0;
        for (int e = 0; e < E; ++e) {
            num += (double)(yh[e] - th[e]) * (yh[e] - th[e]);
            den += (double)th[e] * th[e];
        }
        vr += den > 0 ? num / den : 0;
    }
    val_rel = vr / nval;
    FEM_HIP(hipStreamSynchronize(s));
}

}  // namespace femamd
