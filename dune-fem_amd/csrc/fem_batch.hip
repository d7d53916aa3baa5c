// Batched solves of many independent continuous-Q2 problems of different
// sizes in one set of launches.
//
// No reference counterpart: the reference solves one problem per
// FemScheme::solve call (dune/fem/schemes/femscheme.hh:195-253).  Each
// instance here follows the same CG recurrence as the single-instance solver
// (LinearSolver::cg, dune/fem/solver/linear/cg.hh:19-117, Jacobi
// preconditioned) with its own scalars, inside an FP64 iterative-refinement
// loop; instances that have converged are masked out of every launch.
//
// Layout: all instances' DoFs are concatenated (dof_off), their elements too
// (elem_off, coefficient arrays [E][9]).  Reductions are segmented: every
// reduction block belongs to one instance, the last block of an instance to
// finish sums that instance's partials in block order (deterministic).
// One inner CG step is four launches; they are captured once in a HIP graph
// and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "fem_batch.hpp"
#include "fem_blas.hpp"

namespace femamd {

constexpr int SEG_THREADS = 256;
constexpr int SEG_ITEMS = 8;   // DoFs per thread of a segment
constexpr int SEG_CHUNK = SEG_THREADS * SEG_ITEMS;

struct SegTab {
    const int *inst;       // [nblocks] instance of each block
    const int64_t *lo;     // [nblocks] first DoF
    const int64_t *hi;     // [nblocks] one past the last DoF
    const int *first;      // [count + 1] first block of each instance
    unsigned *counter;     // [count]
    double *partials;      // [3 * nblocks]
};

struct BState {            // per-instance scalars (device arrays)
    double *res, *prev, *alpha, *beta, *tol, *rr, *bn2, *rn, *scale;
    int *it, *done, *conv;
    int maxit;
};

__device__ __forceinline__ double bsum(double v, double *lds)
{
    return block_sum(v, lds);
}

// last block of an instance: sums that instance's partials in a fixed order
// (thread t takes blocks first+t, first+t+256, ...; then a block tree)
__device__ __forceinline__ bool seg_finish(const SegTab &S, int inst, int nq, const double *own, double *tot)
{
    __shared__ bool amLast;
    __shared__ double lds2[SEG_THREADS / 64];
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        for (int q = 0; q < nq; ++q) publish_partial(&S.partials[(size_t)q * nb + blockIdx.x], own[q]);
        const int nblk = S.first[inst + 1] - S.first[inst];
        amLast = atomicAdd(&S.counter[inst], 1u) == (unsigned)(nblk - 1);
    }
    __syncthreads();
    if (!amLast) return false;
    const int b0 = S.first[inst], b1 = S.first[inst + 1];
    for (int q = 0; q < nq; ++q) {
        double s = 0;
        for (int b = b0 + threadIdx.x; b < b1; b += SEG_THREADS)
            s += read_partial(&S.partials[(size_t)q * nb + b]);
        const double t = block_sum(s, lds2);
        if (threadIdx.x == 0) tot[q] = t;
        __syncthreads();
    }
    if (threadIdx.x == 0) reset_counter(&S.counter[inst]);
    return true;
}

enum { SEG_INIT = 0, SEG_QH = 1, SEG_UPD = 2, SEG_RES64 = 3 };

// MODE INIT  : p = b, q = s = dinv b; res = <p,s>, rr = <p,p>, bn2 = <b,b>
// MODE QH    : alpha = res / <q, h>
// MODE UPD   : x += a q, p -= a h, s = dinv p; res, rr, beta, it, done
// MODE RES64 : r = b - h (FP64); rn = ||r||, bn2 = ||b||^2, conv
template <typename T, int MODE>
__global__ void __launch_bounds__(SEG_THREADS)
    k_seg(SegTab S, BState B, T *__restrict__ x, T *__restrict__ p, T *__restrict__ q, const T *__restrict__ h,
          T *__restrict__ s, const T *__restrict__ b, const T *__restrict__ dinv, double eps)
{
    __shared__ double lds[SEG_THREADS / 64];
    const int inst = S.inst[blockIdx.x];
    const bool skip = (MODE == SEG_RES64) ? B.conv[inst] != 0 : B.done[inst] != 0;
    if (skip) return;   // uniform over the block
    const int64_t lo = S.lo[blockIdx.x], hi = S.hi[blockIdx.x];
    double a0 = 0, a1 = 0, a2 = 0;
    const double alpha = (MODE == SEG_UPD) ? B.alpha[inst] : 0.0;
    // a segment is at most SEG_CHUNK = SEG_ITEMS * SEG_THREADS DoFs: every
    // thread issues all of its loads before it uses them (memory-level
    // parallelism), then updates and accumulates
    constexpr int U = SEG_ITEMS;
    T va[U], vb[U], vc[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const int64_t i = lo + threadIdx.x + (int64_t)k * SEG_THREADS;
        va[k] = vb[k] = vc[k] = T(0);
        if (i < hi) {
            if (MODE == SEG_INIT) {
                va[k] = b[i];
                vb[k] = dinv[i];
            } else if (MODE == SEG_QH) {
                va[k] = q[i];
                vb[k] = h[i];
            } else if (MODE == SEG_UPD) {
                va[k] = p[i];
                vb[k] = h[i];
                vc[k] = dinv[i];
            } else {
                va[k] = b[i];
                vb[k] = h[i];
            }
        }
    }
    T vx[U], vq[U];
    if (MODE == SEG_UPD) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int64_t i = lo + threadIdx.x + (int64_t)k * SEG_THREADS;
            vx[k] = vq[k] = T(0);
            if (i < hi) {
                vx[k] = x[i];
                vq[k] = q[i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const int64_t i = lo + threadIdx.x + (int64_t)k * SEG_THREADS;
        if (i >= hi) continue;
        if (MODE == SEG_INIT) {
            const T bv = va[k];
            const T sv = vb[k] * bv;
            p[i] = bv;
            q[i] = sv;
            s[i] = sv;
            a0 += (double)bv * (double)sv;
            a1 += (double)bv * (double)bv;
        } else if (MODE == SEG_QH) {
            a0 += (double)va[k] * (double)vb[k];
        } else if (MODE == SEG_UPD) {
            x[i] = (T)((double)vx[k] + alpha * (double)vq[k]);
            const T pn = (T)((double)va[k] - alpha * (double)vb[k]);
            p[i] = pn;
            const T sv = vc[k] * pn;
            s[i] = sv;
            a0 += (double)pn * (double)sv;
            a1 += (double)pn * (double)pn;
        } else {
            const double r = (double)va[k] - (double)vb[k];
            p[i] = (T)r;
            a0 += r * r;
            a2 += (double)va[k] * (double)va[k];
        }
    }
    double own[3];
    own[0] = bsum(a0, lds);
    __syncthreads();
    own[1] = bsum(a1, lds);
    __syncthreads();
    own[2] = bsum(a2, lds);
    double tot[3];
    if (!seg_finish(S, inst, 3, own, tot) || threadIdx.x != 0) return;
    if (MODE == SEG_INIT) {
        B.res[inst] = tot[0];
        B.rr[inst] = tot[1];
        B.tol[inst] = eps * eps * tot[1];   // relative reduction of the true residual
        B.it[inst] = 0;
        B.done[inst] = !(tot[1] > B.tol[inst]) || B.maxit <= 0;
    } else if (MODE == SEG_QH) {
        B.alpha[inst] = B.res[inst] / tot[0];
    } else if (MODE == SEG_UPD) {
        const double r = tot[0];
        B.prev[inst] = B.res[inst];
        B.beta[inst] = r / B.res[inst];
        B.res[inst] = r;
        B.rr[inst] = tot[1];
        B.it[inst] += 1;
        B.done[inst] = !(tot[1] > B.tol[inst]) || B.it[inst] >= B.maxit;
    } else {
        const double rn = sqrt(tot[0]);
        B.rn[inst] = rn;
        B.bn2[inst] = tot[2];
        B.conv[inst] = rn <= eps * sqrt(tot[2]);
        B.scale[inst] = rn > 0 ? 1.0 / rn : 1.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(SEG_THREADS)
    k_seg_qupd(SegTab S, BState B, T *__restrict__ q, const T *__restrict__ s)
{
    const int inst = S.inst[blockIdx.x];
    if (B.done[inst] || B.it[inst] == 0) return;
    const double beta = B.beta[inst];
    const int64_t lo = S.lo[blockIdx.x], hi = S.hi[blockIdx.x];
    T vq[SEG_ITEMS], vs[SEG_ITEMS];
#pragma unroll
    for (int k = 0; k < SEG_ITEMS; ++k) {
        const int64_t i = lo + threadIdx.x + (int64_t)k * SEG_THREADS;
        vq[k] = vs[k] = T(0);
        if (i < hi) {
            vq[k] = q[i];
            vs[k] = s[i];
        }
    }
#pragma unroll
    for (int k = 0; k < SEG_ITEMS; ++k) {
        const int64_t i = lo + threadIdx.x + (int64_t)k * SEG_THREADS;
        if (i < hi) q[i] = (T)(beta * (double)vq[k] + (double)vs[k]);
    }
}

// r32 = r64 * scale_i, d = 0, inner done = outer converged
__global__ void __launch_bounds__(SEG_THREADS)
    k_seg_to32(SegTab S, BState B, const double *__restrict__ r, float *__restrict__ r32, float *__restrict__ d)
{
    const int inst = S.inst[blockIdx.x];
    const double a = B.scale[inst];
    for (int64_t i = S.lo[blockIdx.x] + threadIdx.x; i < S.hi[blockIdx.x]; i += SEG_THREADS) {
        r32[i] = (float)(r[i] * a);
        d[i] = 0.f;
    }
    if (threadIdx.x == 0 && S.first[inst] == (int)blockIdx.x) B.done[inst] = B.conv[inst];
}

__global__ void __launch_bounds__(SEG_THREADS)
    k_seg_upd64(SegTab S, BState B, double *__restrict__ x, const float *__restrict__ d)
{
    const int inst = S.inst[blockIdx.x];
    if (B.conv[inst]) return;
    const double a = 1.0 / B.scale[inst];
    for (int64_t i = S.lo[blockIdx.x] + threadIdx.x; i < S.hi[blockIdx.x]; i += SEG_THREADS)
        x[i] += a * (double)d[i];
}

__global__ void k_count_active(const int *done, int n, int *out)
{
    __shared__ int c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    int mine = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mine += done[i] ? 0 : 1;
    atomicAdd(&c, mine);
    __syncthreads();
    if (threadIdx.x == 0) *out = c;
}

template <typename T>
__global__ void k_jacobi_inv_b(const double *__restrict__ d, T *__restrict__ dinv, int64_t n)
{
    const double eps = 16.0 * 2.220446049250313e-16;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dinv[i] = (T)(fabs(d[i]) < eps ? 1.0 : 1.0 / d[i]);
}

// ---------------------------------------------------------------------------------
void Batch::init(Context *c, int n, const int *nx_, const int *ny_, double c_)
{
    ctx = c;
    count = n;
    creac = c_;
    FEM_CHECK(n >= 1, FEM_ERR_ARG, "batch: at least one instance");
    nx.assign(nx_, nx_ + n);
    ny.assign(ny_, ny_ + n);
    dof_off.assign(n + 1, 0);
    elem_off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        FEM_CHECK(nx[i] > 0 && ny[i] > 0, FEM_ERR_ARG, "batch: mesh size");
        dof_off[i + 1] = dof_off[i] + (int64_t)(2 * nx[i] + 1) * (2 * ny[i] + 1);
        elem_off[i + 1] = elem_off[i] + (int64_t)nx[i] * ny[i];
    }
    N = dof_off[n];
    E = elem_off[n];
    for (int sh = 0; sh < CG_NSHAPES; ++sh) {
        const int TX = CG_SHAPES[sh][0], TY = CG_SHAPES[sh][1];
        std::vector<CGInst> v(n);
        int off = 0;
        for (int i = 0; i < n; ++i) {
            CGInst &I = v[i];
            I.nx = nx[i];
            I.ny = ny[i];
            I.dof_off = dof_off[i];
            I.elem_off = elem_off[i];
            I.tiles_x = (nx[i] + TX - 1) / TX;
            I.ntiles = I.tiles_x * ((ny[i] + TY - 1) / TY);
            I.tile_off = off;
            I.c = creac;
            I.kappa_const = 1.0;
            off += I.ntiles;
        }
        tiles[sh] = off;
        inst_dev[sh].alloc(sizeof(CGInst) * n);
        FEM_HIP(hipMemcpy(inst_dev[sh].p, v.data(), sizeof(CGInst) * n, hipMemcpyHostToDevice));
    }
    CGTab t = make_cg_tab();
    tab.alloc(sizeof(t));
    FEM_HIP(hipMemcpy(tab.p, &t, sizeof(t), hipMemcpyHostToDevice));
    // segment table
    std::vector<int> sinst, first(n + 1, 0);
    std::vector<int64_t> slo, shi;
    for (int i = 0; i < n; ++i) {
        first[i] = (int)sinst.size();
        for (int64_t a = dof_off[i]; a < dof_off[i + 1]; a += SEG_CHUNK) {
            sinst.push_back(i);
            slo.push_back(a);
            shi.push_back(std::min<int64_t>(a + SEG_CHUNK, dof_off[i + 1]));
        }
    }
    first[n] = (int)sinst.size();
    nseg = (int)sinst.size();
    seg_inst.alloc(sizeof(int) * nseg);
    seg_lo.alloc(sizeof(int64_t) * nseg);
    seg_hi.alloc(sizeof(int64_t) * nseg);
    seg_first.alloc(sizeof(int) * (n + 1));
    seg_counter.alloc(sizeof(unsigned) * n);
    seg_partials.alloc(sizeof(double) * 3 * nseg);
    FEM_HIP(hipMemcpy(seg_inst.p, sinst.data(), sizeof(int) * nseg, hipMemcpyHostToDevice));
    FEM_HIP(hipMemcpy(seg_lo.p, slo.data(), sizeof(int64_t) * nseg, hipMemcpyHostToDevice));
    FEM_HIP(hipMemcpy(seg_hi.p, shi.data(), sizeof(int64_t) * nseg, hipMemcpyHostToDevice));
    FEM_HIP(hipMemcpy(seg_first.p, first.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    FEM_HIP(hipMemset(seg_counter.p, 0, sizeof(unsigned) * n));
    // per-instance state
    sdbl.alloc(sizeof(double) * 9 * n);
    sint.alloc(sizeof(int) * 3 * n);
    FEM_HIP(hipMemset(sdbl.p, 0, sdbl.bytes));
    FEM_HIP(hipMemset(sint.p, 0, sint.bytes));
    active.alloc(sizeof(int));
    FEM_HIP(hipHostMalloc((void **)&hactive, sizeof(int), hipHostMallocDefault));
}

void Batch::drop_graph()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

Batch::~Batch()
{
    drop_graph();
    if (hactive) (void)hipHostFree(hactive);
}

void Batch::set_kappa_qp(const double *kv)
{
    FEM_CHECK(kv, FEM_ERR_ARG, "null kappa");
    // the coefficient and Jacobi buffers keep their addresses after the first
    // call (the captured graph reads kvol32 and dinv32); a first allocation
    // drops any graph
    if (!kvol64.p) {
        drop_graph();
        kvol64.alloc(sizeof(double) * E * 9);
        kvol32.alloc(sizeof(float) * E * 9);
        dinv64.alloc(sizeof(double) * N);
        dinv32.alloc(sizeof(float) * N);
    }
    FEM_HIP(hipMemcpy(kvol64.p, kv, sizeof(double) * E * 9, hipMemcpyHostToDevice));
    std::vector<float> f(E * 9);
    for (int64_t i = 0; i < E * 9; ++i) f[i] = (float)kv[i];
    FEM_HIP(hipMemcpy(kvol32.p, f.data(), sizeof(float) * E * 9, hipMemcpyHostToDevice));
    // Jacobi diagonal (FP64 and FP32 inverses)
    DevBuf d;
    d.alloc(sizeof(double) * N);
    cg_launch_shape<double>(0, CG_MODE_DIAG, inst_dev[0].as<CGInst>(), count, tiles[0], tab.as<CGTab>(),
                            kvol64.as<double>(), nullptr, d.as<double>(), nullptr, nullptr, ctx->stream);
    hipLaunchKernelGGL(k_jacobi_inv_b<double>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, ctx->stream, d.as<double>(), dinv64.as<double>(), N);
    hipLaunchKernelGGL(k_jacobi_inv_b<float>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, ctx->stream, d.as<double>(), dinv32.as<float>(), N);
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(ctx->stream));
}

void Batch::apply_raw(int prec, const void *u, void *w, const int *inst_done) const
{
    FEM_CHECK(kvol64.p, FEM_ERR_STATE, "batch: coefficient data not set");
    if (prec == FEM_F64)
        cg_launch_shape<double>(shape, CG_MODE_APPLY, inst_dev[shape].as<CGInst>(), count, tiles[shape],
                                tab.as<CGTab>(), kvol64.as<double>(), (const double *)u, (double *)w, nullptr,
                                inst_done, ctx->stream);
    else
        cg_launch_shape<float>(shape, CG_MODE_APPLY, inst_dev[shape].as<CGInst>(), count, tiles[shape],
                               tab.as<CGTab>(), kvol32.as<float>(), (const float *)u, (float *)w, nullptr,
                               inst_done, ctx->stream);
}

void Batch::rhs(const double *fq, Vec &F) const
{
    FEM_CHECK(F.n == N && F.prec == FEM_F64, FEM_ERR_ARG, "batch rhs: FP64 vector of the batch size");
    DevBuf f;
    f.alloc(sizeof(double) * E * 9);
    FEM_HIP(hipMemcpy(f.p, fq, sizeof(double) * E * 9, hipMemcpyHostToDevice));
    cg_launch_shape<double>(0, CG_MODE_RHS, inst_dev[0].as<CGInst>(), count, tiles[0], tab.as<CGTab>(),
                            f.as<double>(), nullptr, (double *)F.data, nullptr, nullptr, ctx->stream);
    FEM_HIP(hipStreamSynchronize(ctx->stream));
}

int Batch::autotune(int reps)
{
    // time the FP32 apply of the whole batch for every tile shape; keep the fastest
    DevBuf u, w;
    u.alloc(sizeof(float) * N);
    w.alloc(sizeof(float) * N);
    FEM_HIP(hipMemset(u.p, 0, u.bytes));
    float best = 1e30f;
    int bestsh = 0;
    for (int sh = 0; sh < CG_NSHAPES; ++sh) {
        shape = sh;
        apply_raw(FEM_F32, u.p, w.p, nullptr);
        FEM_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        for (int r = 0; r < reps; ++r) apply_raw(FEM_F32, u.p, w.p, nullptr);
        FEM_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        FEM_HIP(hipEventSynchronize(ctx->ev1));
        float ms;
        FEM_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        tune_ms[sh] = ms / reps;
        if (ms < best) { best = ms; bestsh = sh; }
    }
    shape = bestsh;
    drop_graph();
    return bestsh;
}

void Batch::solve(const Vec &b, Vec &x, double tol, double inner_tol, int inner_maxit, int max_outer, int use_graph,
                  fem_batch_info &info)
{
    FEM_CHECK(b.n == N && x.n == N && b.prec == FEM_F64 && x.prec == FEM_F64, FEM_ERR_ARG,
              "batch solve: FP64 vectors of the batch size");
    FEM_CHECK(dinv32.p, FEM_ERR_STATE, "batch: coefficient data not set");
    hipStream_t st = ctx->stream;
    std::memset(&info, 0, sizeof(info));
    auto t0 = std::chrono::steady_clock::now();
    // work vectors: allocated on the first solve only (N is fixed for a
    // batch), so the pointers captured in the HIP graph stay valid
    if (!h64.p) {
        drop_graph();
        h64.alloc(sizeof(double) * N);
        r64.alloc(sizeof(double) * N);
        for (DevBuf *v : { &r32, &d32, &p32, &q32, &s32, &h32 }) v->alloc(sizeof(float) * N);
    }
    const int n = count;
    double *D = sdbl.as<double>();
    int *I = sint.as<int>();
    BState B{ D, D + n, D + 2 * n, D + 3 * n, D + 4 * n, D + 5 * n, D + 6 * n, D + 7 * n, D + 8 * n,
              I, I + n, I + 2 * n, inner_maxit };
    SegTab S{ seg_inst.as<int>(), seg_lo.as<int64_t>(), seg_hi.as<int64_t>(), seg_first.as<int>(),
              seg_counter.as<unsigned>(), seg_partials.as<double>() };
    float *r32p = r32.as<float>(), *d = d32.as<float>(), *p = p32.as<float>(), *q = q32.as<float>(),
          *s = s32.as<float>(), *h = h32.as<float>();
    const float *dv = dinv32.as<float>();
    auto step = [&](hipStream_t ss) {
        hipLaunchKernelGGL(k_seg_qupd<float>, dim3(nseg), dim3(SEG_THREADS), 0, ss, S, B, q, s);
        if (shape >= 0)
            cg_launch_shape<float>(shape, CG_MODE_APPLY, inst_dev[shape].as<CGInst>(), count, tiles[shape],
                                   tab.as<CGTab>(), kvol32.as<float>(), q, h, nullptr, B.done, ss);
        hipLaunchKernelGGL((k_seg<float, SEG_QH>), dim3(nseg), dim3(SEG_THREADS), 0, ss, S, B, d, p, q, h, s, r32p, dv, 0.0);
        hipLaunchKernelGGL((k_seg<float, SEG_UPD>), dim3(nseg), dim3(SEG_THREADS), 0, ss, S, B, d, p, q, h, s, r32p, dv, 0.0);
    };
    if (use_graph && !graph_exec) {
        FEM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        step(st);
        FEM_HIP(hipStreamEndCapture(st, &graph));
        FEM_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    }
    FEM_HIP(hipMemsetAsync(B.conv, 0, sizeof(int) * n, st));   // not captured
    int outer = 0, total_inner = 0;
    std::vector<int> its(n, 0);
    while (true) {
        // FP64 true residual per instance
        apply_raw(FEM_F64, x.data, h64.p, nullptr);
        hipLaunchKernelGGL((k_seg<double, SEG_RES64>), dim3(nseg), dim3(SEG_THREADS), 0, st, S, B, (double *)nullptr,
                           r64.as<double>(), (double *)nullptr, h64.as<double>(), (double *)nullptr,
                           (const double *)b.data, (const double *)nullptr, tol);
        FEM_HIP(hipGetLastError());
        // clear conv flags only on the first pass (k_seg RES64 skips converged instances)
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, st, B.conv, n, active.as<int>());
        FEM_HIP(hipMemcpyAsync(hactive, active.p, sizeof(int), hipMemcpyDeviceToHost, st));
        FEM_HIP(hipStreamSynchronize(st));
        if (*hactive == 0 || outer >= max_outer) break;
        hipLaunchKernelGGL(k_seg_to32, dim3(nseg), dim3(SEG_THREADS), 0, st, S, B, r64.as<double>(), r32p, d);
        hipLaunchKernelGGL((k_seg<float, SEG_INIT>), dim3(nseg), dim3(SEG_THREADS), 0, st, S, B, d, p, q, h, s, r32p, dv, inner_tol);
        FEM_HIP(hipGetLastError());
        int chunk = 8, done_steps = 0;
        while (true) {
            for (int c = 0; c < chunk; ++c) {
                if (use_graph) FEM_HIP(hipGraphLaunch(graph_exec, st));
                else step(st);
            }
            done_steps += chunk;
            hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, st, B.done, n, active.as<int>());
            FEM_HIP(hipMemcpyAsync(hactive, active.p, sizeof(int), hipMemcpyDeviceToHost, st));
            FEM_HIP(hipStreamSynchronize(st));
            if (*hactive == 0 || done_steps >= inner_maxit) break;
            if (chunk < 64) chunk *= 2;
        }
        std::vector<int> it(n);
        FEM_HIP(hipMemcpy(it.data(), B.it, sizeof(int) * n, hipMemcpyDeviceToHost));
        int mx = 0;
        for (int i = 0; i < n; ++i) { its[i] += it[i]; mx = std::max(mx, it[i]); total_inner += it[i]; }
        info.max_inner_iterations += mx;
        hipLaunchKernelGGL(k_seg_upd64, dim3(nseg), dim3(SEG_THREADS), 0, st, S, B, (double *)x.data, d);
        FEM_HIP(hipGetLastError());
        ++outer;
    }
    FEM_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    std::vector<int> conv(n);
    std::vector<double> rn(n), bn2(n);
    FEM_HIP(hipMemcpy(conv.data(), B.conv, sizeof(int) * n, hipMemcpyDeviceToHost));
    FEM_HIP(hipMemcpy(rn.data(), B.rn, sizeof(double) * n, hipMemcpyDeviceToHost));
    FEM_HIP(hipMemcpy(bn2.data(), B.bn2, sizeof(double) * n, hipMemcpyDeviceToHost));
    info.count = n;
    info.converged = 0;
    info.max_true_rel_residual = 0;
    last_iters = its;
    last_relres.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        info.converged += conv[i] ? 1 : 0;
        const double rel = bn2[i] > 0 ? rn[i] / std::sqrt(bn2[i]) : rn[i];
        last_relres[i] = rel;
        info.max_true_rel_residual = std::max(info.max_true_rel_residual, rel);
    }
    info.outer_iterations = outer;
    info.total_inner_iterations = total_inner;
    info.time_s = std::chrono::duration<double>(t1 - t0).count();
    info.tile_x = CG_SHAPES[shape][0];
    info.tile_y = CG_SHAPES[shape][1];
}

}  // namespace femamd
