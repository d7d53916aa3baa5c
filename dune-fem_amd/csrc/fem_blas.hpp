// BLAS-1 kernels and deterministic reductions on device DoF vectors.
#pragma once

#include "fem_common.hpp"

namespace femamd {

// wavefront (64-lane) sum followed by an LDS tree over the block's waves;
// result valid in thread 0
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds /* >= blockDim/64 */)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) s += lds[i];
    }
    return s;
}

// Publication of block partials across the XCDs without a device-scope
// fence: on this GPU a device-scope fence writes back and invalidates the
// XCD's whole L2, which costs microseconds per block.  Instead a partial is
// written with an atomic exchange (performed at the same coherence point as
// the arrival counter's atomic add), and its returned value is consumed before
// the counter is incremented, so the increment is issued only after the
// partial is in place.  The finishing block reads partials with atomic
// read-modify-writes that change nothing, from that same coherence point.
__device__ __forceinline__ void publish_partial(double *dst, double v)
{
    const unsigned long long old =
        atomicExch(reinterpret_cast<unsigned long long *>(dst), (unsigned long long)__double_as_longlong(v));
    asm volatile("" ::"v"(old));   // wait for the exchange to return
}

__device__ __forceinline__ double read_partial(double *src)
{
    return __longlong_as_double((long long)atomicAdd(reinterpret_cast<unsigned long long *>(src), 0ull));
}

__device__ __forceinline__ void reset_counter(unsigned *c)
{
    (void)atomicExch(c, 0u);
}

// Deterministic finish of a fixed-grid reduction: every block publishes its
// partials, the last block to arrive sums the partials in block order and
// writes result[slot].  The counter is reset by that block.
// Returns true in all threads of the last block (result written).
__device__ __forceinline__ bool finish_reduction(double partial, double *partials,
                                                 unsigned *counter, double *result, int nslots,
                                                 const double *own /* [nslots] in thread 0 */)
{
    __shared__ bool amLast;
    if (threadIdx.x == 0) {
        for (int s = 0; s < nslots; ++s) publish_partial(&partials[(size_t)s * gridDim.x + blockIdx.x], own[s]);
        const unsigned prev = atomicAdd(counter, 1u);
        amLast = (prev == gridDim.x - 1);
    }
    __syncthreads();
    (void)partial;
    if (!amLast) return false;
    __shared__ double red[RED_THREADS / 64];
    for (int s = 0; s < nslots; ++s) {
        double v = 0;
        // fixed order: thread t sums blocks t, t+256, ...
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            v += read_partial(&partials[(size_t)s * gridDim.x + b]);
        double tot = block_sum(v, red);
        if (threadIdx.x == 0) result[s] = tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) reset_counter(counter);
    return true;
}

// fixed grid of RED_DOT_BLOCKS x RED_THREADS threads; each thread accumulates a
// grid-stride slice in FP64 (FP32 inputs are widened before multiplying)
template <typename TX, typename TY>
__global__ void __launch_bounds__(RED_THREADS) k_dot(const TX *__restrict__ x, const TY *__restrict__ y,
                                                     int64_t n, double *partials, unsigned *counter,
                                                     double *result)
{
    __shared__ double lds[RED_THREADS / 64];
    double s = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += (double)x[i] * (double)y[i];
    double b = block_sum(s, lds);
    finish_reduction(0.0, partials, counter, result, 1, &b);
}

// host-side launchers (fem_blas.hip)
void vec_fill(Vec &v, double a);
void vec_copy(Vec &dst, const Vec &src);
void vec_axpy(Vec &y, double a, const Vec &x);
void vec_scal(Vec &x, double a);
// device result in ctx.red_result[slot]
void vec_dot_async(const Vec &x, const Vec &y, int slot);
double vec_dot(const Vec &x, const Vec &y);
// the same scalar product with its reduction scratch in a file-scope device
// array instead of the context's workspace (fem_noalloc.hip): no allocation
double vec_dot_noalloc(const Vec &x, const Vec &y);
// raw-pointer variants used by the solvers
void dot_async(Context &c, int64_t n, const double *x, const double *y, double *result, unsigned *counter, double *partials);
void dot_async(Context &c, int64_t n, const float *x, const float *y, double *result, unsigned *counter, double *partials);

}  // namespace femamd
