// BLAS-1 kernels on device DoF vectors with deterministic fixed-grid
// reductions (scalarProductDofs / normSquaredDofs of
// dune/fem/function/common/scalarproducts.hh:101-127, without the MPI sum).
#include "fem_blas.hpp"

namespace femamd {

template <typename T>
__global__ void k_fill(T *__restrict__ x, int64_t n, T a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = a;
}

template <typename TD, typename TS>
__global__ void k_copy(TD *__restrict__ d, const TS *__restrict__ s, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (TD)s[i];
}

template <typename TY, typename TX>
__global__ void k_axpy(TY *__restrict__ y, const TX *__restrict__ x, int64_t n, double a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (TY)((double)y[i] + a * (double)x[i]);
}

template <typename T>
__global__ void k_scal(T *__restrict__ x, int64_t n, double a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)((double)x[i] * a);
}

static void launch_geom(int64_t n, unsigned &g, unsigned &t)
{
    t = 256;
    g = grid_for(n, 256, 8192);
}

void vec_fill(Vec &v, double a)
{
    unsigned g, t;
    launch_geom(v.n, g, t);
    if (v.prec == FEM_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(t), 0, v.ctx->stream, (double *)v.data, v.n, a);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(t), 0, v.ctx->stream, (float *)v.data, v.n, (float)a);
    FEM_HIP(hipGetLastError());
}

void vec_copy(Vec &d, const Vec &s)
{
    FEM_CHECK(d.n == s.n, FEM_ERR_ARG, "copy: size mismatch");
    unsigned g, t;
    launch_geom(d.n, g, t);
    hipStream_t st = d.ctx->stream;
    if (d.prec == s.prec) {
        FEM_HIP(hipMemcpyAsync(d.data, s.data, d.bytes(), hipMemcpyDeviceToDevice, st));
        return;
    }
    if (d.prec == FEM_F64)
        hipLaunchKernelGGL((k_copy<double, float>), dim3(g), dim3(t), 0, st, (double *)d.data, (const float *)s.data, d.n);
    else
        hipLaunchKernelGGL((k_copy<float, double>), dim3(g), dim3(t), 0, st, (float *)d.data, (const double *)s.data, d.n);
    FEM_HIP(hipGetLastError());
}

void vec_axpy(Vec &y, double a, const Vec &x)
{
    FEM_CHECK(y.n == x.n, FEM_ERR_ARG, "axpy: size mismatch");
    unsigned g, t;
    launch_geom(y.n, g, t);
    hipStream_t st = y.ctx->stream;
    if (y.prec == FEM_F64 && x.prec == FEM_F64)
        hipLaunchKernelGGL((k_axpy<double, double>), dim3(g), dim3(t), 0, st, (double *)y.data, (const double *)x.data, y.n, a);
    else if (y.prec == FEM_F64)
        hipLaunchKernelGGL((k_axpy<double, float>), dim3(g), dim3(t), 0, st, (double *)y.data, (const float *)x.data, y.n, a);
    else if (x.prec == FEM_F64)
        hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(t), 0, st, (float *)y.data, (const double *)x.data, y.n, a);
    else
        hipLaunchKernelGGL((k_axpy<float, float>), dim3(g), dim3(t), 0, st, (float *)y.data, (const float *)x.data, y.n, a);
    FEM_HIP(hipGetLastError());
}

void vec_scal(Vec &x, double a)
{
    unsigned g, t;
    launch_geom(x.n, g, t);
    if (x.prec == FEM_F64)
        hipLaunchKernelGGL(k_scal<double>, dim3(g), dim3(t), 0, x.ctx->stream, (double *)x.data, x.n, a);
    else
        hipLaunchKernelGGL(k_scal<float>, dim3(g), dim3(t), 0, x.ctx->stream, (float *)x.data, x.n, a);
    FEM_HIP(hipGetLastError());
}

template <typename TX, typename TY>
static void dot_launch(Context &c, int64_t n, const TX *x, const TY *y, double *result, unsigned *counter,
                       double *partials)
{
    hipLaunchKernelGGL((k_dot<TX, TY>), dim3(RED_DOT_BLOCKS), dim3(RED_THREADS), 0, c.stream, x, y, n, partials,
                       counter, result);
    FEM_HIP(hipGetLastError());
}

void dot_async(Context &c, int64_t n, const double *x, const double *y, double *result, unsigned *counter,
               double *partials)
{
    dot_launch(c, n, x, y, result, counter, partials);
}

void dot_async(Context &c, int64_t n, const float *x, const float *y, double *result, unsigned *counter,
               double *partials)
{
    dot_launch(c, n, x, y, result, counter, partials);
}

void vec_dot_async(const Vec &x, const Vec &y, int slot)
{
    FEM_CHECK(x.n == y.n, FEM_ERR_ARG, "dot: size mismatch");
    Context &c = *x.ctx;
    double *res = c.red_result + slot;
    unsigned *cnt = c.red_counter + slot;
    double *part = c.red_partial + (size_t)slot * RED_DOT_BLOCKS;
    if (x.prec == FEM_F64 && y.prec == FEM_F64) dot_launch(c, x.n, (const double *)x.data, (const double *)y.data, res, cnt, part);
    else if (x.prec == FEM_F64) dot_launch(c, x.n, (const double *)x.data, (const float *)y.data, res, cnt, part);
    else if (y.prec == FEM_F64) dot_launch(c, x.n, (const float *)x.data, (const double *)y.data, res, cnt, part);
    else dot_launch(c, x.n, (const float *)x.data, (const float *)y.data, res, cnt, part);
}

double vec_dot(const Vec &x, const Vec &y)
{
    vec_dot_async(x, y, 0);
    Context &c = *x.ctx;
    FEM_HIP(hipMemcpyAsync(c.host_result, c.red_result, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FEM_HIP(hipStreamSynchronize(c.stream));
    return c.host_result[0];
}

}  // namespace femamd
