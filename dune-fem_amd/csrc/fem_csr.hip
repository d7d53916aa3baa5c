// Assembled Jacobian of a Galerkin operator as a device CSR matrix: the
// counterpart of DifferentiableGalerkinOperator::jacobian / assemble
// (dune/fem/schemes/galerkin.hh:1195-1314, 1548-1643) into SparseRowMatrix
// (dune/fem/operator/matrix/spmatrix.hh:39-860), with its apply (:148-178)
// and extractDiagonal (:846).
//
// The operators here are linear, so the Jacobian is the operator itself.  It
// is assembled from the production apply kernel by coloured probing instead
// of a second (element-matrix) code path: the apply on the probe "unit
// vector j on every node/element of colour c" gives, in every row, the entry
// of the one column of colour c that row couples with.
//   * SIPG DG: an element couples with itself and its six face neighbours.
//     With colour (x mod 3, y mod 3, z mod 3), the seven elements of a row
//     block have distinct colours: 27 nb applies.  Rows of element e hold
//     the blocks of its existing neighbours in increasing element order.
//   * Continuous Q2: node (I, J) couples with the nodes of the elements
//     around it: columns I-2..I+2 (I a vertex index) or I-1..I+1 (an edge or
//     cell index) per direction, clipped to the lattice.  Colour
//     (I mod 5, J mod 5): 25 applies.  Dirichlet rows are unit rows with
//     zero entries elsewhere in the pattern, as the reference's constraints
//     leave them (dune/fem/schemes/dirichletconstraints.hh:272-338).
// Column indices are 32-bit, row offsets 64-bit.
#include <algorithm>

#include "fem_blas.hpp"
#include "fem_csr.hpp"

namespace femamd {

namespace {

constexpr int CT = 256;

inline unsigned cgrid(int64_t n) { return grid_for(n, CT, 65536); }

// ---- DG pattern -------------------------------------------------------------------
struct DGDims {
    int n[3];
    int nb;
};

// the existing neighbours of element e in increasing element order (the row
// block order), returns the count (1..7)
__device__ __forceinline__ int dg_row_elems(const DGDims &d, int64_t e, int64_t *out)
{
    const int64_t sx = 1, sy = d.n[0], sz = (int64_t)d.n[0] * d.n[1];
    const int x = (int)(e % d.n[0]), y = (int)((e / d.n[0]) % d.n[1]), z = (int)(e / sz);
    int m = 0;
    if (z > 0) out[m++] = e - sz;
    if (y > 0) out[m++] = e - sy;
    if (x > 0) out[m++] = e - sx;
    out[m++] = e;
    if (x < d.n[0] - 1) out[m++] = e + sx;
    if (y < d.n[1] - 1) out[m++] = e + sy;
    if (z < d.n[2] - 1) out[m++] = e + sz;
    return m;
}

__device__ __forceinline__ int dg_color(const DGDims &d, int64_t e)
{
    const int x = (int)(e % d.n[0]), y = (int)((e / d.n[0]) % d.n[1]), z = (int)(e / ((int64_t)d.n[0] * d.n[1]));
    return (x % 3) + 3 * ((y % 3) + 3 * (z % 3));
}

// column indices of every row (the pattern does not depend on the probe)
__global__ void k_csr_dg_cols(DGDims d, const int64_t *__restrict__ rowptr, int *__restrict__ col)
{
    const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2] * d.nb;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = r / d.nb;
        int64_t el[7];
        const int m = dg_row_elems(d, e, el);
        int64_t k = rowptr[r];
        for (int s = 0; s < m; ++s)
            for (int j = 0; j < d.nb; ++j) col[k++] = (int)(el[s] * d.nb + j);
    }
}

template <typename T>
__global__ void k_csr_dg_probe(DGDims d, T *__restrict__ p, int j, int c)
{
    const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2] * d.nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / d.nb;
        p[i] = ((int)(i - e * d.nb) == j && dg_color(d, e) == c) ? T(1) : T(0);
    }
}

template <typename T>
__global__ void k_csr_dg_extract(DGDims d, const T *__restrict__ y, const int64_t *__restrict__ rowptr,
                                 double *__restrict__ val, int j, int c)
{
    const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2] * d.nb;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = r / d.nb;
        int64_t el[7];
        const int m = dg_row_elems(d, e, el);
        for (int s = 0; s < m; ++s)
            if (dg_color(d, el[s]) == c) {
                val[rowptr[r] + (int64_t)s * d.nb + j] = (double)y[r];
                break;
            }
    }
}

// ---- Q2 pattern ---------------------------------------------------------------------
struct Q2Dims {
    int nx, ny;   // cells; the node lattice is (2 nx + 1) x (2 ny + 1)
};

// column range of lattice index I along a direction with `cells` cells
__device__ __forceinline__ void q2_range(int I, int cells, int &lo, int &hi)
{
    const int L = 2 * cells;
    if (I % 2 == 0) {
        lo = I - 2;
        hi = I + 2;
    } else {
        lo = I - 1;
        hi = I + 1;
    }
    lo = lo < 0 ? 0 : lo;
    hi = hi > L ? L : hi;
}

__global__ void k_csr_q2_cols(Q2Dims d, const int64_t *__restrict__ rowptr, int *__restrict__ col)
{
    const int W = 2 * d.nx + 1;
    const int64_t N = (int64_t)W * (2 * d.ny + 1);
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(r % W), J = (int)(r / W);
        int x0, x1, y0, y1;
        q2_range(I, d.nx, x0, x1);
        q2_range(J, d.ny, y0, y1);
        int64_t k = rowptr[r];
        for (int jj = y0; jj <= y1; ++jj)
            for (int ii = x0; ii <= x1; ++ii) col[k++] = jj * W + ii;
    }
}

template <typename T>
__global__ void k_csr_q2_probe(Q2Dims d, T *__restrict__ p, int c)
{
    const int W = 2 * d.nx + 1;
    const int64_t N = (int64_t)W * (2 * d.ny + 1);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(i % W), J = (int)(i / W);
        p[i] = ((I % 5) + 5 * (J % 5)) == c ? T(1) : T(0);
    }
}

template <typename T>
__global__ void k_csr_q2_extract(Q2Dims d, const T *__restrict__ y, const int64_t *__restrict__ rowptr,
                                 double *__restrict__ val, int c)
{
    const int W = 2 * d.nx + 1;
    const int64_t N = (int64_t)W * (2 * d.ny + 1);
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(r % W), J = (int)(r / W);
        int x0, x1, y0, y1;
        q2_range(I, d.nx, x0, x1);
        q2_range(J, d.ny, y0, y1);
        // the one column of colour c in the window
        const int cx = c % 5, cy = c / 5;
        const int ii = x0 + ((cx - x0 % 5) + 5) % 5, jj = y0 + ((cy - y0 % 5) + 5) % 5;
        if (ii > x1 || jj > y1) continue;
        val[rowptr[r] + (int64_t)(jj - y0) * (x1 - x0 + 1) + (ii - x0)] = (double)y[r];
    }
}

// ---- SpMV: one wavefront per row ------------------------------------------------------
template <typename TV, typename TX>
__global__ void __launch_bounds__(256) k_csr_spmv(int64_t nrows, const int64_t *__restrict__ rowptr,
                                                  const int *__restrict__ col, const TV *__restrict__ val,
                                                  const TX *__restrict__ x, TX *__restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r = wave; r < nrows; r += nw) {
        double acc = 0;
        const int64_t b = rowptr[r], e = rowptr[r + 1];
        for (int64_t k = b + lane; k < e; k += 64) acc += (double)val[k] * (double)x[col[k]];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) y[r] = (TX)acc;
    }
}

__global__ void k_csr_diag(int64_t nrows, const int64_t *__restrict__ rowptr, const int *__restrict__ col,
                           const double *__restrict__ val, double *__restrict__ d)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
        double v = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
            if (col[k] == r) {
                v = val[k];
                break;
            }
        d[r] = v;
    }
}

}  // namespace

void CSR::assemble(Operator *o)
{
    op = o;
    ctx = o->ctx;
    hipStream_t s = ctx->stream;
    N = o->N;
    FEM_CHECK(N < ((int64_t)1 << 31), FEM_ERR_UNSUPPORTED, "CSR: more than 2^31 rows");
    // row lengths on the host, then offsets
    std::vector<int64_t> rp(N + 1, 0);
    if (o->is_dg()) {
        const int nb = o->k * o->k * o->k;
        const int *n = o->desc.n;
        int64_t acc = 0;
        for (int64_t e = 0; e < o->E; ++e) {
            const int x = (int)(e % n[0]), y = (int)((e / n[0]) % n[1]), z = (int)(e / ((int64_t)n[0] * n[1]));
            const int m = 1 + (x > 0) + (x < n[0] - 1) + (y > 0) + (y < n[1] - 1) + (z > 0) + (z < n[2] - 1);
            for (int i = 0; i < nb; ++i) {
                rp[e * nb + i] = acc;
                acc += (int64_t)m * nb;
            }
        }
        rp[N] = acc;
    } else {
        const int nx = o->desc.n[0], ny = o->desc.n[1], W = 2 * nx + 1;
        auto len = [](int I, int cells) {
            int lo = (I % 2 == 0) ? I - 2 : I - 1, hi = (I % 2 == 0) ? I + 2 : I + 1;
            lo = std::max(lo, 0);
            hi = std::min(hi, 2 * cells);
            return hi - lo + 1;
        };
        int64_t acc = 0;
        for (int64_t r = 0; r < N; ++r) {
            rp[r] = acc;
            acc += (int64_t)len((int)(r % W), nx) * len((int)(r / W), ny);
        }
        rp[N] = acc;
    }
    nnz = rp[N];
    rowptr.alloc(sizeof(int64_t) * (N + 1));
    col.alloc(sizeof(int) * nnz);
    val.alloc(sizeof(double) * nnz);
    FEM_HIP(hipMemcpyAsync(rowptr.p, rp.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice, s));
    FEM_HIP(hipMemsetAsync(val.p, 0, val.bytes, s));
    DevBuf pb, yb;
    const size_t es = o->vec_prec == FEM_F64 ? 8 : 4;
    pb.alloc(N * es);
    yb.alloc(N * es);
    const int64_t *R = rowptr.as<int64_t>();
    if (o->is_dg()) {
        DGDims d{ { o->desc.n[0], o->desc.n[1], o->desc.n[2] }, o->k * o->k * o->k };
        hipLaunchKernelGGL(k_csr_dg_cols, dim3(cgrid(N)), dim3(CT), 0, s, d, R, col.as<int>());
        for (int c = 0; c < 27; ++c)
            for (int j = 0; j < d.nb; ++j) {
                if (es == 8) {
                    hipLaunchKernelGGL(k_csr_dg_probe<double>, dim3(cgrid(N)), dim3(CT), 0, s, d, pb.as<double>(), j, c);
                    o->apply_raw(pb.p, yb.p, nullptr);
                    hipLaunchKernelGGL(k_csr_dg_extract<double>, dim3(cgrid(N)), dim3(CT), 0, s, d, yb.as<double>(), R,
                                       val.as<double>(), j, c);
                } else {
                    hipLaunchKernelGGL(k_csr_dg_probe<float>, dim3(cgrid(N)), dim3(CT), 0, s, d, pb.as<float>(), j, c);
                    o->apply_raw(pb.p, yb.p, nullptr);
                    hipLaunchKernelGGL(k_csr_dg_extract<float>, dim3(cgrid(N)), dim3(CT), 0, s, d, yb.as<float>(), R,
                                       val.as<double>(), j, c);
                }
                FEM_HIP(hipGetLastError());
            }
    } else {
        Q2Dims d{ o->desc.n[0], o->desc.n[1] };
        hipLaunchKernelGGL(k_csr_q2_cols, dim3(cgrid(N)), dim3(CT), 0, s, d, R, col.as<int>());
        for (int c = 0; c < 25; ++c) {
            if (es == 8) {
                hipLaunchKernelGGL(k_csr_q2_probe<double>, dim3(cgrid(N)), dim3(CT), 0, s, d, pb.as<double>(), c);
                o->apply_raw(pb.p, yb.p, nullptr);
                hipLaunchKernelGGL(k_csr_q2_extract<double>, dim3(cgrid(N)), dim3(CT), 0, s, d, yb.as<double>(), R,
                                   val.as<double>(), c);
            } else {
                hipLaunchKernelGGL(k_csr_q2_probe<float>, dim3(cgrid(N)), dim3(CT), 0, s, d, pb.as<float>(), c);
                o->apply_raw(pb.p, yb.p, nullptr);
                hipLaunchKernelGGL(k_csr_q2_extract<float>, dim3(cgrid(N)), dim3(CT), 0, s, d, yb.as<float>(), R,
                                   val.as<double>(), c);
            }
            FEM_HIP(hipGetLastError());
        }
    }
    FEM_HIP(hipStreamSynchronize(s));
}

void CSR::apply(const Vec &x, Vec &y) const
{
    FEM_CHECK(x.n == N && y.n == N, FEM_ERR_ARG, "CSR apply: vector size");
    FEM_CHECK(x.prec == y.prec, FEM_ERR_ARG, "CSR apply: x and y of one precision");
    FEM_CHECK(x.data != y.data, FEM_ERR_ARG, "CSR apply: x and y must be distinct");
    hipStream_t s = ctx->stream;
    const unsigned G = (unsigned)std::min<int64_t>((N + 3) / 4, 65536);
    if (x.prec == FEM_F64)
        hipLaunchKernelGGL((k_csr_spmv<double, double>), dim3(G), dim3(256), 0, s, N, rowptr.as<int64_t>(),
                           col.as<int>(), val.as<double>(), (const double *)x.data, (double *)y.data);
    else
        hipLaunchKernelGGL((k_csr_spmv<double, float>), dim3(G), dim3(256), 0, s, N, rowptr.as<int64_t>(),
                           col.as<int>(), val.as<double>(), (const float *)x.data, (float *)y.data);
    FEM_HIP(hipGetLastError());
}

void CSR::diagonal(Vec &d) const
{
    FEM_CHECK(d.n == N && d.prec == FEM_F64, FEM_ERR_ARG, "CSR diagonal: an FP64 vector of N entries");
    hipLaunchKernelGGL(k_csr_diag, dim3(cgrid(N)), dim3(CT), 0, ctx->stream, N, rowptr.as<int64_t>(), col.as<int>(),
                       val.as<double>(), (double *)d.data);
    FEM_HIP(hipGetLastError());
}

void CSR::download(int64_t *rp, int *c, double *v) const
{
    hipStream_t s = ctx->stream;
    if (rp) FEM_HIP(hipMemcpyAsync(rp, rowptr.p, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, s));
    if (c) FEM_HIP(hipMemcpyAsync(c, col.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, s));
    if (v) FEM_HIP(hipMemcpyAsync(v, val.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
}

}  // namespace femamd
