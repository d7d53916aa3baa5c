// Exact element block-Jacobi transform of the SIPG operator with cellwise
// kappa and b = 0: the element block is a Kronecker sum of 1D matrices, so
//   y_e = (Qx (x) Qy (x) Qz) diag(1 / (lx + ly + lz)) (Qx (x) Qy (x) Qz)^T x_e
// with the per-element eigenvectors Qe [3][K][K] (Q_d[r][c] = Qe[d*K2 + r*K + c],
// columns are eigenvectors) and eigenvalues le [3][K], read from global
// memory (Qg, lg) into the element's LDS slice Qs [3K^2 + 3K].  One element is
// handled by K*K lanes (i, j); lane (i, j) passes its z pencil x[i][j][0..K)
// in v and receives y[i][j][0..K).  A is the element's K^3 LDS scratch; each
// phase reads and rewrites the pencils of one direction, which partition the
// tensor, so only the direction changes need a barrier.  The transform
// synchronises the workgroup: every lane of it must call it.
#pragma once

#include <hip/hip_runtime.h>

namespace femamd {

// lanes of one workgroup of the block-Jacobi kernels: EPB elements of K*K lanes
template <int K>
struct BJShape {
    static constexpr int K2 = K * K, K3 = K2 * K, EPB = 64 / K2, THREADS = EPB * K2;
};

template <int K, typename T>
__device__ __forceinline__ void bj_element(const T *__restrict__ Qg, const T *__restrict__ lg, T *A, T *Qs, int i,
                                           int j, T (&v)[K])
{
    constexpr int K2 = K * K;
    // every lane of the element needs all of Qe and le: stage them in LDS once
    // (the element's K*K lanes copy 3K^2 + 3K values) instead of each lane
    // loading them from memory in every pass
    const int t = i * K + j;
    for (int k = t; k < 3 * K2; k += K2) Qs[k] = Qg[k];
    for (int k = t; k < 3 * K; k += K2) Qs[3 * K2 + k] = lg[k];
    __syncthreads();
    const T *Qe = Qs, *le = Qs + 3 * K2;
    T tmp[K];
    // Q_z^T along c: lane (i=a, j=b)
#pragma unroll
    for (int r = 0; r < K; ++r) {
        T s = 0;
#pragma unroll
        for (int c = 0; c < K; ++c) s = fma(Qe[2 * K2 + c * K + r], v[c], s);
        tmp[r] = s;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) A[i * K2 + j * K + r] = tmp[r];
    __syncthreads();
    // Q_y^T along b: lane (i=a, j=c')
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[i * K2 + b * K + j];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        T s = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) s = fma(Qe[1 * K2 + b * K + r], v[b], s);
        tmp[r] = s;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) A[i * K2 + r * K + j] = tmp[r];
    __syncthreads();
    // Q_x^T along a, scale, Q_x: lane (i=b', j=c')
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[b * K2 + i * K + j];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        T s = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) s = fma(Qe[0 * K2 + b * K + r], v[b], s);
        tmp[r] = s / (le[0 * K + r] + le[1 * K + i] + le[2 * K + j]);
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
        T s = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) s = fma(Qe[0 * K2 + r * K + b], tmp[b], s);
        v[r] = s;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) A[r * K2 + i * K + j] = v[r];
    __syncthreads();
    // Q_y along b: lane (i=a, j=c')
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[i * K2 + b * K + j];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        T s = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) s = fma(Qe[1 * K2 + r * K + b], v[b], s);
        tmp[r] = s;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) A[i * K2 + r * K + j] = tmp[r];
    __syncthreads();
    // Q_z along c: lane (i=a, j=b)
#pragma unroll
    for (int c = 0; c < K; ++c) v[c] = A[i * K2 + j * K + c];
    __syncthreads();   // A is free again for the caller's next element
#pragma unroll
    for (int r = 0; r < K; ++r) {
        T s = 0;
#pragma unroll
        for (int c = 0; c < K; ++c) s = fma(Qe[2 * K2 + r * K + c], v[c], s);
        tmp[r] = s;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) v[r] = tmp[r];
}

}  // namespace femamd
