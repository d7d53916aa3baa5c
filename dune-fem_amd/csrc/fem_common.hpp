// Shared host/device definitions of the MI355X matrix-free Galerkin library.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "dunefem_amd.h"

namespace femamd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FEM_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::femamd::Error(FEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define FEM_CHECK(cond, code, msg)                                                         \
    do {                                                                                   \
        if (!(cond)) throw ::femamd::Error((code), (msg));                                 \
    } while (0)

// Maximum number of 1D points/modes per direction supported by the kernels
// (polynomial order <= 7).
constexpr int KMAX = 8;
// internal preconditioner kind (not in the C header): an FP32 preconditioner
// applied to the FP64 vectors of an FP64 Krylov solve
constexpr int FEM_PRE_WIDENED = 100;

// 1D tables of one polynomial order, on the m = p+1 point Gauss rule.
//   B[q][a]   modal Legendre basis a at Gauss point q
//   Dt[q][r]  collocation derivative: d/dxi of the Lagrange polynomial of
//             Gauss point r, evaluated at Gauss point q
//   E[s][q]   Lagrange polynomial of Gauss point q evaluated at xi = s (face)
//   dE[s][q]  its derivative at xi = s
//   P[s][a], dP[s][a]  Legendre basis a and derivative at xi = s
//   Bd[q][a]  derivative of Legendre basis a at Gauss point q
//   S[a][b]   1D stiffness int_0^1 phi_a' phi_b'
//   w[q], x[q] Gauss weights and points on [0,1]
//   Dp, Dm    even/odd halves of Dt (skew-centrosymmetric on the symmetric
//             points, Dt[m-1-q][m-1-r] = -Dt[q][r]), with h = m/2, c = m/2:
//             Dp[q][r] = (Dt[q][r] + Dt[q][m-1-r]) / 2, Dp[q][h] = Dt[q][c]
//             (odd m), Dm[q][r] = (Dt[q][r] - Dt[q][m-1-r]) / 2,
//             Dm[h][r] = Dt[c][r] (odd m), for q, r < h
//   DTp, DTm  the same for the transpose Dt^T
template <typename T>
struct Tab1D {
    T B[KMAX][KMAX];
    T Bd[KMAX][KMAX];
    T S[KMAX][KMAX];
    T Dt[KMAX][KMAX];
    T E[2][KMAX];
    T dE[2][KMAX];
    T P[2][KMAX];
    T dP[2][KMAX];
    T w[KMAX];
    T x[KMAX];
    T Dp[KMAX / 2][KMAX / 2 + 1];
    T Dm[KMAX / 2 + 1][KMAX / 2];
    T DTp[KMAX / 2][KMAX / 2 + 1];
    T DTm[KMAX / 2 + 1][KMAX / 2];
};

// host-side table construction (fem_tables.cpp)
void gauss_legendre01(int m, double *x, double *w);
double legendre01(int n, double x);
double legendre01_d(int n, double x);
double lagrange2(int i, double x);
double lagrange2_d(int i, double x);
Tab1D<double> make_tab1d(int k);
Tab1D<float> to_float(const Tab1D<double> &t);

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // reduction workspace: partial sums + counters + result slots
    double *red_partial = nullptr;     // [RED_DOT_BLOCKS * RED_MAX_SLOTS]
    unsigned *red_counter = nullptr;   // [RED_MAX_SLOTS]
    double *red_result = nullptr;      // device scalar results [RED_MAX_SLOTS]
    double *host_result = nullptr;     // pinned mirror
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // one reference for the fem_ctx handle and one for every handle made on
    // it: the stream and workspaces live until the last of them is destroyed
    std::atomic<int> refs{1};
};

constexpr int RED_BLOCKS = 512;      // fixed grid of the reduction kernels
// fixed grid of the single and paired dot products (k_dot, the CG steps)
constexpr int RED_DOT_BLOCKS = 512;
constexpr int RED_THREADS = 256;
constexpr int RED_MAX_SLOTS = 64;    // simultaneous dot products (multi-dot)

struct Vec {
    Context *ctx = nullptr;
    int64_t n = 0;
    int prec = FEM_F64;
    void *data = nullptr;
    bool owner = true;
    size_t bytes() const { return (size_t)n * (prec == FEM_F64 ? 8 : 4); }
};

inline unsigned grid_for(int64_t n, int threads, unsigned cap = 1u << 20)
{
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > (int64_t)cap) g = cap;
    return (unsigned)g;
}

// XCD-aware block remap: consecutive hardware block ids are dealt to the 8
// XCDs round robin; give each XCD a contiguous range of logical blocks so
// neighbouring elements share that XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblocks)
{
    const unsigned nx = 8;
    unsigned q = nblocks / nx, r = nblocks % nx;
    unsigned xcd = bid % nx, local = bid / nx;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

}  // namespace femamd

struct fem_ctx { femamd::Context c; };
struct fem_vec { femamd::Vec v; femamd::Context *held = nullptr; };
