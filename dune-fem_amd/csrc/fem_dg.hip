// SIPG DG operator on a structured hexahedral mesh of [0,1]^3: matrix-free
// apply by sum factorisation.
//
// Reference semantics: GalerkinOperator::evaluate / evaluateImpl
// (dune/fem/schemes/galerkin.hh:811-917, 1460-1496) with the SIPG + upwind
// integrands of dune/fempy/test/parallelTest.py:113-129 and
// pydemo/advectiondiffusion.py:30-56, quadrature orders 2p/2p+1
// (galerkin.hh:131-132) on (p+1)-point Gauss rules.
//
// Design (MI355X): one workgroup handles EPB elements with K*K threads per
// element (K = p+1).  A thread holds one K-long "pencil" of an element tensor
// in registers; contractions along the pencil are register-local, and the
// pencil direction is changed by a write/read through LDS.  The reference's
// dense nq x nb tables (default.hh:276-487) become K x K contractions per
// direction.  Faces are integrated by both adjacent elements ("owner
// computes"): each element evaluates the neighbour trace itself, so every
// element writes only its own DoFs and no atomics or locks are needed
// (the reference visits each face once and locks the scatter,
// galerkin.hh:879,951-991).  Face traces of the own element come from the
// quadrature-point values by 1D extrapolation; neighbour traces are a
// normal-direction contraction of the neighbour's modal DoFs followed by two
// tangential contractions (on the FP64 matrix cores at K = 4 in FP64, with
// DPP cross-lane moves at K = 4 in FP32, through LDS otherwise).  The two
// faces normal to a direction are integrated in the block that computes that
// direction's volume flux, where both pencils along it are in registers.
#include <hip/hip_bf16.h>

#include <type_traits>

#include "fem_dg.hpp"
#include "fem_dg_bj.hpp"

namespace femamd {

template <int K, typename T>
__device__ __forceinline__ void mvT(const T (&M)[KMAX][KMAX], const T *in, T *out)
{
#pragma unroll
    for (int a = 0; a < K; ++a) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < K; ++q) s = fma(M[q][a], in[q], s);
        out[a] = s;
    }
}

template <int K, typename T>
__device__ __forceinline__ T dotk(const T *v, const T *in)
{
    T s = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) s = fma(v[q], in[q], s);
    return s;
}

template <int K>
struct DGShape {
    static constexpr int K2 = K * K, K3 = K * K * K;
    // elements per workgroup: all threads of a workgroup form one wavefront,
    // so intra-element exchanges through LDS need no workgroup barrier
    static constexpr int EPB = 64 / K2;
    static constexpr int THREADS = EPB * K2;
    // LDS tensor layout (a, b, c) -> a*S1 + b*K + swz(c, a).  For K = 4 the
    // stride S1 = 20 plus the XOR swizzle make the x-, y- and z-pencil
    // accesses of the 16 lanes of an element hit 16 distinct 8-byte banks.
    static constexpr int S1 = (K == 4) ? 20 : K2 + 1;
    static constexpr int TS = K * S1;           // one padded K^3 tensor
    // per element: A (u at quadrature points), R (residual at quadrature
    // points), S (neighbour-trace scratch, 4 K^2); the element stride is
    // padded to an odd multiple of 16 doubles to separate the elements' banks
    static constexpr int SZ0 = 2 * TS + 4 * K2;
    static constexpr int SZ = ((SZ0 + 15) / 16) * 16 + (((SZ0 + 15) / 16) % 2 == 0 ? 16 : 0);
};

// Workgroup of the apply: one wavefront of EPB = 64 / K^2 elements, except
// at K = 5, where two wavefronts hold five elements (125 of 128 lanes busy
// instead of 50 of 64); the pencil exchanges of an element may then cross
// the two wavefronts and synchronise the workgroup.
template <int K>
struct ApplyShape {
    static constexpr int WG = (K == 5) ? 128 : 64;
    static constexpr int EPB = WG / (K * K);
};

template <int K>
__device__ __forceinline__ int lidx(int a, int b, int c)
{
    using S = DGShape<K>;
    return a * S::S1 + b * K + ((K == 4) ? (c ^ a) : c);
}

// Orders LDS accesses between the lanes of one wavefront.  LDS instructions of
// a wavefront execute in issue order, so only the compiler must be kept from
// moving memory operations across this point; no s_barrier and no wait on
// outstanding global loads is needed.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the exchange barrier of a workgroup of WG threads: a wavefront fence for
// one wavefront, the workgroup barrier for more
template <int WG>
__device__ __forceinline__ void lds_sync()
{
    if constexpr (WG > 64) __syncthreads();
    else wave_lds_sync();
}

// Cross-lane moves inside a 16-lane DPP row (one K = 4 element of the FP32
// and BF16 kernels; the FP64 kernel uses the matrix cores instead): CTRL is a
// quad_perm pattern (0x00..0xFF) or row_ror:N (0x120 + N).  Every lane of the
// wavefront must be active (no lane reads an inactive or out-of-row source, so
// the destination's previous value is never kept).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, false));
}
// row_ror:4k moves lane (4 i' + j) of a row to lane (4 ((i' + k) mod 4) + j):
// after it, lane (i, j) holds the value of lane (4 ((i - k) mod 4) + j)
constexpr int ROR_SRC_SIGN = -1;

// y(i, j) = sum_m B[j][m] x(i, m): x(i, m) sits in lane m of the lane's quad
template <typename T>
__device__ __forceinline__ T quad_contract(const T *Bj, T x)
{
    T y = Bj[0] * dpp_mov<0x00>(x);
    y = fma(Bj[1], dpp_mov<0x55>(x), y);
    y = fma(Bj[2], dpp_mov<0xAA>(x), y);
    return fma(Bj[3], dpp_mov<0xFF>(x), y);
}
// z(i, j) = sum_m B[i][m] y(m, j): y(m, j) sits in lane 4 m + j of the row
template <typename T>
__device__ __forceinline__ T row_contract(const T *Bi, int i, T y)
{
    T z = Bi[i] * y;
    z = fma(Bi[(i + ROR_SRC_SIGN * 1) & 3], dpp_mov<0x124>(y), z);
    z = fma(Bi[(i + ROR_SRC_SIGN * 2) & 3], dpp_mov<0x128>(y), z);
    return fma(Bi[(i + ROR_SRC_SIGN * 3) & 3], dpp_mov<0x12C>(y), z);
}

// ---- FP64 MFMA tangential contraction (K = 4) -----------------------------
// v_mfma_f64_4x4x4f64 computes four independent products D_b = A_b B_b.  Lane
// layout (measured on the device): A_b[r][k] and B_b[k][r] sit in lane
// 16 k + 4 b + r, D_b[r][c] in lane 16 r + 4 b + c.  An element is block b
// and its lane (i, j) is 16 i + 4 b + j.
// With N(m1, m2) in lane (m1, m2) read as the A operand (A = N^T) and the
// constant B operand Bop[k][c] = B[c][k] (lane (i, j) supplies B[j][i]):
//   D1[r][c] = sum_m1 B[c][m1] N(m1, r)  ->  lane (m2, q1) holds Y(q1, m2)
//   D2 = D1 (as A) x Bop: D2[r][c] = sum_m2 Y(r, m2) B[c][m2] = Z(q1=r, q2=c)
// so lane (i, j) receives Z(q1 = i, q2 = j) = sum B[i][m1] B[j][m2] N(m1, m2).
__device__ __forceinline__ double mf64_tangential(double nm, double bji)
{
    const double y = __builtin_amdgcn_mfma_f64_4x4x4f64(nm, bji, 0.0, 0, 0, 0);
    return __builtin_amdgcn_mfma_f64_4x4x4f64(y, bji, 0.0, 0, 0, 0);
}

// Legendre-basis contractions use the parity of the basis on the symmetric
// Gauss points, B[K-1-q][a] = (-1)^a B[q][a]: out[q] and out[K-1-q] share the
// even and odd partial sums.
template <int K, typename T>
__device__ __forceinline__ void mvB(const T (&M)[KMAX][KMAX], const T *in, T *out)
{
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
        T ev = 0, od = 0;
#pragma unroll
        for (int a = 0; a < K; a += 2) ev = fma(M[q][a], in[a], ev);
#pragma unroll
        for (int a = 1; a < K; a += 2) od = fma(M[q][a], in[a], od);
        out[q] = ev + od;
        out[K - 1 - q] = ev - od;
    }
    if (K % 2) {
        T ev = 0;
#pragma unroll
        for (int a = 0; a < K; a += 2) ev = fma(M[K / 2][a], in[a], ev);
        out[K / 2] = ev;
    }
}

template <int K, typename T>
__device__ __forceinline__ void mvBT(const T (&M)[KMAX][KMAX], const T *in, T *out)
{
    T sp[(K + 1) / 2], sm[(K + 1) / 2];
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
        sp[q] = in[q] + in[K - 1 - q];
        sm[q] = in[q] - in[K - 1 - q];
    }
#pragma unroll
    for (int a = 0; a < K; ++a) {
        T acc = (K % 2 && a % 2 == 0) ? M[K / 2][a] * in[K / 2] : T(0);
#pragma unroll
        for (int q = 0; q < K / 2; ++q) acc = fma(M[q][a], (a % 2) ? sm[q] : sp[q], acc);
        out[a] = acc;
    }
}

// Collocation-derivative contractions by even/odd halves: on the symmetric
// Gauss points Dt is skew-centrosymmetric, so with s_r = in[r] + in[K-1-r],
// a_r = in[r] - in[K-1-r] (r < K/2):
//   out[q] = E_q + O_q, out[K-1-q] = O_q - E_q,
//   E_q = sum_r Dp[q][r] s_r (+ Dp[q][K/2] in[K/2]), O_q = sum_r Dm[q][r] a_r,
// and for odd K the middle row out[K/2] = sum_r Dm[K/2][r] a_r.  Half the
// coefficients of the full product, and fewer operations for odd K.
template <int K, typename T>
__device__ __forceinline__ void mvD(const T (&Pp)[KMAX / 2][KMAX / 2 + 1], const T (&Pm)[KMAX / 2 + 1][KMAX / 2],
                                    const T *in, T *out)
{
    constexpr int H = K / 2;
    T sv[H > 0 ? H : 1], av[H > 0 ? H : 1];
#pragma unroll
    for (int r = 0; r < H; ++r) {
        sv[r] = in[r] + in[K - 1 - r];
        av[r] = in[r] - in[K - 1 - r];
    }
#pragma unroll
    for (int q = 0; q < H; ++q) {
        T e = (K % 2) ? Pp[q][H] * in[H] : T(0), o = 0;
#pragma unroll
        for (int r = 0; r < H; ++r) {
            e = fma(Pp[q][r], sv[r], e);
            o = fma(Pm[q][r], av[r], o);
        }
        out[q] = e + o;
        out[K - 1 - q] = o - e;
    }
    if (K % 2) {
        T m = 0;
#pragma unroll
        for (int r = 0; r < H; ++r) m = fma(Pm[H][r], av[r], m);
        out[H] = m;
    }
}

// ---- BF16 MFMA contractions (v_mfma_f32_4x4x4bf16_1k, 16 blocks) --------
// Operand layout of the 16-block 4x4x4 instruction: lane (i + 4b) supplies
// row i of A_b and column i of B_b (4 values each) and receives column i of
// D_b.  With A_b = M for every block, a lane's own K-pencil is contracted by
// one instruction: D(lane) = M * pencil(lane), FP32 accumulation.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (v_cvt_pk_bf16_f32 on this GPU)
__device__ __forceinline__ short to_bf16(float x)
{
    return __builtin_bit_cast(short, (__bf16)x);
}

struct MFRows {
    bf16x4 B, BT, D, DT;   // row (lane % 4) of B, B^T, Dt, Dt^T (4x4 leading block)
};

template <int K>
__device__ __forceinline__ bf16x4 mf_row(const float (&M)[KMAX][KMAX], bool trans, int r)
{
    bf16x4 v;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float x = (r < K && k < K) ? (trans ? M[k][r] : M[r][k]) : 0.0f;
        v[k] = to_bf16(x);
    }
    return v;
}

// out = M in (trans: M^T in) for a K-pencil; K <= 4 padded to 4, K = 5 as a
// 4x4 MFMA block plus the fifth row/column on the VALU
template <int K>
__device__ __forceinline__ void mf_mv(const bf16x4 &A, const float (&M)[KMAX][KMAX], bool trans, const float *in,
                                      float *out)
{
    bf16x4 b;
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = to_bf16(k < K ? in[k] : 0.0f);
    const f32x4 z = { 0.f, 0.f, 0.f, 0.f };
    const f32x4 d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(A, b, z, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < (K < 4 ? K : 4); ++q) out[q] = d[q];
    if (K == 5) {
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = fmaf(trans ? M[4][q] : M[q][4], in[4], out[q]);
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 5; ++k) s = fmaf(trans ? M[k][4] : M[4][k], in[k], s);
        out[4] = s;
    }
}

// contraction policy: VALU in T, or BF16 MFMA with FP32 accumulation
template <int K, typename T, bool MF>
struct Ctr {
    static __device__ __forceinline__ void B(const Tab1D<T> &tb, const MFRows &r, const T *in, T *out)
    {
        if constexpr (MF) mf_mv<K>(r.B, tb.B, false, in, out);
        else mvB<K>(tb.B, in, out);
    }
    static __device__ __forceinline__ void BT(const Tab1D<T> &tb, const MFRows &r, const T *in, T *out)
    {
        if constexpr (MF) mf_mv<K>(r.BT, tb.B, true, in, out);
        else mvBT<K>(tb.B, in, out);
    }
    static __device__ __forceinline__ void D(const Tab1D<T> &tb, const MFRows &r, const T *in, T *out)
    {
        if constexpr (MF) mf_mv<K>(r.D, tb.Dt, false, in, out);
        else mvD<K>(tb.Dp, tb.Dm, in, out);
    }
    static __device__ __forceinline__ void DT(const Tab1D<T> &tb, const MFRows &r, const T *in, T *out)
    {
        if constexpr (MF) mf_mv<K>(r.DT, tb.Dt, true, in, out);
        else mvD<K>(tb.DTp, tb.DTm, in, out);
    }
};

// KQ: kappa given at the quadrature points (FEM_KAPPA_QP); otherwise one
// value per element (FEM_KAPPA_CELL) or one constant (FEM_KAPPA_CONST).
// SM: the damped block-Jacobi sweep of a multigrid level in the same pass,
// xout = u + omega B^-1 (r - A u), instead of w = A u (sm carries r, the
// element blocks' eigen data, omega and xout; w and dot are not used)
template <int K, typename T, bool MF, bool KQ, bool SM = false, bool RX = false>
__global__ void __launch_bounds__(ApplyShape<K>::WG)
    k_dg_apply(const DGArgs<T> a, const T *__restrict__ u, T *__restrict__ w, const int *__restrict__ skip,
               double *__restrict__ dot, const DGSmooth<T> sm)
{
    using S = DGShape<K>;
    constexpr int K2 = S::K2, K3 = S::K3, EPB = ApplyShape<K>::EPB, WG = ApplyShape<K>::WG;
    if (skip && *skip) return;
    // FP64 MFMA path (K = 4, lane 16 i + 4 el + j): a 16-lane LDS access
    // group holds one i of all four elements.  The layout (a, b, c) ->
    // a * 20 + b * 5 + c with an element stride of 164 doubles puts every
    // pencil access pattern of the kernel on 16 distinct 8-byte banks, with
    // compile-time offsets only (searched exhaustively over additive layouts).
    // Otherwise the layout of DGShape (one element per 16-lane group).
    constexpr bool F64MF_ = (K == 4) && std::is_same<T, double>::value && !MF;
    // K = 3 and 5 without BF16 MFMA (lane el K^2 + i K + j): unpadded
    // tensors (a, b, c) -> a K^2 + b K + c, face-trace scratch rows of 4 at
    // K = 3, and an odd element stride (107 and 361 words); from the same
    // exhaustive bank model this cuts the modelled conflicts of the FP64
    // accesses 3.1x (K = 3) and 2.2x (K = 5)
    constexpr bool D35 = (K == 3 || K == 5) && !MF;
    constexpr int S1 = F64MF_ ? 20 : D35 ? K2 : S::S1;
    constexpr int SB = F64MF_ ? 5 : K;
    constexpr int SXI = (D35 && K == 3) ? 4 : K;   // face-trace scratch: row stride
    constexpr int SXQ = K * SXI;                   // and quantity stride
    // with pointwise kappa a third tensor KA holds the element's volume kappa
    // values: loaded once as x-pencils, the y- and z-pencils come from LDS
    // (K = 4 only: the other orders keep the face-trace scratch there)
    constexpr bool KLDS = (K == 4) && KQ;
    // the pointwise/nonlinear reaction is compiled into the plain applies
    // with constant or cellwise kappa, and into a separate variant (RX) of
    // the pointwise-kappa apply that only operators with a reaction use (not
    // into the BF16-MFMA or fused-smoother variants): the pointwise-kappa
    // kernel without it keeps its code and registers
    constexpr bool CV = (!KQ && !MF && !SM) || RX;
    constexpr int SZ = F64MF_ ? (KLDS ? 244 : 164) : D35 ? (K == 3 ? 107 : 361) : S::SZ;
    static_assert(!F64MF_ || SZ >= (KLDS ? 3 : 2) * K * S1, "tensors per element");
    static_assert(F64MF_ || SZ >= 2 * K * S1 + 4 * SXQ, "tensors and face-trace scratch per element");
    // DGShape layout at K = 4: the third tensor ends at 2 TS + lidx(3,3,3) + 1
    static_assert(F64MF_ || !KLDS || SZ >= 2 * K * S1 + 3 * S1 + 3 * K + 4, "tensors per element");
    __shared__ T smem[EPB * SZ];
    const Tab1D<T> &tb = *a.tab;
    using C = Ctr<K, T, MF>;
    MFRows mr;
    if constexpr (MF) {
        const int r = threadIdx.x & 3;
        mr.B = mf_row<K>(tb.B, false, r);
        mr.BT = mf_row<K>(tb.B, true, r);
        mr.D = mf_row<K>(tb.Dt, false, r);
        mr.DT = mf_row<K>(tb.Dt, true, r);
    }

    const int tid = threadIdx.x;
    // FP64 at K = 4: lane 16 i + 4 el + j, the block layout of the FP64 4x4x4
    // MFMA (one block per element); otherwise lane el K^2 + i K + j
    constexpr bool F64MF = (K == 4) && std::is_same<T, double>::value && !MF;
    const int el0 = F64MF ? (tid >> 2) & 3 : tid / K2;
    const bool wr = el0 < EPB;   // lanes beyond EPB*K2 (K = 3, 5) only shadow element 0
    const int el = wr ? el0 : 0;
    const int i = F64MF ? tid >> 4 : (tid % K2) / K, j = F64MF ? tid & 3 : tid % K;
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
    const int E = nx * ny * nz;
    const int blk = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int e_raw = blk * EPB + el;
    bool active = wr && e_raw < E;
    // shadow lanes (el = 0 when !wr) take the block's first element, so e
    // and its coordinates below always describe the same in-range element:
    // every neighbour index formed from them stays inside the grid
    int e = e_raw < E ? e_raw : 0;
    int ex, ey, ez;
    if (a.slabs) {
        // the workgroup's elements share one x-row.  Rows are visited in
        // slabs of ty (16, 8, 4, 2 or 1, dividing ny) y-rows, each slab
        // through all z, so the rows below and above (z -+ 1) are read a few
        // hundred workgroups apart on the same XCD and hit its L2: the launch
        // then fetches about 1.1x the unique input instead of 1.8x.  The
        // divisions act on the wave-uniform block index only, by host-made
        // fixed-point reciprocals.
        const int rpb = a.rpb, ty = a.ty;
        const int t = (int)udiv((unsigned)blk, a.dv_slab), rem = blk - t * (rpb * ty * nz);
        const int z = (int)udiv((unsigned)rem, a.dv_zrow), rem2 = rem - z * (rpb * ty);
        const int yl = (int)udiv((unsigned)rem2, a.dv_rpb), xu = rem2 - yl * rpb;
        ex = xu * EPB + el;
        // past the end of the row (the last group of a padded row): a
        // shadow of the row's last element, computed and not written
        const bool inrow = ex < nx;
        ex = inrow ? ex : nx - 1;
        active = wr && inrow;
        ey = t * ty + yl;
        ez = z;
        e = ex + nx * (ey + ny * ez);
    } else {
        ex = e % nx;
        ey = (e / nx) % ny;
        ez = e / (nx * ny);
    }
    const int ec[3] = { ex, ey, ez };
    const int stride[3] = { 1, nx, nx * ny };
    // FP64 at K = 4: the four elements of a wavefront are consecutive
    // elements of one x-row (lane 16 i + 4 el + j; in a padded last group the
    // shadows follow the row's last element, whose +x face is a boundary
    // face), so an x-face between two of them takes the neighbour's trace
    // from the neighbour's own lanes
    // (lane +- 4), where it is the neighbour's own extrapolated trace, instead
    // of contracting the neighbour's DoFs.  Only the first element of the
    // wavefront loads its -x neighbour and the last its +x neighbour: one
    // x pencil per lane instead of two.
    constexpr bool XSH = (K == 4) && std::is_same<T, double>::value && !MF;
    const int xs = (el0 == 0) ? 0 : 1;   // the x side whose neighbour this lane loads

    T *A = smem + el * SZ;
    T *R = A + K * S1;
    T *KA = R + K * S1;   // pointwise kappa at the volume points (KLDS only;
                          // at K = 4 the face-trace scratch Sx is not used)
    auto L = [&](int a_, int b_, int c_) {
        if constexpr (F64MF_ || D35) return a_ * S1 + b_ * SB + c_;
        else return lidx<K>(a_, b_, c_);
    };
    T *Sx = R + K * S1;   // [2 faces][2 quantities][K2] (K != 4 only)

    // weights folded per thread: volume point (i, j) of a pencil, scaled by
    // |K| / h_d^2 for the diffusion part; advection scaled by h_d
    const T detJ = a.h[0] * a.h[1] * a.h[2];
    const T wij = tb.w[i] * tb.w[j];
    T v[K], tmp[K];

    // ---- all global loads are issued up front ------------------------------
    // own DoFs (z pencil), the six neighbour pencils through this lane's
    // tangential modes, and the coefficient values this lane needs; the
    // neighbour data is contracted in the normal direction while the volume
    // work runs, so its latency is hidden behind the volume phase
    const T *ue = u + (size_t)e * K3;
#pragma unroll
    for (int c = 0; c < K; ++c) v[c] = ue[i * K2 + j * K + c];
    T nbr[3][2][K];
    bool has[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            has[d][s2] = s2 ? ec[d] + 1 < a.n[d] : ec[d] > 0;
            if (XSH && d == 0) continue;   // one x pencil, below
            const T *un = u + (size_t)(has[d][s2] ? e + (s2 ? stride[d] : -stride[d]) : e) * K3;
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const int off = d == 0 ? m * K2 + i * K + j : d == 1 ? i * K2 + m * K + j : i * K2 + j * K + m;
                nbr[d][s2][m] = un[off];
            }
        }
    T nbx[K];
    if constexpr (XSH) {
        // the middle elements need no x pencil: their own DoFs (cached) stand in
        const bool xload = has[0][xs] && (el0 == 0 || el0 == EPB - 1);
        const T *un = u + (size_t)(xload ? e + (xs ? 1 : -1) : e) * K3;
#pragma unroll
        for (int m = 0; m < K; ++m) nbx[m] = un[m * K2 + i * K + j];
    }
    const bool kcell = a.kappa_mode == FEM_KAPPA_CELL;
    T kx[K], ky[K], kz[K], kf[3][2];
    {
        const T kel0 = KQ ? T(0) : kcell ? a.kcell[e] : a.kappa_const;
        const T *kv = KQ ? a.kvol + (size_t)e * K3 : nullptr;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            kx[q] = KQ ? kv[q + K * i + K2 * j] : kel0;
            if constexpr (!KLDS) {
                ky[q] = KQ ? kv[i + K * q + K2 * j] : kel0;
                kz[q] = KQ ? kv[i + K * j + K2 * q] : kel0;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int N0 = nx + (d == 0), N1 = ny + (d == 1);
            const int fbase = ec[0] + N0 * (ec[1] + N1 * ec[2]);
            const int fstep = d == 0 ? 1 : d == 1 ? nx : nx * ny;
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
                if constexpr (KQ) kf[d][s2] = a.kface[d][(size_t)(fbase + (s2 ? fstep : 0)) * K2 + i + K * j];
                else kf[d][s2] = (kcell && has[d][s2]) ? a.kcell[e + (s2 ? stride[d] : -stride[d])] : kel0;
            }
        }
    }
    // normal contraction of the neighbour pencils: thread (i=m1, j=m2)
    T nval[3][2], nder[3][2];
    // B[j][i]: this lane's entry of the constant MFMA operand
    T bji = T(0);
    if constexpr (F64MF) bji = tb.B[j][i];

    // ---- forward: modal -> quadrature points ------------------------------
    // Z layout: thread (i=a, j=b) holds u[a][b][0..K)
    C::B(tb, mr, v, tmp);
    if (wr)
#pragma unroll
        for (int q = 0; q < K; ++q) A[L(i, j, q)] = tmp[q];
    if constexpr (KLDS)
#pragma unroll
        for (int q = 0; q < K; ++q) KA[L(q, i, j)] = kx[q];
    lds_sync<WG>();
    // Y layout: thread (i=a, j=qz); each thread rewrites the pencil it read
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[L(i, b, j)];
    C::B(tb, mr, v, tmp);
    lds_sync<WG>();
    if (wr)
#pragma unroll
        for (int q = 0; q < K; ++q) A[L(i, q, j)] = tmp[q];
    lds_sync<WG>();
    T nvx = 0, ndx = 0;   // XSH: the loaded x neighbour's normal contraction
    if constexpr (XSH) {
        T val = 0, der = 0;
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const T sg = (xs == 1 && (m & 1)) ? T(-1) : T(1);
            val = fma(sg * tb.P[1][m], nbx[m], val);
            der = fma(xs == 1 ? -sg * tb.dP[1][m] : tb.dP[1][m], nbx[m], der);
        }
        nvx = has[0][xs] ? val : T(0);
        ndx = has[0][xs] ? der : T(0);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            if (XSH && d == 0) continue;
            // the neighbour below (s2 = 0) is evaluated at its face xi = 1,
            // the one above at xi = 0: P(0) = (-1)^m P(1), P'(0) = -(-1)^m P'(1)
            T val = 0, der = 0;
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const T sg = (s2 == 1 && (m & 1)) ? T(-1) : T(1);
                val = fma(sg * tb.P[1][m], nbr[d][s2][m], val);
                der = fma(s2 == 1 ? -sg * tb.dP[1][m] : tb.dP[1][m], nbr[d][s2][m], der);
            }
            nval[d][s2] = has[d][s2] ? val : T(0);
            nder[d][s2] = has[d][s2] ? der : T(0);
        }
    // ---- faces: neighbour traces, numerical fluxes, lifting ----------------
    // neighbour trace value / normal derivative at face point (q1=i, q2=j):
    // two tangential contractions of the lane's modal (m1=i, m2=j) values
    auto neighbour_traces = [&](int d, T (&nvs)[2], T (&nds)[2]) {
        if constexpr (F64MF) {
            // Z = B N B^T on the matrix cores: two chained 4x4x4 products per
            // quantity, no LDS and no cross-lane moves
            if (XSH && d == 0) {
                // the one loaded x neighbour (both entries; face_terms picks)
                nvs[0] = nvs[1] = mf64_tangential(nvx, bji);
                nds[0] = nds[1] = mf64_tangential(ndx, bji);
                return;
            }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                nvs[s] = mf64_tangential(nval[d][s], bji);
                nds[s] = mf64_tangential(nder[d][s], bji);
            }
        } else if constexpr (K == 4) {
            // an element is one 16-lane DPP row: contract across lanes, no LDS
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                nvs[s] = row_contract(tb.B[i], i, quad_contract(tb.B[j], nval[d][s]));
                nds[s] = row_contract(tb.B[i], i, quad_contract(tb.B[j], nder[d][s]));
            }
        } else {
            lds_sync<WG>();
            if (wr) {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    Sx[(s * 2 + 0) * SXQ + i * SXI + j] = nval[d][s];
                    Sx[(s * 2 + 1) * SXQ + i * SXI + j] = nder[d][s];
                }
            }
            lds_sync<WG>();
            // second tangential index, in place: thread (i=m1, j=q2)
            T t2v[4], w2[K];
#pragma unroll
            for (int sq = 0; sq < 4; ++sq) {
#pragma unroll
                for (int m = 0; m < K; ++m) w2[m] = Sx[sq * SXQ + i * SXI + m];
                t2v[sq] = dotk<K>(tb.B[j], w2);
            }
            lds_sync<WG>();
            if (wr)
#pragma unroll
                for (int sq = 0; sq < 4; ++sq) Sx[sq * SXQ + i * SXI + j] = t2v[sq];
            lds_sync<WG>();
            // first tangential index: thread (i=q1, j=q2)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                T nv = 0, nd = 0;
#pragma unroll
                for (int m = 0; m < K; ++m) {
                    nv = fma(tb.B[i][m], Sx[(s * 2 + 0) * SXQ + m * SXI + j], nv);
                    nd = fma(tb.B[i][m], Sx[(s * 2 + 1) * SXQ + m * SXI + j], nd);
                }
                nvs[s] = nv;
                nds[s] = nd;
            }
        }
    };
    // the two faces normal to d: own traces from the u_q pencil along d (up),
    // numerical fluxes, lifting into the residual pencil along d (rp); the
    // lane (i=q1, j=q2) holds both pencils through face point (q1, q2)
    auto face_terms = [&](int d, const T *up, T *rp) {
        T nvs[2], nds[2];
        neighbour_traces(d, nvs, nds);
        const T Wf = wij * a.farea[d];
        const T ihd = a.ih[d];
        // extrapolation to the faces from the symmetric Gauss points:
        // E[1][q] = E[0][K-1-q], dE[1][q] = -dE[0][K-1-q]
        auto Es = [&](int s, int m) { return s ? tb.E[0][K - 1 - m] : tb.E[0][m]; };
        auto dEs = [&](int s, int m) { return s ? -tb.dE[0][K - 1 - m] : tb.dE[0][m]; };
        // own traces (value, d/dxi) on both faces normal to d
        T ut[2], dt[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            T uin = 0, din = 0;
#pragma unroll
            for (int m = 0; m < K; ++m) {
                uin = fma(Es(s, m), up[m], uin);
                din = fma(dEs(s, m), up[m], din);
            }
            ut[s] = uin;
            dt[s] = din;
        }
        if (XSH && d == 0) {
            // the x neighbours inside the wavefront: element el + 1 (lane + 4)
            // holds the trace of the shared face as its own lower face, el - 1
            // (lane - 4) as its upper face; all lanes take part in the moves
            const int lane = threadIdx.x & 63;
            const T pv = __shfl(ut[0], (lane + 4) & 63, 64), pd = __shfl(dt[0], (lane + 4) & 63, 64);
            const T mv = __shfl(ut[1], (lane + 60) & 63, 64), md = __shfl(dt[1], (lane + 60) & 63, 64);
            if (el0 < EPB - 1) { nvs[1] = pv; nds[1] = pd; }
            if (el0 > 0) { nvs[0] = mv; nds[0] = md; }
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (!has[0][s]) { nvs[s] = T(0); nds[s] = T(0); }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const T nrm = s ? T(1) : T(-1);
            const bool interior = has[d][s];
            const T uin = ut[s];
            const T din = dt[s] * (ihd * nrm);
            // upwind split of b.n for the outward normal nrm e_d
            const T bp = s ? a.bpos[d] : a.bneg[d], bm = s ? a.bneg[d] : a.bpos[d];
            T kin, kout, kp;
            if constexpr (!KQ) {
                // cellwise kappa: average of the two elements' values
                // (the constant case has kout = kin)
                kin = kx[0];
                kout = kf[d][s];
                kp = T(0.5) * (kin + kout);
            } else {
                kin = kout = kp = kf[d][s];
            }
            const T pen = a.pen[d];
            // one formula for interior and boundary faces: on a boundary face
            // the neighbour trace is 0 (nval, nder are zeroed), kout = kp =
            // kin, and the consistency and symmetry terms carry weight 1
            // instead of 1/2 (Nitsche terms for u = 0)
            // (the FP64 Q3 kernel takes this form for both; the other orders
            // branch, which keeps their register use down)
            T c0, c1;
            if (F64MF || interior) {
                const T wgt = interior ? T(0.5) : T(1);
                const T nv = nvs[s], nd = nds[s];
                const T dout = nd * (ihd * nrm);
                const T jump = uin - nv;
                const T sigma = wgt * fma(kin, din, kout * dout);
                c0 = Wf * (fma(kp * pen, jump, -sigma) + fma(bp, uin, -bm * nv));
                c1 = Wf * (-wgt * kin * jump);
            } else {
                c0 = Wf * fma(fma(kin, pen, bp), uin, -kin * din);
                c1 = Wf * (-kin * uin);
            }
            c1 *= nrm * ihd;
#pragma unroll
            for (int m = 0; m < K; ++m) rp[m] = fma(Es(s, m), c0, fma(dEs(s, m), c1, rp[m]));
        }
    };

    // ---- volume fluxes, each direction with its two faces -------------------
    // X layout: thread (i=qy, j=qz)
    T uq[K];
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[L(b, i, j)];
    C::B(tb, mr, v, uq);
    lds_sync<WG>();
    if (wr)
#pragma unroll
        for (int q = 0; q < K; ++q) A[L(q, i, j)] = uq[q];

    // x direction: residual pencil kept in registers (X layout)
    T Rx[K];
    {
        const T sc = wij * a.vsc[0], bh = a.bh[0];
        T g[K], r[K];
        C::D(tb, mr, uq, g);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            r[q] = (tb.w[q] * sc) * fma(kx[q], g[q], -bh * uq[q]);
        }
        C::DT(tb, mr, r, Rx);
        if (a.c != T(0)) {
            const T cm = wij * detJ * a.c;
#pragma unroll
            for (int q = 0; q < K; ++q) Rx[q] = fma(tb.w[q] * cm, uq[q], Rx[q]);
        }
        if constexpr (CV) {
            // pointwise and nonlinear reaction at this lane's K volume points
            // (qx = q, qy = i, qz = j): (cvol + nl_coef u^nl_pow) u
            if (a.cvol || a.nl_pow) {
                const T cm = wij * detJ;
                const T *cv = a.cvol ? a.cvol + ((size_t)e * K3 + K * i + K2 * j) : nullptr;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    T r = cv ? cv[q] : T(0);
                    if (a.nl_pow) {
                        T pw = a.nl_coef;
                        for (int t = 0; t < a.nl_pow; ++t) pw *= uq[q];
                        r += pw;
                    }
                    Rx[q] = fma(tb.w[q] * cm * r, uq[q], Rx[q]);
                }
            }
        }
        face_terms(0, uq, Rx);
    }
    lds_sync<WG>();
    // Y layout of u_q: thread (i=qx, j=qz) -> R = y part
    {
        const T sc = wij * a.vsc[1], bh = a.bh[1];
        T g[K], r[K];
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = A[L(i, q, j)];
        if constexpr (KLDS)
#pragma unroll
            for (int q = 0; q < K; ++q) ky[q] = KA[L(i, q, j)];
        C::D(tb, mr, v, g);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            r[q] = (tb.w[q] * sc) * fma(ky[q], g[q], -bh * v[q]);
        }
        C::DT(tb, mr, r, tmp);
        face_terms(1, v, tmp);
        if (wr)
#pragma unroll
            for (int q = 0; q < K; ++q) R[L(i, q, j)] = tmp[q];
    }
    lds_sync<WG>();
    // Z layout of u_q: thread (i=qx, j=qy) -> R += z part
    {
        const T sc = wij * a.vsc[2], bh = a.bh[2];
        T g[K], r[K];
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = A[L(i, j, q)];
        if constexpr (KLDS)
#pragma unroll
            for (int q = 0; q < K; ++q) kz[q] = KA[L(i, j, q)];
        C::D(tb, mr, v, g);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            r[q] = (tb.w[q] * sc) * fma(kz[q], g[q], -bh * v[q]);
        }
        C::DT(tb, mr, r, tmp);
#pragma unroll
        for (int q = 0; q < K; ++q) tmp[q] += R[L(i, j, q)];
        face_terms(2, v, tmp);
        if (wr)
#pragma unroll
            for (int q = 0; q < K; ++q) R[L(i, j, q)] = tmp[q];
    }
    lds_sync<WG>();

    // ---- backward: quadrature points -> modal test functions ---------------
    // X layout: thread (i=qy, j=qz)
    T Rt[K];
#pragma unroll
    for (int q = 0; q < K; ++q) Rt[q] = Rx[q] + R[L(q, i, j)];
    C::BT(tb, mr, Rt, tmp);
    if (wr)
#pragma unroll
        for (int q = 0; q < K; ++q) A[L(q, i, j)] = tmp[q];
    lds_sync<WG>();
    // Y layout: thread (i=a, j=qz)
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = A[L(i, q, j)];
    C::BT(tb, mr, v, tmp);
    lds_sync<WG>();
    if (wr)
#pragma unroll
        for (int q = 0; q < K; ++q) A[L(i, q, j)] = tmp[q];
    lds_sync<WG>();
    // Z layout: thread (i=a, j=b)
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = A[L(i, j, q)];
    C::BT(tb, mr, v, tmp);
    if constexpr (SM) {
        // the element's residual pencil d = r - A u, then the exact element
        // block inverse through the element's LDS slice (K^3 values, then
        // its eigen data), every lane of the workgroup taking part
        static_assert(SZ >= K3 + 3 * K2 + 3 * K, "the element's LDS slice holds the block-inverse scratch");
        const T *re = sm.r + (size_t)e * K3;
#pragma unroll
        for (int c = 0; c < K; ++c) tmp[c] = re[i * K2 + j * K + c] - tmp[c];
        __syncthreads();   // the slice's last reads above are done
        T *Ab = smem + el * SZ;
        bj_element<K>(sm.Q + (size_t)e * 3 * K2, sm.lam + (size_t)e * 3 * K, Ab, Ab + K3, i, j, tmp);
        if (active) {
            T *xe = sm.xout + (size_t)e * K3;
#pragma unroll
            for (int c = 0; c < K; ++c) xe[i * K2 + j * K + c] = ue[i * K2 + j * K + c] + sm.omega * tmp[c];
        }
        return;
    } else if (active) {
        T *we = w + (size_t)e * K3;
#pragma unroll
        for (int c = 0; c < K; ++c) we[i * K2 + j * K + c] = tmp[c];
    }
    if (dot) {
        // this workgroup's <u, w> (fixed summation order)
        double acc = 0;
        if (active)
#pragma unroll
            for (int c = 0; c < K; ++c) acc += (double)ue[i * K2 + j * K + c] * (double)tmp[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if constexpr (WG > 64) {
            __shared__ double wsum[WG / 64];
            if ((tid & 63) == 0) wsum[tid >> 6] = acc;
            __syncthreads();
            if (tid == 0) {
                double t = 0;
                for (int k = 0; k < WG / 64; ++k) t += wsum[k];
                dot[blockIdx.x] = t;
            }
        } else {
            if (tid == 0) dot[blockIdx.x] = acc;
        }
    }
}

template <int K, typename T, bool MF = false>
static void launch_apply(const DGArgs<T> &a, const T *u, T *w, const int *skip, hipStream_t s, double *dot = nullptr)
{
    using S = DGShape<K>;
    using A = ApplyShape<K>;
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    FEM_CHECK(E * S::K3 < (int64_t)1 << 31, FEM_ERR_UNSUPPORTED, "DG apply: more than 2^31 DoFs");
    const unsigned nb = (unsigned)dg_blocks(K, a.n);
    bool rx = false;
    if constexpr (!MF) rx = a.cvol || a.nl_pow;
    if (a.kappa_mode == FEM_KAPPA_QP && rx) {
        if constexpr (!MF)
            hipLaunchKernelGGL((k_dg_apply<K, T, false, true, false, true>), dim3(nb), dim3(A::WG), 0, s, a, u, w,
                               skip, dot, DGSmooth<T>{});
    } else if (a.kappa_mode == FEM_KAPPA_QP)
        hipLaunchKernelGGL((k_dg_apply<K, T, MF, true>), dim3(nb), dim3(A::WG), 0, s, a, u, w, skip, dot,
                           DGSmooth<T>{});
    else
        hipLaunchKernelGGL((k_dg_apply<K, T, MF, false>), dim3(nb), dim3(A::WG), 0, s, a, u, w, skip, dot,
                           DGSmooth<T>{});
    FEM_HIP(hipGetLastError());
}

template <int K, typename T>
static void launch_smooth(const DGArgs<T> &a, const T *u, const DGSmooth<T> &sm, const int *skip, hipStream_t s)
{
    using S = DGShape<K>;
    using A = ApplyShape<K>;
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    FEM_CHECK(E * S::K3 < (int64_t)1 << 31, FEM_ERR_UNSUPPORTED, "DG apply: more than 2^31 DoFs");
    FEM_CHECK(a.kappa_mode != FEM_KAPPA_QP, FEM_ERR_UNSUPPORTED, "smoothing sweep: cellwise or constant kappa");
    FEM_CHECK(sm.r && sm.Q && sm.lam && sm.xout && sm.xout != u, FEM_ERR_ARG,
              "smoothing sweep: r, eigen data and an output other than the input");
    const unsigned nb = (unsigned)dg_blocks(K, a.n);
    hipLaunchKernelGGL((k_dg_apply<K, T, false, false, true>), dim3(nb), dim3(A::WG), 0, s, a, u, (T *)nullptr, skip,
                       (double *)nullptr, sm);
    FEM_HIP(hipGetLastError());
}

template <typename T>
static void dispatch_smooth(const DGArgs<T> &a, int k, const T *u, const DGSmooth<T> &sm, const int *skip,
                            hipStream_t s)
{
    switch (k) {
    case 2: launch_smooth<2, T>(a, u, sm, skip, s); break;
    case 3: launch_smooth<3, T>(a, u, sm, skip, s); break;
    case 4: launch_smooth<4, T>(a, u, sm, skip, s); break;
    case 5: launch_smooth<5, T>(a, u, sm, skip, s); break;
    default: throw Error(FEM_ERR_UNSUPPORTED, "DG apply: order must be 1..4");
    }
}

void dg_smooth_f64(const DGArgs<double> &a, int k, const double *u, const DGSmooth<double> &sm, const int *skip,
                   hipStream_t s)
{
    dispatch_smooth<double>(a, k, u, sm, skip, s);
}

void dg_smooth_f32(const DGArgs<float> &a, int k, const float *u, const DGSmooth<float> &sm, const int *skip,
                   hipStream_t s)
{
    dispatch_smooth<float>(a, k, u, sm, skip, s);
}

template <typename T>
static void dispatch_apply(const DGArgs<T> &a, int k, const T *u, T *w, const int *skip, hipStream_t s, double *dot)
{
    switch (k) {
    case 2: launch_apply<2, T>(a, u, w, skip, s, dot); break;
    case 3: launch_apply<3, T>(a, u, w, skip, s, dot); break;
    case 4: launch_apply<4, T>(a, u, w, skip, s, dot); break;
    case 5: launch_apply<5, T>(a, u, w, skip, s, dot); break;
    default: throw Error(FEM_ERR_UNSUPPORTED, "DG apply: order must be 1..4");
    }
}

void dg_apply_f64(const DGArgs<double> &a, int k, const double *u, double *w, const int *skip, hipStream_t s,
                  double *dot)
{
    dispatch_apply<double>(a, k, u, w, skip, s, dot);
}

void dg_apply_f32(const DGArgs<float> &a, int k, const float *u, float *w, const int *skip, hipStream_t s,
                  double *dot)
{
    dispatch_apply<float>(a, k, u, w, skip, s, dot);
}

int dg_apply_epb(int k)
{
    FEM_CHECK(k >= 2 && k <= 5, FEM_ERR_UNSUPPORTED, "DG: order must be 1..4");
    return k == 5 ? ApplyShape<5>::EPB : 64 / (k * k);
}

int64_t dg_blocks(int k, const int *n)
{
    // the slab traversal: rpb groups per x-row, one workgroup each
    const int epb = dg_apply_epb(k);
    return (int64_t)((n[0] + epb - 1) / epb) * n[1] * n[2];
}

int64_t dg_bj_blocks(int k, int64_t E)
{
    FEM_CHECK(k >= 2 && k <= 5, FEM_ERR_UNSUPPORTED, "DG: order must be 1..4");
    const int epb = 64 / (k * k);
    return (E + epb - 1) / epb;
}

void dg_apply_bf16(const DGArgs<float> &a, int k, const float *u, float *w, const int *skip, hipStream_t s)
{
    switch (k) {
    case 2: launch_apply<2, float, true>(a, u, w, skip, s); break;
    case 3: launch_apply<3, float, true>(a, u, w, skip, s); break;
    case 4: launch_apply<4, float, true>(a, u, w, skip, s); break;
    case 5: launch_apply<5, float, true>(a, u, w, skip, s); break;
    default: throw Error(FEM_ERR_UNSUPPORTED, "DG apply: order must be 1..4");
    }
}

// ---------------------------------------------------------------------------
// load vector F_i = int_K f phi_i (f at volume quadrature points); ADD:
// F_i += int_K f phi_i
// ---------------------------------------------------------------------------
template <int K, bool ADD>
__global__ void __launch_bounds__(DGShape<K>::THREADS)
    k_dg_rhs(const DGArgs<double> a, const double *__restrict__ fq, double *__restrict__ Fv)
{
    using S = DGShape<K>;
    constexpr int K2 = S::K2, K3 = S::K3, EPB = S::EPB;
    __shared__ double smem[EPB * K3];
    const Tab1D<double> &tb = *a.tab;
    const int tid = threadIdx.x, el = tid / K2, t = tid % K2, i = t / K, j = t % K;
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    const int64_t e_raw = (int64_t)blockIdx.x * EPB + el;
    const bool active = e_raw < E;
    const int64_t e = active ? e_raw : 0;
    double *A = smem + el * K3;
    const double detJ = a.h[0] * a.h[1] * a.h[2];
    double v[K], tmp[K];
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = tb.w[q] * tb.w[i] * tb.w[j] * detJ * fq[e * K3 + q + K * i + K2 * j];
    mvT<K>(tb.B, v, tmp);
#pragma unroll
    for (int q = 0; q < K; ++q) A[q * K2 + i * K + j] = tmp[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = A[i * K2 + q * K + j];
    mvT<K>(tb.B, v, tmp);
#pragma unroll
    for (int q = 0; q < K; ++q) A[i * K2 + q * K + j] = tmp[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = A[i * K2 + j * K + q];
    mvT<K>(tb.B, v, tmp);
    if (active)
#pragma unroll
        for (int c = 0; c < K; ++c) {
            double *f = Fv + e * K3 + i * K2 + j * K + c;
            *f = ADD ? *f + tmp[c] : tmp[c];
        }
}

void dg_rhs(const DGArgs<double> &a, int k, const double *fq, double *F, hipStream_t s, bool add)
{
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
#define RHS_CASE(KK)                                                                                   \
    case KK: {                                                                                         \
        unsigned nb = (unsigned)((E + DGShape<KK>::EPB - 1) / DGShape<KK>::EPB);                       \
        if (add)                                                                                       \
            hipLaunchKernelGGL((k_dg_rhs<KK, true>), dim3(nb), dim3(DGShape<KK>::THREADS), 0, s, a, fq, F); \
        else                                                                                           \
            hipLaunchKernelGGL((k_dg_rhs<KK, false>), dim3(nb), dim3(DGShape<KK>::THREADS), 0, s, a, fq, F); \
        break;                                                                                         \
    }
    switch (k) {
        RHS_CASE(2) RHS_CASE(3) RHS_CASE(4) RHS_CASE(5)
    default: throw Error(FEM_ERR_UNSUPPORTED, "DG rhs: order must be 1..4");
    }
#undef RHS_CASE
    FEM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// a DG function at the volume quadrature points, out[e][qx + K qy + K^2 qz]:
// out = base + coef u^pw (base may be null; pw = 1, coef = 1 gives u itself)
// ---------------------------------------------------------------------------
template <int K, typename T>
__global__ void __launch_bounds__(DGShape<K>::THREADS)
    k_dg_eval_qp(const DGArgs<T> a, const T *__restrict__ u, const T *__restrict__ base, T coef, int pw,
                 T *__restrict__ out)
{
    using S = DGShape<K>;
    constexpr int K2 = S::K2, K3 = S::K3, EPB = S::EPB;
    __shared__ T smem[EPB * K3];
    const Tab1D<T> &tb = *a.tab;
    const int tid = threadIdx.x, el = tid / K2, t = tid % K2, i = t / K, j = t % K;
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    const int64_t e_raw = (int64_t)blockIdx.x * EPB + el;
    const bool active = e_raw < E;
    const int64_t e = active ? e_raw : 0;
    T *A = smem + el * K3;
    T v[K], tmp[K];
    // z modes -> qz: lane (i = a, j = b)
#pragma unroll
    for (int c = 0; c < K; ++c) v[c] = u[e * K3 + i * K2 + j * K + c];
    mvB<K>(tb.B, v, tmp);
#pragma unroll
    for (int q = 0; q < K; ++q) A[i * K2 + j * K + q] = tmp[q];
    __syncthreads();
    // y modes -> qy: lane (i = a, j = qz)
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[i * K2 + b * K + j];
    mvB<K>(tb.B, v, tmp);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) A[i * K2 + q * K + j] = tmp[q];
    __syncthreads();
    // x modes -> qx: lane (i = qy, j = qz)
#pragma unroll
    for (int b = 0; b < K; ++b) v[b] = A[b * K2 + i * K + j];
    mvB<K>(tb.B, v, tmp);
    if (active)
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int64_t o = e * K3 + q + K * i + K2 * j;
            T r = coef;
            for (int p = 0; p < pw; ++p) r *= tmp[q];
            out[o] = base ? base[o] + r : r;
        }
}

template <typename T>
void dg_eval_qp(const DGArgs<T> &a, int k, const T *u, const T *base, T coef, int pw, T *out, hipStream_t s)
{
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    FEM_CHECK(pw >= 0, FEM_ERR_ARG, "quadrature-point evaluation: power >= 0");
#define EQ_CASE(KK)                                                                                        \
    case KK: {                                                                                             \
        unsigned nb = (unsigned)((E + DGShape<KK>::EPB - 1) / DGShape<KK>::EPB);                           \
        hipLaunchKernelGGL((k_dg_eval_qp<KK, T>), dim3(nb), dim3(DGShape<KK>::THREADS), 0, s, a, u, base, coef, \
                           pw, out);                                                                       \
        break;                                                                                             \
    }
    switch (k) {
        EQ_CASE(2) EQ_CASE(3) EQ_CASE(4) EQ_CASE(5)
    default: throw Error(FEM_ERR_UNSUPPORTED, "DG: order must be 1..4");
    }
#undef EQ_CASE
    FEM_HIP(hipGetLastError());
}

template void dg_eval_qp<double>(const DGArgs<double> &, int, const double *, const double *, double, int, double *,
                                 hipStream_t);
template void dg_eval_qp<float>(const DGArgs<float> &, int, const float *, const float *, float, int, float *,
                                hipStream_t);

// ---------------------------------------------------------------------------
// boundary data of the load vector: F_i += sum over the element's boundary
// faces of int (kappa beta/h + max(-b.n, 0)) g phi_i - kappa g dphi_i/dn,
// the g-dependent part of the Nitsche and inflow boundary terms
// (dune/fempy/test/parallelTest.py:127-129, pydemo/advectiondiffusion.py:43-49);
// g at the face quadrature points [faces_d][m^2].  One thread per DoF of a
// boundary element (the grid covers all elements; interior ones return).
// ---------------------------------------------------------------------------
__global__ void k_dg_rhs_bnd(const DGArgs<double> a, int K, const double *__restrict__ g0,
                             const double *__restrict__ g1, const double *__restrict__ g2, double *__restrict__ F)
{
    const int K2 = K * K, K3 = K2 * K;
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (gid >= E * K3) return;
    const int64_t e = gid / K3;
    const int li = (int)(gid % K3);
    const int ec[3] = { (int)(e % a.n[0]), (int)((e / a.n[0]) % a.n[1]), (int)(e / ((int64_t)a.n[0] * a.n[1])) };
    bool bnd = false;
    for (int d = 0; d < 3; ++d) bnd = bnd || ec[d] == 0 || ec[d] == a.n[d] - 1;
    if (!bnd) return;
    const Tab1D<double> &tb = *a.tab;
    const int ai[3] = { li / K2, (li / K) % K, li % K };
    const double *gf[3] = { g0, g1, g2 };
    double acc = 0;
    for (int d = 0; d < 3; ++d) {
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        for (int s = 0; s < 2; ++s) {
            if (ec[d] + (s ? 1 : -1) >= 0 && ec[d] + (s ? 1 : -1) < a.n[d]) continue;
            const double nrm = s ? 1.0 : -1.0;
            const double bn = a.b[d] * nrm, bneg = bn < 0 ? -bn : 0.0;
            int cc[3] = { ec[0], ec[1], ec[2] };
            cc[d] += s;
            const int64_t N0 = a.n[0] + (d == 0), N1 = a.n[1] + (d == 1);
            const int64_t fidx = cc[0] + N0 * (cc[1] + N1 * (int64_t)cc[2]);
            for (int q2 = 0; q2 < K; ++q2)
                for (int q1 = 0; q1 < K; ++q1) {
                    const int qf = q1 + K * q2;
                    const double tang = tb.B[q1][ai[t1]] * tb.B[q2][ai[t2]];
                    const double phi = tb.P[s][ai[d]] * tang;
                    const double dn = nrm * tb.dP[s][ai[d]] * tang / a.h[d];
                    const double W = tb.w[q1] * tb.w[q2] * a.h[t1] * a.h[t2];
                    const double kap = a.kappa_mode == FEM_KAPPA_QP ? a.kface[d][fidx * K2 + qf]
                                     : a.kappa_mode == FEM_KAPPA_CELL ? a.kcell[e] : a.kappa_const;
                    const double g = gf[d][fidx * K2 + qf];
                    acc += W * g * ((kap * a.beta / a.h[d] + bneg) * phi - kap * dn);
                }
        }
    }
    F[gid] += acc;
}

void dg_rhs_bnd(const DGArgs<double> &a, int k, const double *const g[3], double *F, hipStream_t s)
{
    const int64_t N = (int64_t)a.n[0] * a.n[1] * a.n[2] * k * k * k;
    hipLaunchKernelGGL(k_dg_rhs_bnd, dim3(grid_for(N, 256)), dim3(256), 0, s, a, k, g[0], g[1], g[2], F);
    FEM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// operator diagonal: one thread per DoF, direct sums over quadrature points
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T kappa_vol(const DGArgs<T> &a, int64_t e, int q, int K3)
{
    if (a.kappa_mode == FEM_KAPPA_QP) return a.kvol[e * K3 + q];
    if (a.kappa_mode == FEM_KAPPA_CELL) return a.kcell[e];
    return a.kappa_const;
}

__global__ void k_dg_diag(const DGArgs<double> a, int K, double *__restrict__ dout)
{
    const int K2 = K * K, K3 = K2 * K;
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (gid >= E * K3) return;
    const Tab1D<double> &tb = *a.tab;
    const int64_t e = gid / K3;
    const int li = (int)(gid % K3);
    const int ai[3] = { li / K2, (li / K) % K, li % K };
    const int ec[3] = { (int)(e % a.n[0]), (int)((e / a.n[0]) % a.n[1]), (int)(e / ((int64_t)a.n[0] * a.n[1])) };
    const int64_t stride[3] = { 1, (int64_t)a.n[0], (int64_t)a.n[0] * a.n[1] };
    const double detJ = a.h[0] * a.h[1] * a.h[2];
    double acc = 0;
    for (int qz = 0; qz < K; ++qz)
        for (int qy = 0; qy < K; ++qy)
            for (int qx = 0; qx < K; ++qx) {
                const int qq[3] = { qx, qy, qz };
                double phi = 1, gr[3];
                for (int d = 0; d < 3; ++d) phi *= tb.B[qq[d]][ai[d]];
                for (int d = 0; d < 3; ++d) {
                    double g = tb.Bd[qq[d]][ai[d]] / a.h[d];
                    for (int d2 = 0; d2 < 3; ++d2)
                        if (d2 != d) g *= tb.B[qq[d2]][ai[d2]];
                    gr[d] = g;
                }
                const double W = tb.w[qx] * tb.w[qy] * tb.w[qz] * detJ;
                const double kap = kappa_vol(a, e, qx + K * qy + K2 * qz, K3);
                acc += W * (kap * (gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]) -
                            (a.b[0] * gr[0] + a.b[1] * gr[1] + a.b[2] * gr[2]) * phi +
                            (a.c + (a.cvol ? a.cvol[e * K3 + qx + K * qy + K2 * qz] : 0.0)) * phi * phi);
            }
    for (int d = 0; d < 3; ++d) {
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        for (int s = 0; s < 2; ++s) {
            const double nrm = s ? 1.0 : -1.0;
            const int nbc = ec[d] + (s ? 1 : -1);
            const bool interior = nbc >= 0 && nbc < a.n[d];
            const double bn = a.b[d] * nrm, bp = bn > 0 ? bn : 0.0;
            for (int q2 = 0; q2 < K; ++q2)
                for (int q1 = 0; q1 < K; ++q1) {
                    const double tang = tb.B[q1][ai[t1]] * tb.B[q2][ai[t2]];
                    const double phi = tb.P[s][ai[d]] * tang;
                    const double dn = nrm * tb.dP[s][ai[d]] * tang / a.h[d];
                    const double Wf = tb.w[q1] * tb.w[q2] * a.h[t1] * a.h[t2];
                    double kin, kp;
                    const int idx = q1 + K * q2;
                    if (a.kappa_mode == FEM_KAPPA_QP) {
                        int cc[3] = { ec[0], ec[1], ec[2] };
                        cc[d] += s;
                        const int64_t N0 = a.n[0] + (d == 0), N1 = a.n[1] + (d == 1);
                        const int64_t fidx = cc[0] + N0 * (cc[1] + N1 * (int64_t)cc[2]);
                        kin = kp = a.kface[d][fidx * K2 + idx];
                    } else if (a.kappa_mode == FEM_KAPPA_CELL) {
                        kin = a.kcell[e];
                        const double kout = interior ? a.kcell[e + (s ? stride[d] : -stride[d])] : kin;
                        kp = 0.5 * (kin + kout);
                    } else {
                        kin = kp = a.kappa_const;
                    }
                    if (interior)
                        acc += Wf * (kp * a.beta / a.h[d] * phi * phi - kin * phi * dn + bp * phi * phi);
                    else
                        acc += Wf * (kin * a.beta / a.h[d] * phi * phi - 2.0 * kin * phi * dn + bp * phi * phi);
                }
        }
    }
    dout[gid] = acc;
}

void dg_diagonal(const DGArgs<double> &a, int k, double *d, hipStream_t s)
{
    const int64_t N = (int64_t)a.n[0] * a.n[1] * a.n[2] * k * k * k;
    hipLaunchKernelGGL(k_dg_diag, dim3(grid_for(N, 256)), dim3(256), 0, s, a, k, d);
    FEM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// exact element-block inverse for cellwise kappa without advection
// ---------------------------------------------------------------------------
// cyclic Jacobi eigenvalue iteration for a symmetric K x K matrix
__device__ void sym_eig(int K, double (*A)[KMAX], double (*Q)[KMAX])
{
    for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) Q[r][c] = (r == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0, tot = 0;
        for (int r = 0; r < K; ++r)
            for (int c = 0; c < K; ++c) {
                tot += A[r][c] * A[r][c];
                if (r != c) off += A[r][c] * A[r][c];
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < K - 1; ++p)
            for (int q = p + 1; q < K; ++q) {
                if (A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(tt * tt + 1.0), sn = tt * cs;
                for (int k = 0; k < K; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = cs * akp - sn * akq;
                    A[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < K; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = cs * apk - sn * aqk;
                    A[q][k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < K; ++k) {
                    const double qkp = Q[k][p], qkq = Q[k][q];
                    Q[k][p] = cs * qkp - sn * qkq;
                    Q[k][q] = sn * qkp + cs * qkq;
                }
            }
    }
}

template <typename TO>
__global__ void k_dg_block_setup(const DGArgs<double> a, int K, TO *__restrict__ Qout, TO *__restrict__ lam,
                                 bool bound)
{
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (gid >= E * 3) return;
    const int64_t e = gid / 3;
    const int d = (int)(gid % 3);
    const Tab1D<double> &tb = *a.tab;
    const int ec[3] = { (int)(e % a.n[0]), (int)((e / a.n[0]) % a.n[1]), (int)(e / ((int64_t)a.n[0] * a.n[1])) };
    const int64_t stride[3] = { 1, (int64_t)a.n[0], (int64_t)a.n[0] * a.n[1] };
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const double cd = a.h[t1] * a.h[t2];
    double kin;
    if (a.kappa_mode == FEM_KAPPA_CELL) kin = a.kcell[e];
    else if (a.kappa_mode == FEM_KAPPA_QP) {
        // element mean of the pointwise coefficient (approximate block)
        double s = 0;
        for (int q = 0; q < K * K * K; ++q) s += a.kvol[e * K * K * K + q];
        kin = s / (K * K * K);
    } else kin = a.kappa_const;
    double M[KMAX][KMAX], Q[KMAX][KMAX];
    for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) M[r][c] = cd * kin / a.h[d] * tb.S[r][c];
    for (int s = 0; s < 2; ++s) {
        const double nrm = s ? 1.0 : -1.0;
        const int nbc = ec[d] + (s ? 1 : -1);
        const bool interior = nbc >= 0 && nbc < a.n[d];
        double kp;
        if (interior) {
            double kout;
            if (a.kappa_mode == FEM_KAPPA_CELL) kout = a.kcell[e + (s ? stride[d] : -stride[d])];
            else kout = kin;
            kp = 0.5 * (kin + kout);
        } else kp = kin;
        const double symw = interior ? 0.5 : 1.0;
        for (int r = 0; r < K; ++r)
            for (int c = 0; c < K; ++c)
                M[r][c] += cd * (kp * a.beta / a.h[d] * tb.P[s][r] * tb.P[s][c] -
                                 symw * kin * nrm / a.h[d] * (tb.dP[s][c] * tb.P[s][r] + tb.P[s][c] * tb.dP[s][r]));
    }
    sym_eig(K, M, Q);
    // the reaction c + cvol (orthonormal basis: mass block |K| I) shifts
    // every eigenvalue sum; taken into the x direction, with the element's
    // weighted mean of cvol (exact for an element-constant reaction), or
    // (bound) its largest value: then the block bounds the element's true
    // diagonal block from above (int cvol v^2 <= max cvol int v^2 under the
    // positive quadrature weights), which keeps a damped sweep with it
    // convergent however cvol varies inside the element
    double shift = 0;
    if (d == 0) {
        double cm = 0, cmax = -INFINITY;
        if (a.cvol)
            for (int qz = 0; qz < K; ++qz)
                for (int qy = 0; qy < K; ++qy)
                    for (int qx = 0; qx < K; ++qx) {
                        const double v = a.cvol[e * K * K * K + qx + K * (qy + K * qz)];
                        cm += tb.w[qx] * tb.w[qy] * tb.w[qz] * v;
                        cmax = fmax(cmax, v);
                    }
        shift = (a.c + (a.cvol && bound ? cmax : cm)) * a.h[0] * a.h[1] * a.h[2];
    }
    for (int r = 0; r < K; ++r) {
        lam[(e * 3 + d) * K + r] = (TO)(M[r][r] + shift);
        for (int c = 0; c < K; ++c) Qout[((e * 3 + d) * K + r) * K + c] = (TO)Q[r][c];
    }
}

template <typename TO>
void dg_block_setup(const DGArgs<double> &a, int k, TO *Q, TO *lam, hipStream_t s, bool bound)
{
    const int64_t E = (int64_t)a.n[0] * a.n[1] * a.n[2];
    hipLaunchKernelGGL(k_dg_block_setup<TO>, dim3(grid_for(E * 3, 128)), dim3(128), 0, s, a, k, Q, lam, bound);
    FEM_HIP(hipGetLastError());
}
template void dg_block_setup<double>(const DGArgs<double> &, int, double *, double *, hipStream_t, bool);
template void dg_block_setup<float>(const DGArgs<double> &, int, float *, float *, hipStream_t, bool);

// one wavefront per block; lanes beyond EPB*K2 shadow element 0 (identical
// LDS writes, no global writes)
template <int K, typename T>
__global__ void __launch_bounds__(64)
    k_dg_block_apply(int64_t E, const T *__restrict__ Q, const T *__restrict__ lam, const T *__restrict__ x,
                     T *__restrict__ y, const int *__restrict__ skip, double *__restrict__ dot)
{
    using S = DGShape<K>;
    constexpr int K2 = S::K2, K3 = S::K3, EPB = S::EPB;
    if (skip && *skip) return;
    __shared__ T smem[EPB * K3];
    __shared__ T qsm[EPB * (3 * K2 + 3 * K)];
    const int tid = threadIdx.x, el0 = tid / K2, t = tid % K2, i = t / K, j = t % K;
    const bool wr = el0 < EPB;
    const int el = wr ? el0 : 0;
    const int64_t e_raw = (int64_t)blockIdx.x * EPB + el;
    const bool active = wr && e_raw < E;
    const int64_t e = e_raw < E ? e_raw : 0;
    T v[K];
#pragma unroll
    for (int c = 0; c < K; ++c) v[c] = x[e * K3 + i * K2 + j * K + c];
    bj_element<K>(Q + e * 3 * K2, lam + e * 3 * K, smem + el * K3, qsm + el * (3 * K2 + 3 * K), i, j, v);
    if (active)
#pragma unroll
        for (int r = 0; r < K; ++r) y[e * K3 + i * K2 + j * K + r] = v[r];
    if (dot) {
        // this workgroup's <x, y> and <x, x> (fixed summation order)
        double a1 = 0, a2 = 0;
        if (active)
#pragma unroll
            for (int r = 0; r < K; ++r) {
                const double xv = (double)x[e * K3 + i * K2 + j * K + r];
                a1 += xv * (double)v[r];
                a2 += xv * xv;
            }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a1 += __shfl_xor(a1, off, 64);
            a2 += __shfl_xor(a2, off, 64);
        }
        if (tid == 0) {
            dot[blockIdx.x] = a1;
            dot[gridDim.x + blockIdx.x] = a2;
        }
    }
}

template <typename T>
void dg_block_apply(int k, int64_t E, const T *Q, const T *lam, const T *x, T *y, const int *skip, hipStream_t s,
                    double *dot)
{
#define BA_CASE(KK)                                                                                         \
    case KK: {                                                                                              \
        unsigned nb = (unsigned)((E + DGShape<KK>::EPB - 1) / DGShape<KK>::EPB);                            \
        hipLaunchKernelGGL((k_dg_block_apply<KK, T>), dim3(nb), dim3(64), 0, s, E, Q, lam, x, y, skip, dot); \
        break;                                                                                              \
    }
    switch (k) {
        BA_CASE(2) BA_CASE(3) BA_CASE(4) BA_CASE(5)
    default: throw Error(FEM_ERR_UNSUPPORTED, "block Jacobi: order must be 1..4");
    }
#undef BA_CASE
    FEM_HIP(hipGetLastError());
}

template void dg_block_apply<double>(int, int64_t, const double *, const double *, const double *, double *, const int *,
                                     hipStream_t, double *);
template void dg_block_apply<float>(int, int64_t, const float *, const float *, const float *, float *, const int *,
                                    hipStream_t, double *);

}  // namespace femamd
