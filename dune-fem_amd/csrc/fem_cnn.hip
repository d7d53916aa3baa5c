// 3D CNN kernels: 3x3x3 dilated convolutions ("same" zero padding, no
// biases, ReLU between layers), channel-last activations.
//
//   hidden layers (32 -> 32): implicit GEMM on v_mfma_f32_16x16x32_bf16,
//     BF16 operands, FP32 accumulation.  Per wave: 16 voxels x 32 output
//     channels; per tap one K = 32 step (the 32 input channels).
//     Operand layout (checked on the device): A[i][k] in lane i + 16*(k/8),
//     B[k][j] in lane j + 16*(k/8), D[i][j] in lane j + 16*(i/4), reg i%4.
//   input layer (2 -> 32) and output layer (32 -> 1): FP32 VALU on the same
//     BF16-rounded weights and activations.
//   weight gradients: per (tap, voxel chunk) partial GEMMs (MFMA for the hidden
//     layers), summed over chunks in a fixed order.
// No reference counterpart.
#include <cmath>
#include <vector>

#include "fem_cnn.hpp"

namespace femamd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int CH = CNN_HID;
constexpr int WG_CHUNK = 1024;   // voxels per weight-gradient partial

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ u16 f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ __bf16 f2b(float x)
{
    const u16 h = f2bf(x);
    return __builtin_bit_cast(__bf16, h);
}

struct Grid {
    int nx, ny, nz, V;
};

// voxel v shifted by sign * dil * offset(t); -1 outside the grid
__device__ __forceinline__ int nbr(const Grid &g, int v, int t, int dil, int sign)
{
    const int x = v % g.nx, y = (v / g.nx) % g.ny, z = v / (g.nx * g.ny);
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int X = x + sign * dil * dx, Y = y + sign * dil * dy, Z = z + sign * dil * dz;
    if (X < 0 || Y < 0 || Z < 0 || X >= g.nx || Y >= g.ny || Z >= g.nz) return -1;
    return X + g.nx * (Y + g.ny * Z);
}

// a voxel with its coordinates, stepped along the voxel order without
// divisions (the weight-gradient kernels walk runs of consecutive voxels)
struct Vox {
    int v, x, y, z;
};
__device__ __forceinline__ Vox vox(const Grid &g, int v)
{
    return Vox{ v, v % g.nx, (v / g.nx) % g.ny, v / (g.nx * g.ny) };
}
__device__ __forceinline__ void vox_next(const Grid &g, Vox &p)
{
    ++p.v;
    if (++p.x == g.nx) {
        p.x = 0;
        if (++p.y == g.ny) {
            p.y = 0;
            ++p.z;
        }
    }
}
__device__ __forceinline__ int nbr(const Grid &g, const Vox &p, int t, int dil, int sign)
{
    const int sx = sign * dil * (t % 3 - 1), sy = sign * dil * ((t / 3) % 3 - 1), sz = sign * dil * (t / 9 - 1);
    const int X = p.x + sx, Y = p.y + sy, Z = p.z + sz;
    if (X < 0 || Y < 0 || Z < 0 || X >= g.nx || Y >= g.ny || Z >= g.nz) return -1;
    return p.v + sx + g.nx * (sy + g.ny * sz);
}

// ---- input layer: 2 -> 32, ReLU --------------------------------------------------
__global__ void __launch_bounds__(256) k_in_fwd(Grid g, int dil, const float *__restrict__ x0,
                                                const float *__restrict__ wq, u16 *__restrict__ a1,
                                                const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ float W[27 * 2 * CH];
    for (int k = threadIdx.x; k < 27 * 2 * CH; k += blockDim.x) W[k] = wq[k];
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    float acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.f;
    for (int t = 0; t < 27; ++t) {
        const int w = nbr(g, v, t, dil, 1);
        if (w < 0) continue;
        const float xa = x0[2 * w], xb = x0[2 * w + 1];
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = fmaf(W[(t * 2 + 0) * CH + c], xa, fmaf(W[(t * 2 + 1) * CH + c], xb, acc[c]));
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) a1[(size_t)v * CH + c] = f2bf(fmaxf(acc[c], 0.f));
}

// ---- output layer: 32 -> 1, linear -------------------------------------------------
__global__ void __launch_bounds__(256) k_out_fwd(Grid g, int dil, const u16 *__restrict__ a3,
                                                 const float *__restrict__ wq, float *__restrict__ y,
                                                 const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ float W[27 * CH];
    for (int k = threadIdx.x; k < 27 * CH; k += blockDim.x) W[k] = wq[k];
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    float acc = 0.f;
    for (int t = 0; t < 27; ++t) {
        const int w = nbr(g, v, t, dil, 1);
        if (w < 0) continue;
        const uint4 *p = reinterpret_cast<const uint4 *>(a3 + (size_t)w * CH);
#pragma unroll
        for (int q = 0; q < CH / 8; ++q) {
            const uint4 pk = p[q];
            const unsigned wd[4] = { pk.x, pk.y, pk.z, pk.w };
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int c = q * 8 + 2 * h;
                acc = fmaf(W[t * CH + c], __uint_as_float(wd[h] << 16), acc);
                acc = fmaf(W[t * CH + c + 1], __uint_as_float(wd[h] & 0xffff0000u), acc);
            }
        }
    }
    y[v] = acc;
}

// ---- hidden layer forward: 32 -> 32, ReLU, BF16 MFMA -----------------------------
__global__ void __launch_bounds__(256) k_hid_fwd(Grid g, int dil, const u16 *__restrict__ ain,
                                                 const u16 *__restrict__ wf /*[27][co][ci]*/, u16 *__restrict__ aout,
                                                 const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int v0 = (blockIdx.x * 4 + wave) * 16;
    if (v0 >= g.V) return;   // uniform per wave
    const int i = lane & 15, kb = lane >> 4;
    const int v = v0 + i;
    f32x4_t acc0 = { 0.f, 0.f, 0.f, 0.f }, acc1 = { 0.f, 0.f, 0.f, 0.f };
    for (int t = 0; t < 27; ++t) {
        const int w = v < g.V ? nbr(g, v, t, dil, 1) : -1;
        bf16x8_t A;
        if (w >= 0) A = *reinterpret_cast<const bf16x8_t *>(ain + (size_t)w * CH + 8 * kb);
        else {
#pragma unroll
            for (int r = 0; r < 8; ++r) A[r] = __builtin_bit_cast(__bf16, (u16)0);
        }
        const bf16x8_t B0 = *reinterpret_cast<const bf16x8_t *>(wf + ((size_t)t * CH + i) * CH + 8 * kb);
        const bf16x8_t B1 = *reinterpret_cast<const bf16x8_t *>(wf + ((size_t)t * CH + 16 + i) * CH + 8 * kb);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int vr = v0 + 4 * kb + r;
        if (vr < g.V) {
            aout[(size_t)vr * CH + i] = f2bf(fmaxf(acc0[r], 0.f));
            aout[(size_t)vr * CH + 16 + i] = f2bf(fmaxf(acc1[r], 0.f));
        }
    }
}

// ---- hidden layer backward data: gin = (W^T * gout shifted) . (ain > 0) ------------
__global__ void __launch_bounds__(256) k_hid_bwd(Grid g, int dil, const float *__restrict__ gout,
                                                 const u16 *__restrict__ wb /*[27][ci][co]*/,
                                                 const u16 *__restrict__ ain, float *__restrict__ gin)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int v0 = (blockIdx.x * 4 + wave) * 16;
    if (v0 >= g.V) return;
    const int i = lane & 15, kb = lane >> 4;
    const int v = v0 + i;
    f32x4_t acc0 = { 0.f, 0.f, 0.f, 0.f }, acc1 = { 0.f, 0.f, 0.f, 0.f };
    for (int t = 0; t < 27; ++t) {
        const int w = v < g.V ? nbr(g, v, t, dil, -1) : -1;
        bf16x8_t A;
        if (w >= 0) {
            const float4 *p = reinterpret_cast<const float4 *>(gout + (size_t)w * CH + 8 * kb);
            const float4 lo = p[0], hi = p[1];
            A[0] = f2b(lo.x); A[1] = f2b(lo.y); A[2] = f2b(lo.z); A[3] = f2b(lo.w);
            A[4] = f2b(hi.x); A[5] = f2b(hi.y); A[6] = f2b(hi.z); A[7] = f2b(hi.w);
        } else {
#pragma unroll
            for (int r = 0; r < 8; ++r) A[r] = __builtin_bit_cast(__bf16, (u16)0);
        }
        const bf16x8_t B0 = *reinterpret_cast<const bf16x8_t *>(wb + ((size_t)t * CH + i) * CH + 8 * kb);
        const bf16x8_t B1 = *reinterpret_cast<const bf16x8_t *>(wb + ((size_t)t * CH + 16 + i) * CH + 8 * kb);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int vr = v0 + 4 * kb + r;
        if (vr < g.V) {
            const size_t o0 = (size_t)vr * CH + i, o1 = o0 + 16;
            gin[o0] = bf2f(ain[o0]) > 0.f ? acc0[r] : 0.f;
            gin[o1] = bf2f(ain[o1]) > 0.f ? acc1[r] : 0.f;
        }
    }
}

// ---- output layer backward data: g3 = (W4 * g4 shifted) . (a3 > 0) ----------------
__global__ void __launch_bounds__(256) k_out_bwd(Grid g, int dil, const float *__restrict__ g4,
                                                 const float *__restrict__ wq, const u16 *__restrict__ a3,
                                                 float *__restrict__ g3)
{
    __shared__ float W[27 * CH];
    for (int k = threadIdx.x; k < 27 * CH; k += blockDim.x) W[k] = wq[k];
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    float acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.f;
    for (int t = 0; t < 27; ++t) {
        const int w = nbr(g, v, t, dil, -1);
        if (w < 0) continue;
        const float gv = bf2f(f2bf(g4[w]));
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = fmaf(W[t * CH + c], gv, acc[c]);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) g3[(size_t)v * CH + c] = bf2f(a3[(size_t)v * CH + c]) > 0.f ? acc[c] : 0.f;
}

// ---- loss: g4 = 2 (y - t) / V, per-block partial sums of (y - t)^2 ----------------
__global__ void __launch_bounds__(256) k_mse(int V, const float *__restrict__ y, const float *__restrict__ t,
                                             float *__restrict__ g4, double *__restrict__ part, float weight)
{
    __shared__ double s[256];
    double acc = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const float d = y[v] - t[v];
        g4[v] = 2.f * d * weight / (float)V;
        acc += (double)d * d * weight;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// ---- weight gradients ----------------------------------------------------------------
// hidden layers: block (tap, chunk), 4 waves = (ci tile, co tile); K =
// voxels.  Per step of 32 voxels the block stages the tap-shifted input rows
// (32 channels, BF16) and the output-gradient rows (rounded to BF16) in LDS,
// transposed to [channel][voxel], so that a lane's MFMA operand (8
// consecutive voxels of one channel) is one 16-byte LDS read; the four
// wavefronts take their 16 x 16 tiles from the shared stage.  Each staged
// row is one coalesced 8- and 16-byte load per thread instead of 16 scalar
// gathers per lane and step.
constexpr int WGS = 40;   // LDS row stride in BF16: 32 voxels + 8 (16-byte aligned rows)
__global__ void __launch_bounds__(256) k_hid_wgrad(Grid g, int dil, const u16 *__restrict__ ain,
                                                   const float *__restrict__ gout, float *__restrict__ wpart,
                                                   int nchunks)
{
    __shared__ __attribute__((aligned(16))) u16 At[CH * WGS];
    __shared__ __attribute__((aligned(16))) u16 Bt[CH * WGS];
    const int t = blockIdx.x, chunk = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tci = wave >> 1, tco = wave & 1;
    const int i = lane & 15, kb = lane >> 4;
    // staging role of the thread: voxel r of the step, channels 4p..4p+3
    const int r = threadIdx.x >> 3, p = threadIdx.x & 7;
    const int c0 = chunk * WG_CHUNK;
    f32x4_t acc = { 0.f, 0.f, 0.f, 0.f };
    for (int k0 = c0; k0 < c0 + WG_CHUNK; k0 += 32) {
        const int v = k0 + r;
        u16 a4[4] = { 0, 0, 0, 0 }, b4[4] = { 0, 0, 0, 0 };
        if (v < g.V) {
            const int w = nbr(g, v, t, dil, 1);
            if (w >= 0) {
                const uint2 q = *reinterpret_cast<const uint2 *>(ain + (size_t)w * CH + 4 * p);
                a4[0] = (u16)(q.x & 0xffffu);
                a4[1] = (u16)(q.x >> 16);
                a4[2] = (u16)(q.y & 0xffffu);
                a4[3] = (u16)(q.y >> 16);
            }
            const float4 gv = *reinterpret_cast<const float4 *>(gout + (size_t)v * CH + 4 * p);
            b4[0] = f2bf(gv.x);
            b4[1] = f2bf(gv.y);
            b4[2] = f2bf(gv.z);
            b4[3] = f2bf(gv.w);
        }
        __syncthreads();   // the previous step's operand reads are done
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            At[(4 * p + j) * WGS + r] = a4[j];
            Bt[(4 * p + j) * WGS + r] = b4[j];
        }
        __syncthreads();
        const bf16x8_t A = *reinterpret_cast<const bf16x8_t *>(At + (16 * tci + i) * WGS + 8 * kb);
        const bf16x8_t B = *reinterpret_cast<const bf16x8_t *>(Bt + (16 * tco + i) * WGS + 8 * kb);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ci = 16 * tci + 4 * kb + r, co = 16 * tco + i;
        wpart[(((size_t)t * nchunks + chunk) * CH + ci) * CH + co] = acc[r];
    }
}

// thin layers (CI x CO <= 64 weights per tap): block (tap, chunk), the 256
// threads share the chunk's voxels (a quarter of them each, strided), every
// thread accumulating all CI x CO products of its voxels; then a fixed-order
// reduction, by shuffles within each wavefront and over the four wavefronts
// in LDS.  Input FP32 (layer 1) or BF16 (layer 4); the output gradient is
// rounded to BF16 as the MFMA layers see it.
template <bool BF, int CI, int CO>
__global__ void __launch_bounds__(256) k_thin_wgrad(Grid g, int dil, const void *__restrict__ ain_,
                                                    const float *__restrict__ gout, float *__restrict__ wpart,
                                                    int nchunks)
{
    constexpr int NW = CI * CO;
    static_assert(NW <= 64, "thin layer: at most 64 weights per tap");
    const int t = blockIdx.x, chunk = blockIdx.y;
    const int c0 = chunk * WG_CHUNK, c1 = min(c0 + WG_CHUNK, g.V);
    float acc[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) acc[k] = 0.f;
    for (int v = c0 + (int)threadIdx.x; v < c1; v += 256) {
        const int w = nbr(g, v, t, dil, 1);
        if (w < 0) continue;
        float a[CI], gg[CO];
#pragma unroll
        for (int ci = 0; ci < CI; ++ci)
            a[ci] = BF ? bf2f(((const u16 *)ain_)[(size_t)w * CI + ci]) : ((const float *)ain_)[(size_t)w * CI + ci];
#pragma unroll
        for (int co = 0; co < CO; ++co) gg[co] = bf2f(f2bf(gout[(size_t)v * CO + co]));
#pragma unroll
        for (int ci = 0; ci < CI; ++ci)
#pragma unroll
            for (int co = 0; co < CO; ++co) acc[ci * CO + co] = fmaf(a[ci], gg[co], acc[ci * CO + co]);
    }
    __shared__ float red[4][NW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        float s = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NW; k += 256)
        wpart[((size_t)t * nchunks + chunk) * NW + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
}

// grad[t][ci][co] += sum over chunks (fixed order)
__global__ void k_wgrad_reduce(const float *__restrict__ wpart, int nchunks, int cc, float *__restrict__ grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;   // over 27 * cc
    if (k >= 27 * cc) return;
    const int t = k / cc, r = k % cc;
    float s = 0.f;
    for (int c = 0; c < nchunks; ++c) s += wpart[((size_t)t * nchunks + c) * cc + r];
    grad[k] += s;
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float scale, float b1t, float b2t)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gg = g[k] * scale;
    const float mm = 0.9f * m[k] + 0.1f * gg, vv = 0.999f * v[k] + 0.001f * gg * gg;
    m[k] = mm;
    v[k] = vv;
    w[k] -= lr * (mm / (1.f - b1t)) / (sqrtf(vv / (1.f - b2t)) + 1e-8f);
}

// master -> BF16-rounded FP32 copy and the MFMA operand layouts
__global__ void k_refresh(const float *__restrict__ wm, float *__restrict__ wq, u16 *__restrict__ wf,
                          u16 *__restrict__ wb, int n, int h2, int h3, int h2end)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const u16 b = f2bf(wm[k]);
    wq[k] = bf2f(b);
    // hidden layers occupy [h2, h2end): two blocks of 27*32*32
    if (k >= h2 && k < h2end) {
        const int l = k < h3 ? 0 : 1;
        const int r = k - (l ? h3 : h2);
        const int t = r / (CH * CH), ci = (r / CH) % CH, co = r % CH;
        wb[(size_t)l * 27 * CH * CH + r] = b;
        wf[(size_t)l * 27 * CH * CH + ((size_t)t * CH + co) * CH + ci] = b;
    }
}

// ---------------------------------------------------------------------------------
void CNN::init(Context *c, const int nn[3], const int d[CNN_LAYERS])
{
    ctx = c;
    for (int i = 0; i < 3; ++i) n[i] = nn[i];
    for (int l = 0; l < CNN_LAYERS; ++l) dil[l] = d[l];
    V = (int64_t)n[0] * n[1] * n[2];
    FEM_CHECK(V > 0 && V < (1ll << 31) / CH, FEM_ERR_ARG, "cnn: grid size");
    woff[0] = 0;
    for (int l = 0; l < CNN_LAYERS; ++l) woff[l + 1] = woff[l] + (size_t)27 * ch[l] * ch[l + 1];
    nparams = woff[CNN_LAYERS];
    nchunks = (int)((V + WG_CHUNK - 1) / WG_CHUNK);
    for (DevBuf *b : { &wmaster, &wq, &grad, &adam_m, &adam_v }) b->alloc(sizeof(float) * nparams);
    wfwd.alloc(sizeof(u16) * 2 * 27 * CH * CH);
    wbwd.alloc(sizeof(u16) * 2 * 27 * CH * CH);
    wpart.alloc(sizeof(float) * (size_t)27 * nchunks * CH * CH);
    x0.alloc(sizeof(float) * V * 2);
    for (DevBuf *b : { &a1, &a2, &a3 }) b->alloc(sizeof(u16) * V * CH);
    for (DevBuf *b : { &g1, &g2, &g3 }) b->alloc(sizeof(float) * V * CH);
    y.alloc(sizeof(float) * V);
    g4.alloc(sizeof(float) * V);
    lossp.alloc(sizeof(double) * 256);
    FEM_HIP(hipMemset(adam_m.p, 0, adam_m.bytes));
    FEM_HIP(hipMemset(adam_v.p, 0, adam_v.bytes));
    FEM_HIP(hipMemset(grad.p, 0, grad.bytes));
    // deterministic He-style initialisation (host LCG)
    std::vector<float> w(nparams);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const float sc = std::sqrt(2.0f / (27.0f * ch[l])) * (l == CNN_LAYERS - 1 ? 0.1f : 1.0f);
        for (size_t k = woff[l]; k < woff[l + 1]; ++k) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            const double u1 = ((st >> 11) + 1.0) / 9007199254740993.0;
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            const double u2 = (st >> 11) / 9007199254740992.0;
            w[k] = (float)(sc * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
    set_weights(w.data());
}

void CNN::set_weights(const float *host)
{
    FEM_HIP(hipMemcpy(wmaster.p, host, sizeof(float) * nparams, hipMemcpyHostToDevice));
    refresh();
    FEM_HIP(hipStreamSynchronize(ctx->stream));
}

void CNN::get_weights(float *host) const
{
    FEM_HIP(hipStreamSynchronize(ctx->stream));
    FEM_HIP(hipMemcpy(host, wmaster.p, sizeof(float) * nparams, hipMemcpyDeviceToHost));
}

void CNN::refresh()
{
    const int nn = (int)nparams;
    hipLaunchKernelGGL(k_refresh, dim3((nn + 255) / 256), dim3(256), 0, ctx->stream, wmaster.as<float>(), wq.as<float>(),
                       wfwd.as<u16>(), wbwd.as<u16>(), nn, (int)woff[1], (int)woff[2], (int)woff[3]);
    FEM_HIP(hipGetLastError());
}

void CNN::forward(const float *x, float *out, hipStream_t s, const int *skip)
{
    Grid g{ n[0], n[1], n[2], (int)V };
    const unsigned gv = (unsigned)((V + 255) / 256), gw = (unsigned)((V + 63) / 64);
    hipLaunchKernelGGL(k_in_fwd, dim3(gv), dim3(256), 0, s, g, dil[0], x, wq.as<float>() + woff[0], a1.as<u16>(), skip);
    hipLaunchKernelGGL(k_hid_fwd, dim3(gw), dim3(256), 0, s, g, dil[1], a1.as<u16>(), wfwd.as<u16>(), a2.as<u16>(), skip);
    hipLaunchKernelGGL(k_hid_fwd, dim3(gw), dim3(256), 0, s, g, dil[2], a2.as<u16>(), wfwd.as<u16>() + 27 * CH * CH,
                       a3.as<u16>(), skip);
    hipLaunchKernelGGL(k_out_fwd, dim3(gv), dim3(256), 0, s, g, dil[3], a3.as<u16>(), wq.as<float>() + woff[3], out, skip);
    FEM_HIP(hipGetLastError());
}

void CNN::zero_grad(hipStream_t s)
{
    FEM_HIP(hipMemsetAsync(grad.p, 0, grad.bytes, s));
}

void CNN::backward_accumulate(const float *t, hipStream_t s, float weight)
{
    // expects forward(x0, y) to have run on the same stream
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, s, (int)V, y.as<float>(), t, g4.as<float>(), lossp.as<double>(),
                       weight);
    backward_from_g4(s);
}

void CNN::backward_from_g4(hipStream_t s)
{
    Grid g{ n[0], n[1], n[2], (int)V };
    const unsigned gv = (unsigned)((V + 255) / 256), gw = (unsigned)((V + 63) / 64);
    // layer 4 (32 -> 1)
    hipLaunchKernelGGL((k_thin_wgrad<true, CH, 1>), dim3(27, nchunks), dim3(256), 0, s, g, dil[3], (const void *)a3.p,
                       g4.as<float>(), wpart.as<float>(), nchunks);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((27 * CH + 255) / 256), dim3(256), 0, s, wpart.as<float>(), nchunks, CH,
                       grad.as<float>() + woff[3]);
    hipLaunchKernelGGL(k_out_bwd, dim3(gv), dim3(256), 0, s, g, dil[3], g4.as<float>(), wq.as<float>() + woff[3],
                       a3.as<u16>(), g3.as<float>());
    // layer 3 (32 -> 32)
    hipLaunchKernelGGL(k_hid_wgrad, dim3(27, nchunks), dim3(256), 0, s, g, dil[2], a2.as<u16>(), g3.as<float>(),
                       wpart.as<float>(), nchunks);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((27 * CH * CH + 255) / 256), dim3(256), 0, s, wpart.as<float>(), nchunks,
                       CH * CH, grad.as<float>() + woff[2]);
    hipLaunchKernelGGL(k_hid_bwd, dim3(gw), dim3(256), 0, s, g, dil[2], g3.as<float>(), wbwd.as<u16>() + 27 * CH * CH,
                       a2.as<u16>(), g2.as<float>());
    // layer 2 (32 -> 32)
    hipLaunchKernelGGL(k_hid_wgrad, dim3(27, nchunks), dim3(256), 0, s, g, dil[1], a1.as<u16>(), g2.as<float>(),
                       wpart.as<float>(), nchunks);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((27 * CH * CH + 255) / 256), dim3(256), 0, s, wpart.as<float>(), nchunks,
                       CH * CH, grad.as<float>() + woff[1]);
    hipLaunchKernelGGL(k_hid_bwd, dim3(gw), dim3(256), 0, s, g, dil[1], g2.as<float>(), wbwd.as<u16>(), a1.as<u16>(),
                       g1.as<float>());
    // layer 1 (2 -> 32)
    hipLaunchKernelGGL((k_thin_wgrad<false, 2, CH>), dim3(27, nchunks), dim3(256), 0, s, g, dil[0], (const void *)x0.p,
                       g1.as<float>(), wpart.as<float>(), nchunks);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((27 * 2 * CH + 255) / 256), dim3(256), 0, s, wpart.as<float>(), nchunks,
                       2 * CH, grad.as<float>() + woff[0]);
    FEM_HIP(hipGetLastError());
}

// out = A_c c (7-point stencil on the grid)
__global__ void k_coarse_stencil_apply(Grid g, const float *__restrict__ st, const float *__restrict__ c,
                                       float *__restrict__ out)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    const int x = v % g.nx, y = (v / g.nx) % g.ny, z = v / (g.nx * g.ny);
    const float *w = st + (size_t)v * 7;
    float a = w[0] * c[v];
    if (x > 0) a -= w[1] * c[v - 1];
    if (x < g.nx - 1) a -= w[2] * c[v + 1];
    if (y > 0) a -= w[3] * c[v - g.nx];
    if (y < g.ny - 1) a -= w[4] * c[v + g.nx];
    if (z > 0) a -= w[5] * c[v - g.nx * g.ny];
    if (z < g.nz - 1) a -= w[6] * c[v + g.nx * g.ny];
    out[v] = a;
}

void CNN::set_problem(const double *kcell, double beta)
{
    // kinv feature and the Galerkin coarse operator of element means: for
    // piecewise constants only the SIPG penalty survives,
    // a(Pc, Pc) = sum_F kp beta / h_F |F| [c]^2  (kp = avg kappa, boundary: kappa)
    std::vector<float> k(V), st((size_t)V * 7);
    const double h[3] = { 1.0 / n[0], 1.0 / n[1], 1.0 / n[2] };
    const int stride[3] = { 1, n[0], n[0] * n[1] };
    for (int64_t v = 0; v < V; ++v) {
        k[v] = (float)(1.0 / kcell[v]);
        const int c[3] = { (int)(v % n[0]), (int)((v / n[0]) % n[1]), (int)(v / ((int64_t)n[0] * n[1])) };
        double diag = 0;
        for (int d = 0; d < 3; ++d) {
            const double area = h[(d + 1) % 3] * h[(d + 2) % 3];
            for (int s2 = 0; s2 < 2; ++s2) {
                const bool interior = s2 ? c[d] + 1 < n[d] : c[d] > 0;
                double wgt;
                if (interior) {
                    const double kn = kcell[v + (s2 ? stride[d] : -stride[d])];
                    wgt = 0.5 * (kcell[v] + kn) * beta / h[d] * area;
                    st[(size_t)v * 7 + 1 + 2 * d + s2] = (float)wgt;
                } else {
                    wgt = kcell[v] * beta / h[d] * area;
                    st[(size_t)v * 7 + 1 + 2 * d + s2] = 0.f;
                }
                diag += wgt;
            }
        }
        st[(size_t)v * 7] = (float)diag;
    }
    kinv.alloc(sizeof(float) * V);
    stencil.alloc(sizeof(float) * V * 7);
    FEM_HIP(hipMemcpy(kinv.p, k.data(), sizeof(float) * V, hipMemcpyHostToDevice));
    FEM_HIP(hipMemcpy(stencil.p, st.data(), sizeof(float) * V * 7, hipMemcpyHostToDevice));
}

void CNN::coarse_apply(const float *c, float *out, hipStream_t s) const
{
    Grid g{ n[0], n[1], n[2], (int)V };
    hipLaunchKernelGGL(k_coarse_stencil_apply, dim3((V + 255) / 256), dim3(256), 0, s, g, stencil.as<float>(), c, out);
    FEM_HIP(hipGetLastError());
}

double CNN::loss_value(hipStream_t s)
{
    double h[256];
    FEM_HIP(hipMemcpyAsync(h, lossp.p, sizeof(h), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    double t = 0;
    for (int i = 0; i < 256; ++i) t += h[i];
    return t / (double)V;
}

void CNN::adam_step(float lr, float scale, hipStream_t s)
{
    ++adam_t;
    const float b1t = std::pow(0.9f, (float)adam_t), b2t = std::pow(0.999f, (float)adam_t);
    const int nn = (int)nparams;
    hipLaunchKernelGGL(k_adam, dim3((nn + 255) / 256), dim3(256), 0, s, wmaster.as<float>(), grad.as<float>(),
                       adam_m.as<float>(), adam_v.as<float>(), nn, lr, scale, b1t, b2t);
    FEM_HIP(hipGetLastError());
    refresh();
}

}  // namespace femamd
