// Allocation-free scalar product for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The reduction scratch of the scalar product -- one FP64 partial per
// workgroup of the fixed RED_DOT_BLOCKS grid, the arrival counter and the
// result slot -- lives in the file-scope device array `scratch` below
// (4,352 bytes per GPU, held until the process ends), not in the context's
// workspace.  Its size does not depend on the vector length: every length
// the vectors can have is accepted.  Host code never passes `scratch`
// itself to a kernel (on the host the name is only a stand-in for the
// symbol); it looks up the device address once per device with
// hipGetSymbolAddress and uses that pointer.  All calls on one GPU share the
// array: calls that overlap on one GPU (several host threads or streams at
// once) race, and the array keeps the last call's contents.
#include <mutex>

#include "fem_blas.hpp"

namespace femamd {

namespace {

constexpr int NA_PARTIALS = RED_DOT_BLOCKS;              // one per workgroup
constexpr int NA_RESULT = NA_PARTIALS;                   // result slot
constexpr int NA_COUNTER = NA_PARTIALS + 1;              // arrival counter (low 32 bits)
constexpr int NA_DOUBLES = ((NA_PARTIALS + 2 + 31) / 32) * 32;

static __device__ __align__(256) double scratch[NA_DOUBLES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
double *na_ptr[NA_MAX_DEVICES] = {};

// device address of this GPU's copy of `scratch`, looked up on the first call
// on each device (with that device current)
double *scratch_ptr(int device)
{
    FEM_CHECK(device >= 0 && device < NA_MAX_DEVICES, FEM_ERR_ARG, "no-alloc scratch: device index");
    std::lock_guard<std::mutex> g(na_mutex);
    if (!na_ptr[device]) {
        int cur = -1;
        FEM_HIP(hipGetDevice(&cur));
        if (cur != device) FEM_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess)
            throw Error(FEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        na_ptr[device] = (double *)p;
    }
    return na_ptr[device];
}

// the counter must start at zero: cleared by a kernel (not a memset, so the
// call stays capturable in a HIP graph)
__global__ void k_na_clear(double *s)
{
    if (threadIdx.x == 0) reinterpret_cast<unsigned *>(s + NA_COUNTER)[0] = 0u;
}

template <typename TX, typename TY>
void na_launch(hipStream_t st, int64_t n, const TX *x, const TY *y, double *s)
{
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, st, s);
    hipLaunchKernelGGL((k_dot<TX, TY>), dim3(NA_PARTIALS), dim3(RED_THREADS), 0, st, x, y, n, s,
                       reinterpret_cast<unsigned *>(s + NA_COUNTER), s + NA_RESULT);
    FEM_HIP(hipGetLastError());
}

}  // namespace

double vec_dot_noalloc(const Vec &x, const Vec &y)
{
    FEM_CHECK(x.n == y.n, FEM_ERR_ARG, "dot: size mismatch");
    FEM_CHECK(x.ctx == y.ctx, FEM_ERR_ARG, "dot: vectors of different contexts");
    Context &c = *x.ctx;
    double *s = scratch_ptr(c.device);
    if (x.prec == FEM_F64 && y.prec == FEM_F64) na_launch(c.stream, x.n, (const double *)x.data, (const double *)y.data, s);
    else if (x.prec == FEM_F64) na_launch(c.stream, x.n, (const double *)x.data, (const float *)y.data, s);
    else if (y.prec == FEM_F64) na_launch(c.stream, x.n, (const float *)x.data, (const double *)y.data, s);
    else na_launch(c.stream, x.n, (const float *)x.data, (const float *)y.data, s);
    // the context's pinned result slot (allocated with the context) receives
    // the scalar; no memory is allocated here
    double *h = c.host_result + (RED_MAX_SLOTS - 1);
    FEM_HIP(hipMemcpyAsync(h, s + NA_RESULT, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FEM_HIP(hipStreamSynchronize(c.stream));
    return *h;
}

}  // namespace femamd
