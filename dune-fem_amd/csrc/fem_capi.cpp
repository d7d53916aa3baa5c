// extern "C" boundary: opaque handles, int status codes, no exceptions
// across the ABI (the reference throws Dune exceptions; non-convergence is a
// negative iteration count, not an error).
#include <algorithm>
#include <climits>
#include <cstddef>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "fem_batch.hpp"
#include "fem_blas.hpp"
#include "fem_cnn.hpp"
#include "fem_csr.hpp"
#include "fem_learned.hpp"
#include "fem_newton.hpp"
#include "fem_op.hpp"
#include "fem_solver.hpp"

using namespace femamd;

static thread_local std::string g_last_error;

template <typename F>
static int guarded(F &&f)
{
    try {
        f();
        return FEM_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return FEM_ERR_ARG;
    } catch (...) {
        g_last_error = "unknown error";
        return FEM_ERR_ARG;
    }
}

extern "C" {

const char *fem_last_error(void) { return g_last_error.c_str(); }
int fem_version(void) { return 10000; }

int fem_ctx_create(int device, fem_ctx **out)
{
    return guarded([&] {
        FEM_CHECK(out, FEM_ERR_ARG, "null output handle");
        int n = 0;
        FEM_HIP(hipGetDeviceCount(&n));
        FEM_CHECK(device >= 0 && device < n, FEM_ERR_ARG, "no such HIP device");
        FEM_HIP(hipSetDevice(device));
        fem_ctx *h = new fem_ctx();
        Context &c = h->c;
        c.device = device;
        try {
            FEM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            FEM_HIP(hipMalloc(&c.red_partial, sizeof(double) * RED_DOT_BLOCKS * RED_MAX_SLOTS));
            FEM_HIP(hipMalloc(&c.red_counter, sizeof(unsigned) * RED_MAX_SLOTS));
            FEM_HIP(hipMemset(c.red_counter, 0, sizeof(unsigned) * RED_MAX_SLOTS));
            FEM_HIP(hipMalloc(&c.red_result, sizeof(double) * RED_MAX_SLOTS));
            FEM_HIP(hipHostMalloc(&c.host_result, sizeof(double) * RED_MAX_SLOTS, hipHostMallocDefault));
            FEM_HIP(hipEventCreate(&c.ev0));
            FEM_HIP(hipEventCreate(&c.ev1));
            FEM_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

static_assert(std::is_standard_layout<fem_ctx>::value && offsetof(fem_ctx, c) == 0,
              "a Context pointer converts back to its fem_ctx");

// Every handle made on a context holds a reference to it, so a handle that
// outlives fem_ctx_destroy still finds its stream and reduction workspace.
static Context *ctx_retain(Context *c)
{
    c->refs.fetch_add(1);
    return c;
}

static void ctx_release(Context *c)
{
    if (!c || c->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->red_partial) (void)hipFree(c->red_partial);
    if (c->red_counter) (void)hipFree(c->red_counter);
    if (c->red_result) (void)hipFree(c->red_result);
    if (c->host_result) (void)hipHostFree(c->host_result);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete reinterpret_cast<fem_ctx *>(c);
}

int fem_ctx_destroy(fem_ctx *h)
{
    return guarded([&] {
        if (h) ctx_release(&h->c);
    });
}

int fem_ctx_synchronize(fem_ctx *h)
{
    return guarded([&] {
        FEM_CHECK(h, FEM_ERR_ARG, "null context");
        FEM_HIP(hipStreamSynchronize(h->c.stream));
    });
}

int fem_ctx_mark(fem_ctx *h, int which)
{
    return guarded([&] {
        FEM_CHECK(h, FEM_ERR_ARG, "null context");
        FEM_HIP(hipEventRecord(which ? h->c.ev1 : h->c.ev0, h->c.stream));
    });
}

int fem_ctx_elapsed_ms(fem_ctx *h, float *ms)
{
    return guarded([&] {
        FEM_CHECK(h && ms, FEM_ERR_ARG, "null argument");
        FEM_HIP(hipEventSynchronize(h->c.ev1));
        FEM_HIP(hipEventElapsedTime(ms, h->c.ev0, h->c.ev1));
    });
}

// ---- vectors -----------------------------------------------------------------
int fem_vec_create(fem_ctx *ctx, int64_t n, int precision, fem_vec **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && out, FEM_ERR_ARG, "null argument");
        FEM_CHECK(n >= 0, FEM_ERR_ARG, "negative size");
        FEM_CHECK(precision == FEM_F64 || precision == FEM_F32, FEM_ERR_ARG, "vector precision must be FEM_F64 or FEM_F32");
        fem_vec *h = new fem_vec();
        h->v.ctx = &ctx->c;
        h->v.n = n;
        h->v.prec = precision;
        h->held = ctx_retain(&ctx->c);
        if (n > 0) {
            hipError_t e = hipMalloc(&h->v.data, h->v.bytes());
            if (e != hipSuccess) {
                delete h;
                ctx_release(&ctx->c);
                throw Error(FEM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
            }
            FEM_HIP(hipMemsetAsync(h->v.data, 0, h->v.bytes(), ctx->c.stream));
        }
        *out = h;
    });
}

int fem_vec_wrap(fem_ctx *ctx, int64_t n, int precision, void *device_ptr, fem_vec **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && out, FEM_ERR_ARG, "null argument");
        FEM_CHECK(n >= 0, FEM_ERR_ARG, "negative size");
        FEM_CHECK(precision == FEM_F64 || precision == FEM_F32, FEM_ERR_ARG, "vector precision must be FEM_F64 or FEM_F32");
        FEM_CHECK(n == 0 || device_ptr, FEM_ERR_ARG, "null device pointer");
        FEM_CHECK(((uintptr_t)device_ptr & 255) == 0, FEM_ERR_ARG, "device pointer must be 256-byte aligned");
        hipPointerAttribute_t attr;
        if (n > 0) {
            FEM_CHECK(hipPointerGetAttributes(&attr, device_ptr) == hipSuccess && attr.type == hipMemoryTypeDevice,
                      FEM_ERR_ARG, "not a device allocation");
        }
        fem_vec *h = new fem_vec();
        h->v.ctx = &ctx->c;
        h->v.n = n;
        h->v.prec = precision;
        h->v.data = n > 0 ? device_ptr : nullptr;
        h->v.owner = false;
        h->held = ctx_retain(&ctx->c);
        *out = h;
    });
}

int fem_vec_destroy(fem_vec *v)
{
    return guarded([&] {
        if (!v) return;
        if (v->v.data && v->v.owner) {
            (void)hipStreamSynchronize(v->v.ctx->stream);
            (void)hipFree(v->v.data);
        }
        Context *c = v->held;
        delete v;
        ctx_release(c);
    });
}

int64_t fem_vec_size(const fem_vec *v) { return v ? v->v.n : -1; }
int fem_vec_precision(const fem_vec *v) { return v ? v->v.prec : -1; }
void *fem_vec_device_ptr(fem_vec *v) { return v ? v->v.data : nullptr; }

int fem_vec_upload(fem_vec *v, const double *host)
{
    return guarded([&] {
        FEM_CHECK(v && host, FEM_ERR_ARG, "null argument");
        Vec &x = v->v;
        if (x.n == 0) return;
        if (x.prec == FEM_F64) {
            FEM_HIP(hipMemcpyAsync(x.data, host, x.bytes(), hipMemcpyHostToDevice, x.ctx->stream));
        } else {
            std::vector<float> tmp(x.n);
            for (int64_t i = 0; i < x.n; ++i) tmp[i] = (float)host[i];
            FEM_HIP(hipMemcpyAsync(x.data, tmp.data(), x.bytes(), hipMemcpyHostToDevice, x.ctx->stream));
            FEM_HIP(hipStreamSynchronize(x.ctx->stream));
            return;
        }
        FEM_HIP(hipStreamSynchronize(x.ctx->stream));
    });
}

int fem_vec_upload_f32(fem_vec *v, const float *host)
{
    return guarded([&] {
        FEM_CHECK(v && host, FEM_ERR_ARG, "null argument");
        Vec &x = v->v;
        FEM_CHECK(x.prec == FEM_F32, FEM_ERR_ARG, "upload_f32 needs an FP32 vector");
        if (x.n == 0) return;
        FEM_HIP(hipMemcpyAsync(x.data, host, x.bytes(), hipMemcpyHostToDevice, x.ctx->stream));
        FEM_HIP(hipStreamSynchronize(x.ctx->stream));
    });
}

int fem_vec_download(const fem_vec *v, double *host)
{
    return guarded([&] {
        FEM_CHECK(v && host, FEM_ERR_ARG, "null argument");
        const Vec &x = v->v;
        if (x.n == 0) return;
        if (x.prec == FEM_F64) {
            FEM_HIP(hipMemcpyAsync(host, x.data, x.bytes(), hipMemcpyDeviceToHost, x.ctx->stream));
            FEM_HIP(hipStreamSynchronize(x.ctx->stream));
        } else {
            std::vector<float> tmp(x.n);
            FEM_HIP(hipMemcpyAsync(tmp.data(), x.data, x.bytes(), hipMemcpyDeviceToHost, x.ctx->stream));
            FEM_HIP(hipStreamSynchronize(x.ctx->stream));
            for (int64_t i = 0; i < x.n; ++i) host[i] = tmp[i];
        }
    });
}

// ---- DoF vectors in the reference's binary stream format ---------------------------
}  // extern "C"

namespace {

constexpr unsigned kDuneFemVersionId = (2u << 24) | (12u << 16);   // DUNE_VERSION_ID(2,12,0)
constexpr unsigned kOldestVersionId = (0u << 24) | (9u << 16) | 1u;  // DUNE_VERSION_ID(0,9,1)
constexpr unsigned kSpaceIdVersionId = (1u << 24) | (5u << 16);     // DUNE_VERSION_ID(1,5,0)

// little-endian primitives, byte by byte (the stream's byte order whatever
// the host's)
template <typename U>
void put_le(std::vector<unsigned char> &out, U v)
{
    unsigned char b[sizeof(U)];
    std::memcpy(b, &v, sizeof(U));
    const uint16_t probe = 1;
    const bool little = *reinterpret_cast<const unsigned char *>(&probe) == 1;
    for (size_t i = 0; i < sizeof(U); ++i) out.push_back(b[little ? i : sizeof(U) - 1 - i]);
}

template <typename U>
U get_le(const unsigned char *&p, const unsigned char *end)
{
    FEM_CHECK(end - p >= (ptrdiff_t)sizeof(U), FEM_ERR_ARG, "DoF stream: truncated");
    unsigned char b[sizeof(U)];
    const uint16_t probe = 1;
    const bool little = *reinterpret_cast<const unsigned char *>(&probe) == 1;
    for (size_t i = 0; i < sizeof(U); ++i) b[little ? i : sizeof(U) - 1 - i] = p[i];
    p += sizeof(U);
    U v;
    std::memcpy(&v, b, sizeof(U));
    return v;
}

std::vector<unsigned char> dofs_encode(const double *dofs, int64_t n, const char *name, int space_id)
{
    FEM_CHECK(n >= 0 && n <= INT32_MAX, FEM_ERR_ARG, "DoF stream: the size is written as a 32-bit integer");
    FEM_CHECK(n == 0 || dofs, FEM_ERR_ARG, "null DoFs");
    const std::string nm = name ? name : "";
    std::vector<unsigned char> out;
    out.reserve(16 + nm.size() + 8 * (size_t)n);
    put_le<uint32_t>(out, kDuneFemVersionId);
    put_le<int32_t>(out, space_id);
    put_le<uint32_t>(out, (uint32_t)nm.size());
    for (char c : nm) out.push_back((unsigned char)c);
    put_le<int32_t>(out, (int32_t)n);
    for (int64_t i = 0; i < n; ++i) put_le<double>(out, dofs[i]);
    return out;
}

// the DoFs into dofs[n]; returns the name
std::string dofs_decode(const unsigned char *buf, int64_t len, int space_id, double *dofs, int64_t n)
{
    FEM_CHECK(buf || len == 0, FEM_ERR_ARG, "null buffer");
    const unsigned char *p = buf, *end = buf + len;
    const uint32_t version = get_le<uint32_t>(p, end);
    FEM_CHECK(version >= kOldestVersionId, FEM_ERR_ARG, "DoF stream: trying to read outdated file");
    if (version >= kSpaceIdVersionId) {
        const int32_t sid = get_le<int32_t>(p, end);
        FEM_CHECK(sid == space_id, FEM_ERR_ARG, "DoF stream: written from a different space");
    }
    const uint32_t nl = get_le<uint32_t>(p, end);
    FEM_CHECK(end - p >= (ptrdiff_t)nl, FEM_ERR_ARG, "DoF stream: truncated");
    std::string nm(reinterpret_cast<const char *>(p), nl);
    p += nl;
    const int32_t size = get_le<int32_t>(p, end);
    FEM_CHECK(size == n, FEM_ERR_ARG, "DoF stream: trying to read a discrete function of different size");
    FEM_CHECK(n == 0 || dofs, FEM_ERR_ARG, "null DoFs");
    for (int64_t i = 0; i < n; ++i) dofs[i] = get_le<double>(p, end);
    return nm;
}

void copy_name(const std::string &nm, char *name, int64_t cap)
{
    if (!name || cap <= 0) return;
    const size_t k = std::min<size_t>(nm.size(), (size_t)(cap - 1));
    std::memcpy(name, nm.data(), k);
    name[k] = 0;
}

}  // namespace

extern "C" {

int fem_dofs_encode(const double *dofs, int64_t n, const char *name, int space_id, unsigned char *buf, int64_t cap,
                    int64_t *len)
{
    return guarded([&] {
        FEM_CHECK(len, FEM_ERR_ARG, "null length");
        const std::vector<unsigned char> out = dofs_encode(dofs, n, name, space_id);
        *len = (int64_t)out.size();
        FEM_CHECK(buf && cap >= (int64_t)out.size(), FEM_ERR_ARG, "DoF stream: buffer too small");
        std::memcpy(buf, out.data(), out.size());
    });
}

int fem_dofs_decode(const unsigned char *buf, int64_t len, int space_id, double *dofs, int64_t n, char *name,
                    int64_t name_cap)
{
    return guarded([&] { copy_name(dofs_decode(buf, len, space_id, dofs, n), name, name_cap); });
}

int fem_vec_write(const fem_vec *v, const char *path, const char *name, int space_id)
{
    return guarded([&] {
        FEM_CHECK(v && path, FEM_ERR_ARG, "null argument");
        std::vector<double> host(v->v.n);
        FEM_CHECK(fem_vec_download(v, host.data()) == FEM_OK, FEM_ERR_HIP, g_last_error);
        const std::vector<unsigned char> out = dofs_encode(host.data(), v->v.n, name, space_id);
        std::FILE *f = std::fopen(path, "wb");
        FEM_CHECK(f, FEM_ERR_ARG, std::string("cannot open for writing: ") + path);
        const size_t w = std::fwrite(out.data(), 1, out.size(), f);
        const int rc = std::fclose(f);
        FEM_CHECK(w == out.size() && rc == 0, FEM_ERR_ARG, std::string("write failed: ") + path);
    });
}

int fem_vec_read(fem_vec *v, const char *path, int space_id, char *name, int64_t name_cap)
{
    return guarded([&] {
        FEM_CHECK(v && path, FEM_ERR_ARG, "null argument");
        std::FILE *f = std::fopen(path, "rb");
        FEM_CHECK(f, FEM_ERR_ARG, std::string("cannot open for reading: ") + path);
        std::vector<unsigned char> buf;
        unsigned char chunk[1 << 16];
        size_t k;
        while ((k = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + k);
        std::fclose(f);
        std::vector<double> host(v->v.n);
        const std::string nm = dofs_decode(buf.data(), (int64_t)buf.size(), space_id, host.data(), v->v.n);
        FEM_CHECK(fem_vec_upload(v, host.data()) == FEM_OK, FEM_ERR_HIP, g_last_error);
        copy_name(nm, name, name_cap);
    });
}

int fem_vec_fill(fem_vec *v, double a)
{
    return guarded([&] {
        FEM_CHECK(v, FEM_ERR_ARG, "null vector");
        if (v->v.n) vec_fill(v->v, a);
    });
}

int fem_vec_copy(fem_vec *d, const fem_vec *s)
{
    return guarded([&] {
        FEM_CHECK(d && s, FEM_ERR_ARG, "null vector");
        if (d->v.n) vec_copy(d->v, s->v);
    });
}

int fem_vec_axpy(fem_vec *y, double a, const fem_vec *x)
{
    return guarded([&] {
        FEM_CHECK(y && x, FEM_ERR_ARG, "null vector");
        if (y->v.n) vec_axpy(y->v, a, x->v);
    });
}

int fem_vec_scal(fem_vec *x, double a)
{
    return guarded([&] {
        FEM_CHECK(x, FEM_ERR_ARG, "null vector");
        if (x->v.n) vec_scal(x->v, a);
    });
}

int fem_vec_dot(const fem_vec *x, const fem_vec *y, double *out)
{
    return guarded([&] {
        FEM_CHECK(x && y && out, FEM_ERR_ARG, "null argument");
        *out = vec_dot(x->v, y->v);
    });
}

int fem_vec_dot_noalloc(const fem_vec *x, const fem_vec *y, double *out)
{
    return guarded([&] {
        FEM_CHECK(x && y && out, FEM_ERR_ARG, "null argument");
        *out = vec_dot_noalloc(x->v, y->v);
    });
}

int fem_vec_norm2sq(const fem_vec *x, double *out)
{
    return guarded([&] {
        FEM_CHECK(x && out, FEM_ERR_ARG, "null argument");
        *out = vec_dot(x->v, x->v);
    });
}

// ---- operators -----------------------------------------------------------------
int fem_op_create(fem_ctx *ctx, const fem_op_desc *desc, fem_op **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && desc && out, FEM_ERR_ARG, "null argument");
        fem_op *h = new fem_op();
        try {
            h->o.init(&ctx->c, *desc);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&ctx->c);
        *out = h;
    });
}

int fem_op_destroy(fem_op *op)
{
    return guarded([&] {
        if (op) {
            (void)hipStreamSynchronize(op->o.ctx->stream);
            Context *c = op->held;
            delete op;
            ctx_release(c);
        }
    });
}

int64_t fem_op_ndofs(const fem_op *op) { return op ? op->o.N : -1; }

int fem_op_set_kappa_qp(fem_op *op, const double *kvol, const double *kx, const double *ky, const double *kz)
{
    return guarded([&] {
        FEM_CHECK(op, FEM_ERR_ARG, "null operator");
        op->o.set_kappa_qp(kvol, kx, ky, kz);
    });
}

int fem_op_set_kappa_cell(fem_op *op, const double *kc)
{
    return guarded([&] {
        FEM_CHECK(op, FEM_ERR_ARG, "null operator");
        op->o.set_kappa_cell(kc);
    });
}

int fem_op_apply(fem_op *op, const fem_vec *u, fem_vec *w)
{
    return guarded([&] {
        FEM_CHECK(op && u && w, FEM_ERR_ARG, "null argument");
        op->o.apply(u->v, w->v);
    });
}

int fem_op_assemble_rhs(fem_op *op, const double *fq, fem_vec *F)
{
    return guarded([&] {
        FEM_CHECK(op && fq && F, FEM_ERR_ARG, "null argument");
        op->o.rhs(fq, F->v);
    });
}

int fem_op_diagonal(fem_op *op, fem_vec *d)
{
    return guarded([&] {
        FEM_CHECK(op && d, FEM_ERR_ARG, "null argument");
        op->o.diagonal(d->v);
    });
}

int fem_op_set_boundary_data(fem_op *op, const double *g_x, const double *g_y, const double *g_z)
{
    return guarded([&] {
        FEM_CHECK(op, FEM_ERR_ARG, "null operator");
        op->o.set_boundary_data(g_x, g_y, g_z);
    });
}

int fem_op_constrain(fem_op *op, fem_vec *x, const fem_vec *add)
{
    return guarded([&] {
        FEM_CHECK(op && x, FEM_ERR_ARG, "null argument");
        op->o.constrain(x->v, add ? &add->v : nullptr);
    });
}

int fem_op_block_diagonal(fem_op *op, fem_vec *blocks)
{
    return guarded([&] {
        FEM_CHECK(op && blocks, FEM_ERR_ARG, "null argument");
        op->o.block_diagonal(blocks->v);
    });
}

// ---- pointwise and nonlinear reaction ----------------------------------------------
int fem_op_set_reaction_qp(fem_op *op, const double *cvol)
{
    return guarded([&] {
        FEM_CHECK(op, FEM_ERR_ARG, "null operator");
        op->o.set_reaction_qp(cvol);
    });
}

int fem_op_set_reaction_law(fem_op *op, double coef, int pw)
{
    return guarded([&] {
        FEM_CHECK(op, FEM_ERR_ARG, "null operator");
        op->o.set_reaction_law(coef, pw);
    });
}

int fem_op_set_reaction_plugin(fem_op *op, const char *path)
{
    return guarded([&] {
        FEM_CHECK(op, FEM_ERR_ARG, "null operator");
        op->o.set_reaction_plugin(path);
    });
}

int fem_op_nonlinear(const fem_op *op)
{
    if (!op) return FEM_ERR_ARG;
    return op->o.nonlinear() ? 1 : 0;
}

int fem_op_eval_qp(fem_op *op, const fem_vec *u, fem_vec *uq)
{
    return guarded([&] {
        FEM_CHECK(op && u && uq, FEM_ERR_ARG, "null argument");
        op->o.eval_qp(u->v, uq->v);
    });
}

int fem_op_jacobian_create(const fem_op *op, fem_op **jop)
{
    return guarded([&] {
        FEM_CHECK(op && jop, FEM_ERR_ARG, "null argument");
        fem_op *h = new fem_op();
        try {
            op->o.make_jacobian(h->o);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&*op->o.ctx);
        *jop = h;
    });
}

int fem_op_jacobian(const fem_op *op, const fem_vec *u, fem_op *jop)
{
    return guarded([&] {
        FEM_CHECK(op && u && jop, FEM_ERR_ARG, "null argument");
        op->o.jacobian(u->v, jop->o);
    });
}

// ---- assembled Jacobian ----------------------------------------------------------
int fem_op_assemble_csr(fem_op *op, fem_csr **out)
{
    return guarded([&] {
        FEM_CHECK(op && out, FEM_ERR_ARG, "null argument");
        FEM_CHECK(op->o.is_dg() ? op->o.a64.kappa_mode >= 0
                                : (op->o.desc.kappa_mode == FEM_KAPPA_CONST || op->o.kvol64.p),
                  FEM_ERR_STATE, "coefficient data not set");
        FEM_CHECK(!op->o.nonlinear(), FEM_ERR_STATE, "assemble: the operator is nonlinear (take its Jacobian)");
        fem_csr *h = new fem_csr();
        try {
            h->m.assemble(&op->o);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&*op->o.ctx);
        *out = h;
    });
}

int fem_csr_destroy(fem_csr *A)
{
    return guarded([&] {
        if (!A) return;
        Context *c = A->held;
        delete A;
        ctx_release(c);
    });
}

int64_t fem_csr_rows(const fem_csr *A) { return A ? A->m.N : -1; }
int64_t fem_csr_nnz(const fem_csr *A) { return A ? A->m.nnz : -1; }

int fem_csr_apply(const fem_csr *A, const fem_vec *x, fem_vec *y)
{
    return guarded([&] {
        FEM_CHECK(A && x && y, FEM_ERR_ARG, "null argument");
        A->m.apply(x->v, y->v);
    });
}

int fem_csr_diagonal(const fem_csr *A, fem_vec *d)
{
    return guarded([&] {
        FEM_CHECK(A && d, FEM_ERR_ARG, "null argument");
        A->m.diagonal(d->v);
    });
}

int fem_csr_download(const fem_csr *A, int64_t *rowptr, int *col, double *val)
{
    return guarded([&] {
        FEM_CHECK(A, FEM_ERR_ARG, "null argument");
        A->m.download(rowptr, col, val);
    });
}

// ---- preconditioners -----------------------------------------------------------
int fem_precond_create(fem_op *op, int kind, fem_precond **out)
{
    return guarded([&] {
        FEM_CHECK(op && out, FEM_ERR_ARG, "null argument");
        fem_precond *h = new fem_precond();
        try {
            h->p.init(&op->o, kind);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&*op->o.ctx);
        *out = h;
    });
}

int fem_precond_destroy(fem_precond *p)
{
    return guarded([&] {
        if (!p) return;
        Context *c = p->held;
        delete p;
        ctx_release(c);
    });
}

int fem_precond_apply(fem_precond *p, const fem_vec *x, fem_vec *y)
{
    return guarded([&] {
        FEM_CHECK(p && x && y, FEM_ERR_ARG, "null argument");
        FEM_CHECK(x->v.n == p->p.op->N && y->v.n == p->p.op->N, FEM_ERR_ARG, "precond: vector size");
        FEM_CHECK(x->v.prec == p->p.prec && y->v.prec == p->p.prec, FEM_ERR_ARG, "precond: vector precision");
        p->p.apply_raw(x->v.data, y->v.data, nullptr);
    });
}

// ---- solvers -------------------------------------------------------------------
void fem_solver_default_params(fem_solver_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->method = FEM_SOLVER_GMRES;     // KrylovInverseOperator default (krylovinverseoperators.hh:103-107)
    p->tolerance = 1e-8;              // parameter.hh:118-141
    p->errormeasure = FEM_ERRMEASURE_ABSOLUTE;
    p->maxiterations = 1 << 30;
    p->restart = 20;                  // parameter.hh:197-201
    p->verbose = 0;
    p->refinement = 0;
    p->inner_tolerance = 1e-3;
    p->inner_maxiterations = 1000;
    p->preconditioning = FEM_PRE_NONE;   // parameter.hh:206-217 (default: the first method, none)
    p->relaxation = 1.1;                 // parameter.hh:219-222
    p->precond_iterations = 1;           // parameter.hh:224-233
}

// "true"/"false" as ParameterReader::getValue<bool> reads them (also 1/0, yes/no)
static int parse_bool(const std::string &k, const std::string &v)
{
    if (v == "true" || v == "1" || v == "yes") return 1;
    if (v == "false" || v == "0" || v == "no") return 0;
    throw Error(FEM_ERR_ARG, k + ": expected true or false, got " + v);
}

static double parse_double(const std::string &k, const std::string &v)
{
    size_t pos = 0;
    double d = 0;
    try { d = std::stod(v, &pos); } catch (...) { pos = 0; }
    FEM_CHECK(pos == v.size() && pos > 0, FEM_ERR_ARG, k + ": not a number: " + v);
    return d;
}

static int parse_int(const std::string &k, const std::string &v)
{
    size_t pos = 0;
    long d = 0;
    try { d = std::stol(v, &pos); } catch (...) { pos = 0; }
    FEM_CHECK(pos == v.size() && pos > 0, FEM_ERR_ARG, k + ": not an integer: " + v);
    return (int)d;
}

int fem_solver_params_set(fem_solver_params *p, const char *key, const char *value)
{
    return guarded([&] {
        FEM_CHECK(p && key && value, FEM_ERR_ARG, "null argument");
        std::string k(key), v(value);
        // SolverParameter's key prefix (parameter.hh:74-78), then the scheme
        // prefixes of the linear solver (python/dune/fem/scheme/_schemes.py:40-60)
        auto strip = [&](const char *pre) {
            const size_t n = std::strlen(pre);
            if (k.compare(0, n, pre) == 0) { k = k.substr(n); return true; }
            return false;
        };
        strip("fem.solver.");
        strip("newton.linear.") || strip("nonlinear.linear.") || strip("linear.");
        FEM_CHECK(k.compare(0, 10, "nonlinear.") != 0 && k.compare(0, 7, "newton.") != 0, FEM_ERR_ARG,
                  std::string("a Newton parameter, not a linear-solver one: ") + key);
        if (k == "method") {
            if (v == "cg") p->method = FEM_SOLVER_CG;
            else if (v == "gmres") p->method = FEM_SOLVER_GMRES;
            else if (v == "fgmres") p->method = FEM_SOLVER_FGMRES;
            else if (v == "bicgstab") p->method = FEM_SOLVER_BICGSTAB;
            else throw Error(FEM_ERR_UNSUPPORTED, "method: " + v + " (supported: cg, gmres, fgmres, bicgstab)");
        } else if (k == "tolerance" || k == "absolutetol" || k == "reductiontol") {
            p->tolerance = parse_double(k, v);
            FEM_CHECK(p->tolerance >= 0, FEM_ERR_ARG, k + " must be >= 0");
        } else if (k == "errormeasure") {
            if (v == "absolute") p->errormeasure = FEM_ERRMEASURE_ABSOLUTE;
            else if (v == "relative") p->errormeasure = FEM_ERRMEASURE_RELATIVE;
            else if (v == "residualreduction") p->errormeasure = FEM_ERRMEASURE_REDUCTION;
            else throw Error(FEM_ERR_ARG, "errormeasure: " + v);
        } else if (k == "maxiterations" || k == "maxlineariterations") {
            p->maxiterations = parse_int(k, v);
            FEM_CHECK(p->maxiterations >= 0, FEM_ERR_ARG, k + " must be >= 0");
        } else if (k == "gmres.restart") {
            p->restart = parse_int(k, v);
        } else if (k == "verbose") {
            p->verbose = parse_bool(k, v);
        } else if (k == "residualcheck") {
            p->residual_check = parse_bool(k, v);
        } else if (k == "preconditioning.method") {
            // the reference's names (parameter.hh:48-59; underscores read as dashes)
            if (v == "none") p->preconditioning = FEM_PRE_NONE;
            else if (v == "jacobi") p->preconditioning = FEM_PRE_JACOBI;
            else if (v == "blockjacobi" || v == "block-jacobi") p->preconditioning = FEM_PRE_BLOCK_JACOBI;
            else if (v == "multigrid" || v == "pmg") p->preconditioning = FEM_PRE_MULTIGRID;
            else if (v == "sor" || v == "ssor" || v == "gauss-seidel" || v == "ilu" || v == "amg-ilu" ||
                     v == "amg-jacobi" || v == "ildl" || v == "oas" || v == "icc" || v == "kspoptions")
                throw Error(FEM_ERR_UNSUPPORTED, "preconditioning.method " + v +
                                                     ": needs an assembled matrix; the matrix-free path offers none, "
                                                     "jacobi, blockjacobi and multigrid");
            else throw Error(FEM_ERR_ARG, "preconditioning.method: unknown value " + v);
        } else if (k == "preconditioning.relaxation") {
            p->relaxation = parse_double(k, v);
        } else if (k == "preconditioning.iterations" || k == "preconditioning.iteration") {
            p->precond_iterations = parse_int(k, v);
            FEM_CHECK(p->precond_iterations >= 1, FEM_ERR_ARG, k + " must be >= 1");
        } else if (k == "preconditioning.level") {
            FEM_CHECK(parse_int(k, v) == 0, FEM_ERR_UNSUPPORTED, "preconditioning.level: only 0 (no ILU levels)");
        } else if (k == "threading") {
            (void)parse_bool(k, v);   // one device stream per context: nothing to switch
        } else if (k == "knolltrick") {
            FEM_CHECK(parse_bool(k, v) == 0, FEM_ERR_UNSUPPORTED, "knolltrick: only false");
        } else {
            throw Error(FEM_ERR_ARG, std::string("unknown solver parameter: ") + key);
        }
    });
}

int fem_solver_create(fem_ctx *ctx, const fem_solver_params *params, fem_solver **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && params && out, FEM_ERR_ARG, "null argument");
        fem_solver *h = new fem_solver();
        try {
            h->s.init(&ctx->c, *params);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&ctx->c);
        *out = h;
    });
}

int fem_solver_destroy(fem_solver *s)
{
    return guarded([&] {
        if (!s) return;
        Context *c = s->held;
        delete s;
        ctx_release(c);
    });
}

int fem_solver_bind(fem_solver *s, fem_op *op, fem_precond *pre, fem_op *iop, fem_precond *ipre)
{
    return guarded([&] {
        FEM_CHECK(s && op, FEM_ERR_ARG, "null argument");
        s->s.bind(&op->o, pre ? &pre->p : nullptr, iop ? &iop->o : nullptr, ipre ? &ipre->p : nullptr);
    });
}

int fem_solver_unbind(fem_solver *s)
{
    return guarded([&] {
        FEM_CHECK(s, FEM_ERR_ARG, "null solver");
        s->s.unbind();
    });
}

int fem_solver_solve(fem_solver *s, const fem_vec *b, fem_vec *x, fem_solver_info *info)
{
    return guarded([&] {
        FEM_CHECK(s && b && x, FEM_ERR_ARG, "null argument");
        fem_solver_info tmp;
        s->s.solve(b->v, x->v, info ? *info : tmp);
    });
}

// ---- Newton ----------------------------------------------------------------------
void fem_newton_default_params(fem_newton_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->tolerance = 1e-6;                  // newtoninverseoperator.hh:203-207
    p->maxiterations = INT_MAX;           // :231-235
    p->maxlineariterations = -1;          // the linear solver's maxiterations (:246-250)
    p->linesearch = FEM_LINESEARCH_NONE;  // :277-288
    p->maxlinesearchiterations = INT_MAX; // :259-264
    p->forcing = FEM_FORCING_NONE;        // :298-315
    p->eisenstatwalker_etamax = 0.99;     // :326-331
    p->eisenstatwalker_gamma = 0.1;       // :343-348
    p->simplified = 0;                    // :359-362
    p->verbose = 0;
}

int fem_newton_params_set(fem_newton_params *p, const char *key, const char *value)
{
    return guarded([&] {
        FEM_CHECK(p && key && value, FEM_ERR_ARG, "null argument");
        std::string k(key), v(value);
        auto strip = [&](const char *pre) {
            const size_t n = std::strlen(pre);
            if (k.compare(0, n, pre) == 0) { k = k.substr(n); return true; }
            return false;
        };
        strip("fem.solver.");
        FEM_CHECK(strip("nonlinear.") || strip("newton."), FEM_ERR_ARG,
                  std::string("not a Newton parameter (nonlinear.*): ") + key);
        FEM_CHECK(k.compare(0, 7, "linear.") != 0 || k == "linear.tolerance.strategy", FEM_ERR_ARG,
                  std::string("a linear-solver parameter, set with fem_solver_params_set: ") + key);
        if (k == "tolerance") {
            p->tolerance = parse_double(k, v);
            FEM_CHECK(p->tolerance > 0, FEM_ERR_ARG, "tolerance must be > 0");
        } else if (k == "maxiterations") {
            p->maxiterations = parse_int(k, v);
            FEM_CHECK(p->maxiterations >= 0, FEM_ERR_ARG, "maxiterations must be >= 0");
        } else if (k == "maxlineariterations") {
            p->maxlineariterations = parse_int(k, v);
            FEM_CHECK(p->maxlineariterations >= 0, FEM_ERR_ARG, "maxlineariterations must be >= 0");
        } else if (k == "linesearch" || k == "lineSearch") {
            if (v == "none") p->linesearch = FEM_LINESEARCH_NONE;
            else if (v == "simple") p->linesearch = FEM_LINESEARCH_SIMPLE;
            else throw Error(FEM_ERR_ARG, "linesearch: none or simple, got " + v);
        } else if (k == "maxlinesearchiterations") {
            p->maxlinesearchiterations = parse_int(k, v);
            FEM_CHECK(p->maxlinesearchiterations >= 0, FEM_ERR_ARG, "maxlinesearchiterations must be >= 0");
        } else if (k == "forcing" || k == "linear.tolerance.strategy") {
            if (v == "none") p->forcing = FEM_FORCING_NONE;
            else if (v == "eisenstatwalker") p->forcing = FEM_FORCING_EISENSTATWALKER;
            else throw Error(FEM_ERR_ARG, "forcing: none or eisenstatwalker, got " + v);
        } else if (k == "eisenstatwalker.etamax") {
            p->eisenstatwalker_etamax = parse_double(k, v);
        } else if (k == "eisenstatwalker.gamma") {
            p->eisenstatwalker_gamma = parse_double(k, v);
        } else if (k == "simplified") {
            p->simplified = parse_bool(k, v);
        } else if (k == "verbose") {
            p->verbose = parse_bool(k, v);
        } else {
            throw Error(FEM_ERR_ARG, std::string("unknown Newton parameter: ") + key);
        }
    });
}

int fem_newton_create(fem_ctx *ctx, const fem_newton_params *params, const fem_solver_params *linear,
                      fem_newton **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && params && linear && out, FEM_ERR_ARG, "null argument");
        fem_newton *h = new fem_newton();
        try {
            h->n.init(&ctx->c, *params, *linear);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&ctx->c);
        *out = h;
    });
}

int fem_newton_destroy(fem_newton *n)
{
    return guarded([&] {
        if (!n) return;
        Context *c = n->held;
        delete n;
        ctx_release(c);
    });
}

int fem_newton_bind(fem_newton *n, fem_op *op, fem_precond *pre)
{
    return guarded([&] {
        FEM_CHECK(n && op, FEM_ERR_ARG, "null argument");
        n->n.bind(&op->o, pre ? &pre->p : nullptr);
    });
}

int fem_newton_unbind(fem_newton *n)
{
    return guarded([&] {
        FEM_CHECK(n, FEM_ERR_ARG, "null argument");
        n->n.unbind();
    });
}

int fem_newton_solve(fem_newton *n, const fem_vec *u, fem_vec *w, fem_newton_info *info)
{
    return guarded([&] {
        FEM_CHECK(n && w, FEM_ERR_ARG, "null argument");
        fem_newton_info tmp;
        n->n.solve(u ? &u->v : nullptr, w->v, info ? *info : tmp);
    });
}

// ---- learned preconditioner ------------------------------------------------------
int fem_cnn_create(fem_ctx *ctx, int nx, int ny, int nz, const int *dil, fem_cnn **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && out, FEM_ERR_ARG, "null argument");
        const int n[3] = { nx, ny, nz };
        const int d0[CNN_LAYERS] = { 1, 2, 4, 1 };
        fem_cnn *h = new fem_cnn();
        try {
            h->c.init(&ctx->c, n, dil ? dil : d0);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&ctx->c);
        *out = h;
    });
}

int fem_cnn_destroy(fem_cnn *c)
{
    return guarded([&] {
        if (c) {
            (void)hipStreamSynchronize(c->c.ctx->stream);
            Context *cx = c->held;
            delete c;
            ctx_release(cx);
        }
    });
}

int64_t fem_cnn_nparams(const fem_cnn *c) { return c ? (int64_t)c->c.nparams : -1; }

int fem_cnn_set_weights(fem_cnn *c, const float *w)
{
    return guarded([&] {
        FEM_CHECK(c && w, FEM_ERR_ARG, "null argument");
        c->c.set_weights(w);
    });
}

int fem_cnn_get_weights(const fem_cnn *c, float *w)
{
    return guarded([&] {
        FEM_CHECK(c && w, FEM_ERR_ARG, "null argument");
        c->c.get_weights(w);
    });
}

int fem_cnn_forward(fem_cnn *c, const float *x, float *y)
{
    return guarded([&] {
        FEM_CHECK(c && x && y, FEM_ERR_ARG, "null argument");
        CNN &n = c->c;
        hipStream_t s = n.ctx->stream;
        FEM_HIP(hipMemcpyAsync(n.x0.p, x, sizeof(float) * 2 * n.V, hipMemcpyHostToDevice, s));
        n.forward(n.x0.as<float>(), n.y.as<float>(), s);
        FEM_HIP(hipMemcpyAsync(y, n.y.p, sizeof(float) * n.V, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
    });
}

int fem_cnn_loss_grad(fem_cnn *c, const float *x, const float *t, double *loss, float *grad)
{
    return guarded([&] {
        FEM_CHECK(c && x && t, FEM_ERR_ARG, "null argument");
        CNN &n = c->c;
        hipStream_t s = n.ctx->stream;
        DevBuf td;
        td.alloc(sizeof(float) * n.V);
        FEM_HIP(hipMemcpyAsync(n.x0.p, x, sizeof(float) * 2 * n.V, hipMemcpyHostToDevice, s));
        FEM_HIP(hipMemcpyAsync(td.p, t, sizeof(float) * n.V, hipMemcpyHostToDevice, s));
        n.forward(n.x0.as<float>(), n.y.as<float>(), s);
        n.zero_grad(s);
        n.backward_accumulate(td.as<float>(), s);
        if (loss) *loss = n.loss_value(s);
        if (grad) FEM_HIP(hipMemcpyAsync(grad, n.grad.p, sizeof(float) * n.nparams, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
    });
}

int fem_cnn_set_coefficient(fem_cnn *c, const double *kcell, double beta)
{
    return guarded([&] {
        FEM_CHECK(c && kcell && beta > 0, FEM_ERR_ARG, "null argument or beta <= 0");
        cnn_set_coefficient(c->c, kcell, beta);
    });
}

int fem_cnn_coarse_apply(fem_cnn *c, const float *x, float *y)
{
    return guarded([&] {
        FEM_CHECK(c && x && y && c->c.stencil.p, FEM_ERR_ARG, "null argument or coefficient not set");
        CNN &n = c->c;
        hipStream_t s = n.ctx->stream;
        DevBuf a, b;
        a.alloc(sizeof(float) * n.V);
        b.alloc(sizeof(float) * n.V);
        FEM_HIP(hipMemcpyAsync(a.p, x, sizeof(float) * n.V, hipMemcpyHostToDevice, s));
        n.coarse_apply(a.as<float>(), b.as<float>(), s);
        FEM_HIP(hipMemcpyAsync(y, b.p, sizeof(float) * n.V, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
    });
}

int fem_cnn_coarse_solve(fem_cnn *c, const float *b, int steps, float *x, double *rz)
{
    return guarded([&] {
        FEM_CHECK(c && b && x, FEM_ERR_ARG, "null argument");
        CNN &n = c->c;
        hipStream_t s = n.ctx->stream;
        DevBuf db, dx;
        db.alloc(sizeof(float) * n.V);
        dx.alloc(sizeof(float) * n.V);
        FEM_HIP(hipMemcpyAsync(db.p, b, sizeof(float) * n.V, hipMemcpyHostToDevice, s));
        double h[2];
        cnn_coarse_solve(n, db.as<float>(), steps, dx.as<float>(), h);
        FEM_HIP(hipMemcpyAsync(x, dx.p, sizeof(float) * n.V, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (rz) {
            rz[0] = h[0];
            rz[1] = h[1];
        }
    });
}

int fem_cnn_collect_pairs(fem_cnn *c, fem_op *op64, fem_precond *bj64, const fem_vec *b, int max_pairs, int every,
                          int *npairs)
{
    return guarded([&] {
        FEM_CHECK(c && op64 && bj64 && b, FEM_ERR_ARG, "null argument");
        const int n = cnn_collect_pairs(c->c, &op64->o, &bj64->p, b->v, max_pairs, every);
        if (npairs) *npairs = n;
    });
}

int fem_cnn_train(fem_cnn *c, int epochs, int batch, float lr, float val_frac, unsigned seed, double *train_loss,
                  double *val_rel)
{
    return guarded([&] {
        FEM_CHECK(c && epochs >= 0 && batch >= 1, FEM_ERR_ARG, "train arguments");
        double tl = 0, vr = 0;
        cnn_train(c->c, epochs, batch, lr, val_frac, seed, tl, vr);
        if (train_loss) *train_loss = tl;
        if (val_rel) *val_rel = vr;
    });
}

int fem_precond_create_learned_ex(fem_op *op, fem_cnn *cnn, int coarse_steps, int smoother, fem_precond **out)
{
    return guarded([&] {
        FEM_CHECK(op && cnn && out, FEM_ERR_ARG, "null argument");
        fem_precond *h = new fem_precond();
        try {
            h->p.op = &op->o;
            h->p.kind = FEM_PRE_LEARNED;
            h->p.prec = op->o.vec_prec;
            h->p.learned = new Learned();
            h->p.learned->init(&op->o, &cnn->c, coarse_steps, smoother);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&*op->o.ctx);
        *out = h;
    });
}

int fem_precond_create_learned(fem_op *op, fem_cnn *cnn, int coarse_steps, fem_precond **out)
{
    return fem_precond_create_learned_ex(op, cnn, coarse_steps, FEM_PRE_BLOCK_JACOBI, out);
}

int fem_precond_set_learned_mode(fem_precond *p, int mode)
{
    return guarded([&] {
        FEM_CHECK(p && p->p.kind == FEM_PRE_LEARNED, FEM_ERR_ARG, "not a learned preconditioner");
        FEM_CHECK(mode == FEM_LEARNED_OFF || mode == FEM_LEARNED_NET || mode == FEM_LEARNED_COARSE_ONLY, FEM_ERR_ARG,
                  "learned mode: FEM_LEARNED_OFF, FEM_LEARNED_NET or FEM_LEARNED_COARSE_ONLY");
        p->p.learned->mode = mode;
    });
}

// ---- batched solves --------------------------------------------------------------
int fem_batch_create(fem_ctx *ctx, int count, const int *nx, const int *ny, double c, fem_batch **out)
{
    return guarded([&] {
        FEM_CHECK(ctx && nx && ny && out, FEM_ERR_ARG, "null argument");
        fem_batch *h = new fem_batch();
        try {
            h->b.init(&ctx->c, count, nx, ny, c);
        } catch (...) {
            delete h;
            throw;
        }
        h->held = ctx_retain(&ctx->c);
        *out = h;
    });
}

int fem_batch_destroy(fem_batch *b)
{
    return guarded([&] {
        if (b) {
            (void)hipStreamSynchronize(b->b.ctx->stream);
            Context *c = b->held;
            delete b;
            ctx_release(c);
        }
    });
}

int64_t fem_batch_ndofs(const fem_batch *b) { return b ? b->b.N : -1; }

int fem_batch_dof_offsets(const fem_batch *b, int64_t *off)
{
    return guarded([&] {
        FEM_CHECK(b && off, FEM_ERR_ARG, "null argument");
        for (int i = 0; i <= b->b.count; ++i) off[i] = b->b.dof_off[i];
    });
}

int fem_batch_set_kappa_qp(fem_batch *b, const double *kvol)
{
    return guarded([&] {
        FEM_CHECK(b, FEM_ERR_ARG, "null batch");
        b->b.set_kappa_qp(kvol);
    });
}

int fem_batch_apply(fem_batch *b, const fem_vec *u, fem_vec *w)
{
    return guarded([&] {
        FEM_CHECK(b && u && w, FEM_ERR_ARG, "null argument");
        FEM_CHECK(u->v.n == b->b.N && w->v.n == b->b.N && u->v.prec == w->v.prec, FEM_ERR_ARG,
                  "batch apply: vectors of the batch size and one precision");
        FEM_CHECK(u->v.data != w->v.data, FEM_ERR_ARG, "batch apply: u and w must be distinct");
        b->b.apply_raw(u->v.prec, u->v.data, w->v.data, nullptr);
    });
}

int fem_batch_assemble_rhs(fem_batch *b, const double *fq, fem_vec *F)
{
    return guarded([&] {
        FEM_CHECK(b && fq && F, FEM_ERR_ARG, "null argument");
        b->b.rhs(fq, F->v);
    });
}

int fem_batch_autotune(fem_batch *b, int *tx, int *ty, float *ms)
{
    return guarded([&] {
        FEM_CHECK(b, FEM_ERR_ARG, "null batch");
        FEM_CHECK(b->b.kvol32.p, FEM_ERR_STATE, "batch: coefficient data not set");
        const int sh = b->b.autotune(10);
        if (tx) *tx = CG_SHAPES[sh][0];
        if (ty) *ty = CG_SHAPES[sh][1];
        if (ms)
            for (int i = 0; i < CG_NSHAPES; ++i) ms[i] = b->b.tune_ms[i];
    });
}

int fem_batch_solve(fem_batch *b, const fem_vec *rhs, fem_vec *x, double tol, double inner_tol, int inner_maxit,
                    int max_outer, int use_graph, fem_batch_info *info)
{
    return guarded([&] {
        FEM_CHECK(b && rhs && x, FEM_ERR_ARG, "null argument");
        fem_batch_info tmp;
        b->b.solve(rhs->v, x->v, tol, inner_tol, inner_maxit, max_outer, use_graph, info ? *info : tmp);
    });
}

int fem_batch_instance_results(const fem_batch *b, int *iterations, double *rel_residual)
{
    return guarded([&] {
        FEM_CHECK(b, FEM_ERR_ARG, "null batch");
        FEM_CHECK(!b->b.last_iters.empty(), FEM_ERR_STATE, "no solve yet");
        for (int i = 0; i < b->b.count; ++i) {
            if (iterations) iterations[i] = b->b.last_iters[i];
            if (rel_residual) rel_residual[i] = b->b.last_relres[i];
        }
    });
}

int fem_gauss_points(int m, double *x, double *w)
{
    return guarded([&] {
        FEM_CHECK(m >= 1 && m <= 32 && x && w, FEM_ERR_ARG, "gauss: 1 <= m <= 32");
        gauss_legendre01(m, x, w);
    });
}

}  // extern "C"
