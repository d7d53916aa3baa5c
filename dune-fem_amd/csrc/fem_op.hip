// Operator and preconditioner objects: set-up of coefficient data and
// dispatch to the DG / CG kernels.
#include <dlfcn.h>

#include <cmath>

#include "fem_blas.hpp"
#include "fem_learned.hpp"
#include "fem_mg.hpp"
#include "fem_op.hpp"

namespace femamd {

template <typename T>
__global__ void k_jacobi_inv(const double *__restrict__ d, T *__restrict__ dinv, int64_t n)
{
    // near-zero diagonal entries become 1 (fempreconditioning.hh:99-106)
    const double eps = 16.0 * 2.220446049250313e-16;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dinv[i] = (T)(fabs(d[i]) < eps ? 1.0 : 1.0 / d[i]);
}

template <typename T>
__global__ void k_pointwise_mul(const T *__restrict__ a, const T *__restrict__ x, T *__restrict__ y, int64_t n,
                                const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a[i] * x[i];
}

template <typename T>
static void upload_as(DevBuf &b, const double *host, size_t n, hipStream_t s)
{
    b.alloc(n * sizeof(T));
    if (std::is_same<T, double>::value) {
        FEM_HIP(hipMemcpyAsync(b.p, host, n * sizeof(double), hipMemcpyHostToDevice, s));
        FEM_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
        FEM_HIP(hipMemcpyAsync(b.p, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
        FEM_HIP(hipStreamSynchronize(s));
    }
}

void Operator::init(Context *c, const fem_op_desc &d)
{
    ctx = c;
    desc = d;
    FEM_CHECK(d.precision == FEM_F64 || d.precision == FEM_F32 || d.precision == FEM_BF16MFMA, FEM_ERR_ARG,
              "operator precision");
    vec_prec = d.precision == FEM_F64 ? FEM_F64 : FEM_F32;
    if (d.kind == FEM_OP_DG_SIPG_3D) {
        FEM_CHECK(d.order >= 1 && d.order <= 4, FEM_ERR_UNSUPPORTED, "DG order must be 1..4");
        FEM_CHECK(d.n[0] > 0 && d.n[1] > 0 && d.n[2] > 0, FEM_ERR_ARG, "mesh size");
        FEM_CHECK(d.kappa_mode >= 0 && d.kappa_mode <= 2, FEM_ERR_ARG, "kappa mode");
        k = d.order + 1;
        E = (int64_t)d.n[0] * d.n[1] * d.n[2];
        N = E * k * k * k;
        Tab1D<double> t = make_tab1d(k);
        Tab1D<float> tf = to_float(t);
        tab64.alloc(sizeof(t));
        tab32.alloc(sizeof(tf));
        FEM_HIP(hipMemcpy(tab64.p, &t, sizeof(t), hipMemcpyHostToDevice));
        FEM_HIP(hipMemcpy(tab32.p, &tf, sizeof(tf), hipMemcpyHostToDevice));
        for (int i = 0; i < 3; ++i) {
            a64.n[i] = a32.n[i] = d.n[i];
            a64.h[i] = 1.0 / d.n[i];
            a32.h[i] = (float)(1.0 / d.n[i]);
            a64.ih[i] = (double)d.n[i];
            a32.ih[i] = (float)d.n[i];
            a64.b[i] = d.b[i];
            a32.b[i] = (float)d.b[i];
        }
        a64.beta = d.beta;
        a32.beta = (float)d.beta;
        a64.c = d.c;
        a32.c = (float)d.c;
        a64.kappa_mode = a32.kappa_mode = d.kappa_mode == FEM_KAPPA_CONST ? FEM_KAPPA_CONST : -1;  // set on data upload
        a64.kappa_const = d.kappa_const;
        a32.kappa_const = (float)d.kappa_const;
        a64.tab = tab64.as<Tab1D<double>>();
        a32.tab = tab32.as<Tab1D<float>>();
        dg_derived(a64, k);
        dg_derived(a32, k);
    } else if (d.kind == FEM_OP_CG_Q2_2D) {
        FEM_CHECK(d.order == 2, FEM_ERR_UNSUPPORTED, "CG operator supports order 2");
        FEM_CHECK(d.n[0] > 0 && d.n[1] > 0, FEM_ERR_ARG, "mesh size");
        FEM_CHECK(d.kappa_mode == FEM_KAPPA_CONST || d.kappa_mode == FEM_KAPPA_QP, FEM_ERR_ARG, "kappa mode");
        k = 3;
        E = (int64_t)d.n[0] * d.n[1];
        N = (int64_t)(2 * d.n[0] + 1) * (2 * d.n[1] + 1);
        inst.nx = d.n[0];
        inst.ny = d.n[1];
        inst.dof_off = 0;
        inst.elem_off = 0;
        inst.tiles_x = (d.n[0] + CG_TX - 1) / CG_TX;
        inst.ntiles = inst.tiles_x * ((d.n[1] + CG_TY - 1) / CG_TY);
        inst.tile_off = 0;
        inst.c = d.c;
        inst.kappa_const = d.kappa_const;
        tiles = inst.ntiles;
        inst_dev.alloc(sizeof(CGInst));
        FEM_HIP(hipMemcpy(inst_dev.p, &inst, sizeof(CGInst), hipMemcpyHostToDevice));
        CGTab t = make_cg_tab();
        cgtab.alloc(sizeof(t));
        FEM_HIP(hipMemcpy(cgtab.p, &t, sizeof(t), hipMemcpyHostToDevice));
    } else {
        throw Error(FEM_ERR_ARG, "unknown operator kind");
    }
}

void Operator::set_kappa_qp(const double *kvol, const double *kx, const double *ky, const double *kz)
{
    hipStream_t s = ctx->stream;
    if (is_dg()) {
        FEM_CHECK(desc.kappa_mode == FEM_KAPPA_QP, FEM_ERR_STATE, "operator not created with FEM_KAPPA_QP");
        FEM_CHECK(kvol && kx && ky && kz, FEM_ERR_ARG, "null kappa array");
        const int k2 = k * k;
        const size_t nv = (size_t)E * k * k2;
        const double *kf[3] = { kx, ky, kz };
        size_t nf[3];
        for (int d = 0; d < 3; ++d)
            nf[d] = (size_t)(desc.n[0] + (d == 0)) * (desc.n[1] + (d == 1)) * (desc.n[2] + (d == 2)) * k2;
        if (vec_prec == FEM_F64) {
            upload_as<double>(kvol64, kvol, nv, s);
            for (int d = 0; d < 3; ++d) upload_as<double>(kface64[d], kf[d], nf[d], s);
        }
        upload_as<float>(kvol32, kvol, nv, s);
        for (int d = 0; d < 3; ++d) upload_as<float>(kface32[d], kf[d], nf[d], s);
        if (vec_prec != FEM_F64) {
            // FP64 copies are kept only for FP64 operators; the diagonal and
            // block set-up of FP32 operators read the FP32 data widened
            upload_as<double>(kvol64, kvol, nv, s);
            for (int d = 0; d < 3; ++d) upload_as<double>(kface64[d], kf[d], nf[d], s);
        }
        a64.kvol = kvol64.as<double>();
        a32.kvol = kvol32.as<float>();
        for (int d = 0; d < 3; ++d) {
            a64.kface[d] = kface64[d].as<double>();
            a32.kface[d] = kface32[d].as<float>();
        }
        a64.kappa_mode = a32.kappa_mode = FEM_KAPPA_QP;
    } else {
        FEM_CHECK(desc.kappa_mode == FEM_KAPPA_QP, FEM_ERR_STATE, "operator not created with FEM_KAPPA_QP");
        FEM_CHECK(kvol, FEM_ERR_ARG, "null kappa array");
        upload_as<double>(kvol64, kvol, (size_t)E * 9, s);
        upload_as<float>(kvol32, kvol, (size_t)E * 9, s);
    }
}

void Operator::set_kappa_cell(const double *kc)
{
    FEM_CHECK(is_dg() && desc.kappa_mode == FEM_KAPPA_CELL, FEM_ERR_STATE, "operator not created with FEM_KAPPA_CELL");
    FEM_CHECK(kc, FEM_ERR_ARG, "null kappa array");
    upload_as<double>(kcell64, kc, (size_t)E, ctx->stream);
    upload_as<float>(kcell32, kc, (size_t)E, ctx->stream);
    a64.kcell = kcell64.as<double>();
    a32.kcell = kcell32.as<float>();
    a64.kappa_mode = a32.kappa_mode = FEM_KAPPA_CELL;
}

// ---- reaction beyond the constant c: pointwise and nonlinear ---------------------
static void copy_buf(DevBuf &dst, const DevBuf &src, hipStream_t s)
{
    dst.free();
    if (!src.p) return;
    dst.alloc(src.bytes);
    FEM_HIP(hipMemcpyAsync(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice, s));
}

// a raw device array as a (non-owning) vector
static Vec raw_vec(Context *c, void *p, int64_t n, int prec)
{
    Vec v;
    v.ctx = c; v.n = n; v.prec = prec; v.data = p; v.owner = false;
    return v;
}

static void reaction_checks(const Operator &o)
{
    FEM_CHECK(o.is_dg() || o.desc.kind == FEM_OP_CG_Q2_2D, FEM_ERR_UNSUPPORTED,
              "pointwise/nonlinear reaction: DG or Q2 Lagrange operators");
    FEM_CHECK(o.desc.precision != FEM_BF16MFMA, FEM_ERR_UNSUPPORTED,
              "pointwise/nonlinear reaction: FP64 or FP32 operators");
}

void Operator::set_reaction_qp(const double *cvol)
{
    reaction_checks(*this);
    if (!cvol) {
        cvol64.free();
        cvol32.free();
    } else {
        const size_t nq = (size_t)nqp();
        upload_as<double>(cvol64, cvol, nq, ctx->stream);
        upload_as<float>(cvol32, cvol, nq, ctx->stream);
    }
    a64.cvol = cvol64.as<double>();
    a32.cvol = cvol32.as<float>();
}

Operator::~Operator()
{
    if (plugin_dl) dlclose(plugin_dl);
}

fem_qp_geometry Operator::qp_geometry() const
{
    fem_qp_geometry g{};
    g.dim = is_dg() ? 3 : 2;
    for (int d = 0; d < 3; ++d) {
        g.n[d] = d < g.dim ? desc.n[d] : 1;
        g.h[d] = d < g.dim ? 1.0 / desc.n[d] : 0.0;
    }
    g.m = k;
    double w[KMAX];
    gauss_legendre01(k, g.gx, w);
    g.nq = nqp();
    return g;
}

void Operator::set_reaction_plugin(const char *path)
{
    reaction_checks(*this);
    FEM_CHECK(desc.precision == FEM_F64, FEM_ERR_UNSUPPORTED, "compiled reaction: FP64 operators");
    if (plugin_dl) {
        dlclose(plugin_dl);
        plugin_dl = nullptr;
        plugin = nullptr;
    }
    pq.free();
    ps.free();
    if (!path) return;
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    FEM_CHECK(h, FEM_ERR_ARG, std::string("compiled reaction: cannot load ") + path + ": " + dlerror());
    auto fn = (fem_reaction_eval_fn)dlsym(h, "fem_reaction_eval");
    if (!fn) {
        dlclose(h);
        throw Error(FEM_ERR_ARG, std::string("compiled reaction: no fem_reaction_eval in ") + path);
    }
    const size_t nq = (size_t)nqp();
    pq.alloc(nq * sizeof(double));
    ps.alloc(nq * sizeof(double));
    plugin_dl = h;
    plugin = fn;
}

void Operator::set_reaction_law(double coef, int pw)
{
    reaction_checks(*this);
    FEM_CHECK(coef == 0 || (pw >= 1 && pw <= 4), FEM_ERR_ARG, "reaction law: power in 1..4");
    nl_coef = coef;
    nl_pow = coef == 0 ? 0 : pw;
    a64.nl_coef = nl_coef;
    a32.nl_coef = (float)nl_coef;
    a64.nl_pow = a32.nl_pow = nl_pow;
}

void Operator::eval_qp(const Vec &u, Vec &out) const
{
    FEM_CHECK(u.n == N && out.n == nqp(), FEM_ERR_ARG, "quadrature-point evaluation: vector sizes");
    FEM_CHECK(u.prec == vec_prec && out.prec == vec_prec, FEM_ERR_ARG,
              "quadrature-point evaluation: vectors of the operator's precision");
    if (!is_dg()) {
        if (vec_prec == FEM_F64)
            cg_eval_qp<double>(desc.n[0], desc.n[1], cgtab.as<CGTab>(), (const double *)u.data, nullptr, 1.0, 1,
                               (double *)out.data, ctx->stream);
        else
            cg_eval_qp<float>(desc.n[0], desc.n[1], cgtab.as<CGTab>(), (const float *)u.data, nullptr, 1.0f, 1,
                              (float *)out.data, ctx->stream);
        return;
    }
    if (vec_prec == FEM_F64)
        dg_eval_qp<double>(a64, k, (const double *)u.data, nullptr, 1.0, 1, (double *)out.data, ctx->stream);
    else
        dg_eval_qp<float>(a32, k, (const float *)u.data, nullptr, 1.0f, 1, (float *)out.data, ctx->stream);
}

void Operator::make_jacobian(Operator &jop, int precision) const
{
    reaction_checks(*this);
    FEM_CHECK(precision < 0 || precision == desc.precision || (desc.precision == FEM_F64 && precision == FEM_F32),
              FEM_ERR_ARG, "jacobian operator: the operator's precision or FP32 for an FP64 operator");
    hipStream_t s = ctx->stream;
    fem_op_desc d = desc;
    if (precision >= 0) d.precision = precision;
    jop.init(ctx, d);
    for (int d = 0; d < 3; ++d) {
        copy_buf(jop.kface64[d], kface64[d], s);
        copy_buf(jop.kface32[d], kface32[d], s);
        copy_buf(jop.gbnd64[d], gbnd64[d], s);
    }
    copy_buf(jop.kvol64, kvol64, s);
    copy_buf(jop.kvol32, kvol32, s);
    copy_buf(jop.kcell64, kcell64, s);
    copy_buf(jop.kcell32, kcell32, s);
    copy_buf(jop.cvol64, cvol64, s);
    copy_buf(jop.cvol32, cvol32, s);
    jop.has_bnd = has_bnd;
    jop.a64.kappa_mode = a64.kappa_mode;
    jop.a32.kappa_mode = a32.kappa_mode;
    // the kernel arguments point at the copies (pointwise kappa: volume and
    // face values)
    jop.a64.kcell = jop.kcell64.as<double>();
    jop.a32.kcell = jop.kcell32.as<float>();
    jop.a64.kvol = jop.kvol64.as<double>();
    jop.a32.kvol = jop.kvol32.as<float>();
    for (int d = 0; d < 3; ++d) {
        jop.a64.kface[d] = jop.kface64[d].as<double>();
        jop.a32.kface[d] = jop.kface32[d].as<float>();
    }
    jop.a64.cvol = jop.cvol64.as<double>();
    jop.a32.cvol = jop.cvol32.as<float>();
    FEM_HIP(hipStreamSynchronize(s));
}

void Operator::jacobian(const Vec &u, Operator &jop) const
{
    reaction_checks(*this);
    FEM_CHECK(jop.desc.kind == desc.kind &&
                  (jop.desc.precision == desc.precision || (desc.precision == FEM_F64 && jop.desc.precision == FEM_F32)) &&
                  jop.k == k && jop.E == E &&
                  jop.desc.n[0] == desc.n[0] && jop.desc.n[1] == desc.n[1] && jop.desc.n[2] == desc.n[2] &&
                  jop.desc.kappa_mode == desc.kappa_mode && jop.a64.kappa_mode == a64.kappa_mode,
              FEM_ERR_ARG, "jacobian: the Jacobian operator must be made from this operator");
    FEM_CHECK(&jop != this, FEM_ERR_ARG, "jacobian: an operator other than this one");
    FEM_CHECK(u.n == N && u.prec == vec_prec, FEM_ERR_ARG, "jacobian: vector size or precision");
    hipStream_t s = ctx->stream;
    const int64_t nq = nqp();
    FEM_CHECK(!jop.plugin, FEM_ERR_ARG, "jacobian: the Jacobian operator must be linear");
    // out = base + coef u^pw at the quadrature points, by the operator's kind
    auto eval64 = [&](const double *uu, const double *base, double coef, int pw, double *out) {
        if (is_dg()) dg_eval_qp<double>(a64, k, uu, base, coef, pw, out, s);
        else cg_eval_qp<double>(desc.n[0], desc.n[1], cgtab.as<CGTab>(), uu, base, coef, pw, out, s);
    };
    auto eval32 = [&](const float *uu, const float *base, float coef, int pw, float *out) {
        if (is_dg()) dg_eval_qp<float>(a32, k, uu, base, coef, pw, out, s);
        else cg_eval_qp<float>(desc.n[0], desc.n[1], cgtab.as<CGTab>(), uu, base, coef, pw, out, s);
    };
    jop.set_reaction_law(0.0, 0);
    if (nl_pow == 0) {
        // no law: the linear part's cvol, if any
        copy_buf(jop.cvol64, cvol64, s);
        copy_buf(jop.cvol32, cvol32, s);
    } else {
        // d/du (nl_coef u^pw u) = (pw + 1) nl_coef u^pw
        if (!jop.cvol64.p || jop.cvol64.bytes != (size_t)nq * 8) jop.cvol64.alloc((size_t)nq * 8);
        if (!jop.cvol32.p || jop.cvol32.bytes != (size_t)nq * 4) jop.cvol32.alloc((size_t)nq * 4);
        const double cj = (nl_pow + 1) * nl_coef;
        if (vec_prec == FEM_F64) {
            eval64((const double *)u.data, cvol64.as<double>(), cj, nl_pow, jop.cvol64.as<double>());
            Vec lo = raw_vec(ctx, jop.cvol32.p, nq, FEM_F32), hi = raw_vec(ctx, jop.cvol64.p, nq, FEM_F64);
            vec_copy(lo, hi);
        } else {
            eval32((const float *)u.data, cvol32.as<float>(), (float)cj, nl_pow, jop.cvol32.as<float>());
            Vec lo = raw_vec(ctx, jop.cvol32.p, nq, FEM_F32), hi = raw_vec(ctx, jop.cvol64.p, nq, FEM_F64);
            vec_copy(hi, lo);
        }
    }
    if (plugin) {
        // + dr/du(u) from the compiled reaction
        eval64((const double *)u.data, nullptr, 1.0, 1, pq.as<double>());
        const fem_qp_geometry g = qp_geometry();
        FEM_CHECK(plugin(&g, pq.as<double>(), nullptr, ps.as<double>(), (void *)s) == 0, FEM_ERR_HIP,
                  "compiled reaction: launch failed");
        Vec dr = raw_vec(ctx, ps.p, nq, FEM_F64);
        if (jop.cvol64.p) {
            Vec c = raw_vec(ctx, jop.cvol64.p, nq, FEM_F64);
            vec_axpy(c, 1.0, dr);
        } else {
            jop.cvol64.alloc((size_t)nq * 8);
            jop.cvol32.alloc((size_t)nq * 4);
            Vec c = raw_vec(ctx, jop.cvol64.p, nq, FEM_F64);
            vec_copy(c, dr);
        }
        Vec lo = raw_vec(ctx, jop.cvol32.p, nq, FEM_F32), hi = raw_vec(ctx, jop.cvol64.p, nq, FEM_F64);
        vec_copy(lo, hi);
    }
    jop.a64.cvol = jop.cvol64.as<double>();
    jop.a32.cvol = jop.cvol32.as<float>();
    FEM_HIP(hipStreamSynchronize(s));
}

// Dirichlet nodes of the (2 nx + 1) x (2 ny + 1) Q2 node lattice
__device__ __forceinline__ bool q2_boundary(int64_t i, int nx, int ny)
{
    const int64_t W = 2 * (int64_t)nx + 1;
    const int64_t I = i % W, J = i / W;
    return I == 0 || J == 0 || I == 2 * nx || J == 2 * ny;
}

// x_D = (add ? x_D : 0) + g_D on the Dirichlet nodes, interior untouched
template <typename T>
__global__ void k_q2_set_bnd(T *__restrict__ x, const double *__restrict__ g, int64_t n, int nx, int ny, int add)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (q2_boundary(i, nx, ny)) x[i] = (T)((add ? (double)x[i] : 0.0) + g[i]);
}

// x_D = g_D + r_D (r may be null)
template <typename T>
__global__ void k_q2_constrain(T *__restrict__ x, const double *__restrict__ g, const T *__restrict__ r, int64_t n,
                               int nx, int ny)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (q2_boundary(i, nx, ny)) x[i] = (T)((g ? g[i] : 0.0) + (r ? (double)r[i] : 0.0));
}

void Operator::set_boundary_data(const double *g0, const double *g1, const double *g2)
{
    hipStream_t s = ctx->stream;
    if (is_dg()) {
        FEM_CHECK(g0 && g1 && g2, FEM_ERR_ARG, "boundary data: three face arrays");
        const double *g[3] = { g0, g1, g2 };
        for (int d = 0; d < 3; ++d) {
            const size_t nf = (size_t)(desc.n[0] + (d == 0)) * (desc.n[1] + (d == 1)) * (desc.n[2] + (d == 2)) * k * k;
            upload_as<double>(gbnd64[d], g[d], nf, s);
        }
    } else {
        FEM_CHECK(g0, FEM_ERR_ARG, "boundary data: nodal array");
        upload_as<double>(gbnd64[0], g0, (size_t)N, s);
    }
    has_bnd = true;
}

void Operator::constrain(Vec &x, const Vec *add) const
{
    FEM_CHECK(x.n == N && (!add || add->n == N), FEM_ERR_ARG, "constrain: vector size");
    FEM_CHECK(!add || add->prec == x.prec, FEM_ERR_ARG, "constrain: vector precisions differ");
    if (is_dg()) return;   // weak boundary conditions: nothing to set
    hipStream_t s = ctx->stream;
    const double *g = has_bnd ? gbnd64[0].as<double>() : nullptr;
    if (x.prec == FEM_F64)
        hipLaunchKernelGGL(k_q2_constrain<double>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, s, (double *)x.data, g,
                           add ? (const double *)add->data : nullptr, N, desc.n[0], desc.n[1]);
    else
        hipLaunchKernelGGL(k_q2_constrain<float>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, s, (float *)x.data, g,
                           add ? (const float *)add->data : nullptr, N, desc.n[0], desc.n[1]);
    FEM_HIP(hipGetLastError());
}

int64_t Operator::dot_blocks() const
{
    return is_dg() ? dg_blocks(k, desc.n) : 0;
}

void Operator::apply_raw(const void *x, void *y, const int *skip, double *dot) const
{
    hipStream_t s = ctx->stream;
    if (is_dg()) {
        FEM_CHECK(a64.kappa_mode >= 0, FEM_ERR_STATE, "coefficient data not set");
        // every coefficient array the kernel of this mode reads is in place
        const bool f64 = desc.precision == FEM_F64;
        const bool qp_ok = f64 ? (a64.kvol && a64.kface[0] && a64.kface[1] && a64.kface[2])
                               : (a32.kvol && a32.kface[0] && a32.kface[1] && a32.kface[2]);
        FEM_CHECK(a64.kappa_mode != FEM_KAPPA_QP || qp_ok, FEM_ERR_STATE, "pointwise kappa data not set");
        FEM_CHECK(a64.kappa_mode != FEM_KAPPA_CELL || (f64 ? a64.kcell != nullptr : a32.kcell != nullptr),
                  FEM_ERR_STATE, "cellwise kappa data not set");
        FEM_CHECK(!dot || desc.precision != FEM_BF16MFMA, FEM_ERR_UNSUPPORTED, "fused dot: FP64 or FP32 operator");
        FEM_CHECK(!plugin || (!dot && !skip), FEM_ERR_STATE,
                  "apply: an operator with a compiled reaction is nonlinear (no fused dot, no captured solve)");
        if (desc.precision == FEM_F64) dg_apply_f64(a64, k, (const double *)x, (double *)y, skip, s, dot);
        else if (desc.precision == FEM_F32) dg_apply_f32(a32, k, (const float *)x, (float *)y, skip, s, dot);
        else dg_apply_bf16(a32, k, (const float *)x, (float *)y, skip, s);
        if (plugin) {
            // y += int r(u, x) v: u at the points, r by the compiled code,
            // then the load-vector kernel accumulating into y
            dg_eval_qp<double>(a64, k, (const double *)x, nullptr, 1.0, 1, pq.as<double>(), s);
            const fem_qp_geometry g = qp_geometry();
            FEM_CHECK(plugin(&g, pq.as<double>(), ps.as<double>(), nullptr, (void *)s) == 0, FEM_ERR_HIP,
                      "compiled reaction: launch failed");
            dg_rhs(a64, k, ps.as<double>(), (double *)y, s, true);
        }
    } else {
        FEM_CHECK(!dot, FEM_ERR_UNSUPPORTED, "fused dot: DG operators only");
        FEM_CHECK(desc.kappa_mode == FEM_KAPPA_CONST || kvol64.p, FEM_ERR_STATE, "coefficient data not set");
        FEM_CHECK(!plugin || !skip, FEM_ERR_STATE,
                  "apply: an operator with a compiled reaction is nonlinear (no captured solve)");
        if (vec_prec == FEM_F64) {
            CGReaction<double> rx;
            rx.cvol = cvol64.as<double>();
            rx.coef = nl_coef;
            rx.pw = nl_pow;
            cg_launch<double>(CG_MODE_APPLY, inst_dev.as<CGInst>(), 1, tiles, cgtab.as<CGTab>(),
                              desc.kappa_mode == FEM_KAPPA_QP ? kvol64.as<double>() : nullptr, (const double *)x,
                              (double *)y, skip, s, rx);
        } else {
            CGReaction<float> rx;
            rx.cvol = cvol32.as<float>();
            rx.coef = (float)nl_coef;
            rx.pw = nl_pow;
            cg_launch<float>(CG_MODE_APPLY, inst_dev.as<CGInst>(), 1, tiles, cgtab.as<CGTab>(),
                             desc.kappa_mode == FEM_KAPPA_QP ? kvol32.as<float>() : nullptr, (const float *)x,
                             (float *)y, skip, s, rx);
        }
        if (plugin) {
            // y += int r(u, x) v on the interior rows (the Dirichlet rows
            // keep y_D = u_D)
            cg_eval_qp<double>(desc.n[0], desc.n[1], cgtab.as<CGTab>(), (const double *)x, nullptr, 1.0, 1,
                               pq.as<double>(), s);
            const fem_qp_geometry g = qp_geometry();
            FEM_CHECK(plugin(&g, pq.as<double>(), ps.as<double>(), nullptr, (void *)s) == 0, FEM_ERR_HIP,
                      "compiled reaction: launch failed");
            cg_launch<double>(CG_MODE_RHS_ADD, inst_dev.as<CGInst>(), 1, tiles, cgtab.as<CGTab>(), ps.as<double>(),
                              nullptr, (double *)y, nullptr, s);
        }
    }
}

void Operator::apply(const Vec &x, Vec &y, const int *skip) const
{
    FEM_CHECK(x.n == N && y.n == N, FEM_ERR_ARG, "apply: vector size does not match the operator");
    FEM_CHECK(x.prec == vec_prec && y.prec == vec_prec, FEM_ERR_ARG, "apply: vector precision does not match the operator");
    FEM_CHECK(x.data != y.data, FEM_ERR_ARG, "apply: u and w must be distinct");
    apply_raw(x.data, y.data, skip);
}

void Operator::rhs(const double *fq_host, Vec &F) const
{
    FEM_CHECK(F.n == N, FEM_ERR_ARG, "rhs: vector size");
    hipStream_t s = ctx->stream;
    DevBuf fq, out;
    const size_t nq = is_dg() ? (size_t)E * k * k * k : (size_t)E * 9;
    upload_as<double>(fq, fq_host, nq, s);
    double *dst;
    if (F.prec == FEM_F64) dst = (double *)F.data;
    else { out.alloc(N * sizeof(double)); dst = out.as<double>(); }
    if (is_dg()) {
        dg_rhs(a64, k, fq.as<double>(), dst, s);
        if (has_bnd) {
            const double *g[3] = { gbnd64[0].as<double>(), gbnd64[1].as<double>(), gbnd64[2].as<double>() };
            dg_rhs_bnd(a64, k, g, dst, s);
        }
    } else {
        cg_launch<double>(CG_MODE_RHS, inst_dev.as<CGInst>(), 1, tiles, cgtab.as<CGTab>(), fq.as<double>(), nullptr, dst, nullptr, s);
        // R(u) = A u - F has Dirichlet rows u_D - g_D, as the reference's
        // DirichletWrapperOperator (dirichletwrapper.hh:101-105 with
        // subConstraints, dirichletconstraints.hh:494-554): F_D = g_D
        if (has_bnd)
            hipLaunchKernelGGL(k_q2_set_bnd<double>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, s, dst,
                               gbnd64[0].as<double>(), N, desc.n[0], desc.n[1], 1);
    }
    if (F.prec != FEM_F64) {
        Vec tmp;
        tmp.ctx = ctx; tmp.n = N; tmp.prec = FEM_F64; tmp.data = dst; tmp.owner = false;
        vec_copy(F, tmp);
    }
    FEM_HIP(hipStreamSynchronize(s));
}

void Operator::diagonal(Vec &d) const
{
    FEM_CHECK(d.n == N, FEM_ERR_ARG, "diagonal: vector size");
    FEM_CHECK(!nonlinear(), FEM_ERR_STATE, "diagonal: the operator is nonlinear (take its Jacobian)");
    hipStream_t s = ctx->stream;
    DevBuf out;
    double *dst;
    if (d.prec == FEM_F64) dst = (double *)d.data;
    else { out.alloc(N * sizeof(double)); dst = out.as<double>(); }
    if (is_dg()) {
        FEM_CHECK(a64.kappa_mode >= 0, FEM_ERR_STATE, "coefficient data not set");
        dg_diagonal(a64, k, dst, s);
    } else {
        CGReaction<double> rx;
        rx.cvol = cvol64.as<double>();
        cg_launch<double>(CG_MODE_DIAG, inst_dev.as<CGInst>(), 1, tiles, cgtab.as<CGTab>(),
                          desc.kappa_mode == FEM_KAPPA_QP ? kvol64.as<double>() : nullptr, nullptr, dst, nullptr, s, rx);
    }
    if (d.prec != FEM_F64) {
        Vec tmp;
        tmp.ctx = ctx; tmp.n = N; tmp.prec = FEM_F64; tmp.data = dst; tmp.owner = false;
        vec_copy(d, tmp);
    }
    FEM_HIP(hipStreamSynchronize(s));
}

// ---- element blocks by coloured probing -------------------------------------------
// The SIPG operator couples an element only with its six face neighbours.
// With elements coloured by the parity of their grid coordinates (8
// colours), no two elements of one colour are face neighbours, so applying
// the operator to the probe "phi_j on every element of colour c" gives, on
// each element e of colour c, column j of its diagonal block A_ee and
// nothing else.  8 nb applies of the production kernel yield every block,
// for every coefficient mode and with advection.
template <typename T>
__global__ void k_probe_fill(T *__restrict__ p, int nx, int ny, int64_t E, int nb, int j, int col, int ncol)
{
    const int64_t n = E * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / nb;
        const int l = (int)(i - e * nb);
        const int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
        const int c = (ex % ncol) + ncol * ((ey % ncol) + ncol * (ez % ncol));
        p[i] = (l == j && c == col) ? T(1) : T(0);
    }
}

template <typename T>
__global__ void k_probe_diag_extract(const T *__restrict__ w, double *__restrict__ blocks, int nx, int ny, int64_t E,
                                     int nb, int j, int col)
{
    const int64_t n = E * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / nb;
        const int r = (int)(i - e * nb);
        const int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
        const int c = (ex & 1) + 2 * ((ey & 1) + 2 * (ez & 1));
        if (c == col) blocks[(e * nb + r) * nb + j] = (double)w[i];
    }
}

void Operator::block_diagonal(Vec &blocks) const
{
    FEM_CHECK(is_dg(), FEM_ERR_UNSUPPORTED, "block diagonal: DG operators only");
    FEM_CHECK(a64.kappa_mode >= 0, FEM_ERR_STATE, "coefficient data not set");
    FEM_CHECK(!nonlinear(), FEM_ERR_STATE, "block diagonal: the operator is nonlinear (take its Jacobian)");
    const int nb = k * k * k;
    FEM_CHECK(blocks.prec == FEM_F64 && blocks.n == E * nb * nb, FEM_ERR_ARG,
              "block diagonal: an FP64 vector of E nb^2 entries");
    hipStream_t s = ctx->stream;
    DevBuf pb, wb;
    const size_t es = vec_prec == FEM_F64 ? 8 : 4;
    pb.alloc(N * es);
    wb.alloc(N * es);
    const unsigned G = grid_for(N, 256, 8192);
    for (int col = 0; col < 8; ++col)
        for (int j = 0; j < nb; ++j) {
            if (vec_prec == FEM_F64) {
                hipLaunchKernelGGL(k_probe_fill<double>, dim3(G), dim3(256), 0, s, pb.as<double>(), desc.n[0], desc.n[1], E, nb, j, col, 2);
                apply_raw(pb.p, wb.p, nullptr);
                hipLaunchKernelGGL(k_probe_diag_extract<double>, dim3(G), dim3(256), 0, s, wb.as<double>(), (double *)blocks.data, desc.n[0], desc.n[1], E, nb, j, col);
            } else {
                hipLaunchKernelGGL(k_probe_fill<float>, dim3(G), dim3(256), 0, s, pb.as<float>(), desc.n[0], desc.n[1], E, nb, j, col, 2);
                apply_raw(pb.p, wb.p, nullptr);
                hipLaunchKernelGGL(k_probe_diag_extract<float>, dim3(G), dim3(256), 0, s, wb.as<float>(), (double *)blocks.data, desc.n[0], desc.n[1], E, nb, j, col);
            }
            FEM_HIP(hipGetLastError());
        }
    FEM_HIP(hipStreamSynchronize(s));
}

static bool same_desc(const fem_op_desc &a, const fem_op_desc &b)
{
    return a.kind == b.kind && a.n[0] == b.n[0] && a.n[1] == b.n[1] && a.n[2] == b.n[2] && a.order == b.order &&
           a.beta == b.beta && a.b[0] == b.b[0] && a.b[1] == b.b[1] && a.b[2] == b.b[2] && a.c == b.c &&
           a.kappa_mode == b.kappa_mode && a.kappa_const == b.kappa_const && a.precision == b.precision;
}

void Precond::init(Operator *o, int k_, bool bound_)
{
    op = o;
    kind = k_;
    prec = o->vec_prec;
    bound = bound_;
    desc = o->desc;
    if (kind == FEM_PRE_JACOBI) {
        update();
    } else if (kind == FEM_PRE_BLOCK_JACOBI) {
        FEM_CHECK(o->is_dg(), FEM_ERR_UNSUPPORTED, "block Jacobi needs a DG operator");
        FEM_CHECK(o->desc.b[0] == 0 && o->desc.b[1] == 0 && o->desc.b[2] == 0, FEM_ERR_UNSUPPORTED,
                  "block Jacobi: operators without advection (b = 0)");
        const size_t es = prec == FEM_F64 ? 8 : 4;
        Q.alloc((size_t)o->E * 3 * o->k * o->k * es);
        lam.alloc((size_t)o->E * 3 * o->k * es);
        update();
    } else if (kind == FEM_PRE_MULTIGRID) {
        mg = new Multigrid();
        mg->init(o);
    } else if (kind == FEM_PRE_NONE) {
    } else {
        throw Error(FEM_ERR_UNSUPPORTED, "preconditioner kind (the learned one is made by fem_precond_create_learned)");
    }
}

void Precond::update()
{
    Operator *o = op;
    hipStream_t s = o->ctx->stream;
    if (kind == FEM_PRE_JACOBI) {
        DevBuf d;
        d.alloc(o->N * sizeof(double));
        Vec dv;
        dv.ctx = o->ctx; dv.n = o->N; dv.prec = FEM_F64; dv.data = d.p; dv.owner = false;
        o->diagonal(dv);
        dinv.ensure(o->N * (prec == FEM_F64 ? 8 : 4));
        if (prec == FEM_F64)
            hipLaunchKernelGGL(k_jacobi_inv<double>, dim3(grid_for(o->N, 256, 8192)), dim3(256), 0, s, d.as<double>(), dinv.as<double>(), o->N);
        else
            hipLaunchKernelGGL(k_jacobi_inv<float>, dim3(grid_for(o->N, 256, 8192)), dim3(256), 0, s, d.as<double>(), dinv.as<float>(), o->N);
        FEM_HIP(hipGetLastError());
        FEM_HIP(hipStreamSynchronize(s));
    } else if (kind == FEM_PRE_BLOCK_JACOBI) {
        FEM_CHECK(!o->nonlinear(), FEM_ERR_STATE, "block Jacobi: the operator is nonlinear (take its Jacobian)");
        FEM_CHECK(o->a64.kappa_mode >= 0, FEM_ERR_STATE, "coefficient data not set");
        // the eigen data are computed in FP64 and stored in the
        // preconditioner's precision
        if (prec == FEM_F64)
            dg_block_setup<double>(o->a64, o->k, Q.as<double>(), lam.as<double>(), s, bound);
        else
            dg_block_setup<float>(o->a64, o->k, Q.as<float>(), lam.as<float>(), s, bound);
        FEM_HIP(hipStreamSynchronize(s));
    } else if (kind == FEM_PRE_MULTIGRID) {
        mg->update();
    } else {
        FEM_CHECK(kind == FEM_PRE_NONE, FEM_ERR_UNSUPPORTED, "preconditioner update: Jacobi, block Jacobi, multigrid");
    }
}

bool Precond::updatable_for(const Operator *o, int k_) const
{
    if (o != op || k_ != kind || lowp != nullptr || o->vec_prec != prec || !same_desc(o->desc, desc)) return false;
    if (kind == FEM_PRE_MULTIGRID) return mg && mg->updatable_for(o);
    return kind == FEM_PRE_JACOBI || kind == FEM_PRE_BLOCK_JACOBI;
}

// y (FP64) = widen(M32 round(x)): the FP32 preconditioner inside an FP64
// Krylov solve
__global__ void k_round32(const double *__restrict__ x, float *__restrict__ y, int64_t n, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

__global__ void k_widen64(const float *__restrict__ x, double *__restrict__ y, int64_t n, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (double)x[i];
}

void Precond::init_widening(Operator *o64, const Precond *p32)
{
    FEM_CHECK(o64 && p32, FEM_ERR_ARG, "null argument");
    FEM_CHECK(o64->vec_prec == FEM_F64 && p32->prec == FEM_F32, FEM_ERR_ARG,
              "an FP32 preconditioner widened to an FP64 operator");
    FEM_CHECK(p32->op->N == o64->N, FEM_ERR_ARG, "preconditioner does not match the operator");
    op = o64;
    kind = FEM_PRE_WIDENED;
    prec = FEM_F64;
    lowp = p32;
    cx.alloc(o64->N * sizeof(float));
    cy.alloc(o64->N * sizeof(float));
}

Precond::~Precond()
{
    delete learned;
    delete mg;
}

void Precond::apply_raw(const void *x, void *y, const int *skip, double *dot) const
{
    hipStream_t s = op->ctx->stream;
    const int64_t N = op->N;
    FEM_CHECK(!dot || fused_dot(), FEM_ERR_UNSUPPORTED, "fused dot: block Jacobi only");
    if (lowp) {
        const unsigned G = grid_for(N, 256, 8192);
        hipLaunchKernelGGL(k_round32, dim3(G), dim3(256), 0, s, (const double *)x, cx.as<float>(), N, skip);
        lowp->apply_raw(cx.p, cy.p, skip);
        hipLaunchKernelGGL(k_widen64, dim3(G), dim3(256), 0, s, (const float *)cy.as<float>(), (double *)y, N, skip);
        FEM_HIP(hipGetLastError());
    } else if (kind == FEM_PRE_LEARNED) {
        learned->apply_raw(x, y, skip);
    } else if (kind == FEM_PRE_MULTIGRID) {
        mg->apply_raw(x, y, skip);
    } else if (kind == FEM_PRE_JACOBI) {
        if (prec == FEM_F64)
            hipLaunchKernelGGL(k_pointwise_mul<double>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, s, dinv.as<double>(), (const double *)x, (double *)y, N, skip);
        else
            hipLaunchKernelGGL(k_pointwise_mul<float>, dim3(grid_for(N, 256, 8192)), dim3(256), 0, s, dinv.as<float>(), (const float *)x, (float *)y, N, skip);
        FEM_HIP(hipGetLastError());
    } else if (kind == FEM_PRE_BLOCK_JACOBI) {
        if (prec == FEM_F64)
            dg_block_apply<double>(op->k, op->E, Q.as<double>(), lam.as<double>(), (const double *)x, (double *)y, skip, s,
                                   dot);
        else
            dg_block_apply<float>(op->k, op->E, Q.as<float>(), lam.as<float>(), (const float *)x, (float *)y, skip, s, dot);
    } else {
        FEM_HIP(hipMemcpyAsync(y, x, N * (prec == FEM_F64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace femamd
