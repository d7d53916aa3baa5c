/* C++ host facade over the C ABI of dunefem_amd.h.
 *
 * Mirrors the dune-fem interfaces a caller of the matrix-free apply and the
 * Krylov solvers uses, with the same names, argument meanings and error
 * behaviour:
 *   DiscreteFunction   AdaptiveDiscreteFunction dof storage and BLAS-1
 *                      (dune/fem/function/common/discretefunction.hh:731-835,
 *                       function/blockvectors/defaultblockvectors.hh:110-135)
 *   GalerkinOperator   Operator<DF,DF>::operator()(u, w), w overwritten
 *                      (dune/fem/operator/common/operator.hh:55,
 *                       dune/fem/schemes/galerkin.hh:1460-1496); nonlinear()
 *                      (operator.hh:64); jacobian(u, jOp) of the linear
 *                      operator is the operator itself
 *                      (operator/common/differentiableoperator.hh:54)
 *   SolverParameter    fem.solver.* keys (dune/fem/solver/parameter.hh:71-250)
 *   SolverInfo         {converged, linearIterations, nonlinearIterations,
 *                       timing} (dune/fem/solver/inverseoperatorinterface.hh:15-30)
 *   InverseOperator    bind(op), bind(op, precond), unbind(), operator()(b, x)
 *                      with x the initial guess, iterations() < 0 when not
 *                      converged, info(), setMaxIterations, setParameters
 *                      (inverseoperatorinterface.hh:66-166); CgInverseOperator,
 *                      GmresInverseOperator (krylovinverseoperators.hh:280-295)
 *                      and FgmresInverseOperator
 *   AffineGalerkinOperator, NewtonInverseOperator
 *                      R(u) = A u - F and the Newton solve of R(w) = u
 *                      (dune/fem/solver/newtoninverseoperator.hh:688-802)
 * Errors: every non-zero status of the C ABI becomes a dunefem_amd::Exception
 * (the reference throws Dune::Exception subclasses); non-convergence is not an
 * exception but a negative iteration count, as in the reference.
 *
 * Header only; needs nothing but dunefem_amd.h and libdunefem_amd.so. */
#ifndef DUNEFEM_AMD_HH
#define DUNEFEM_AMD_HH

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dunefem_amd.h"

namespace dunefem_amd {

class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int status)
{
    if (status != FEM_OK) throw Exception(status, fem_last_error());
}

// one device, one stream: every call below is stream-ordered on it
class Context {
public:
    explicit Context(int device = 0) { check(fem_ctx_create(device, &h_)); }
    ~Context() { if (h_) fem_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    void synchronize() const { check(fem_ctx_synchronize(h_)); }
    fem_ctx *handle() const { return h_; }

private:
    fem_ctx *h_ = nullptr;
};

class DiscreteFunction {
public:
    using RangeFieldType = double;

    DiscreteFunction(const std::string &name, Context &ctx, int64_t size, int precision = FEM_F64)
        : name_(name), ctx_(&ctx)
    {
        check(fem_vec_create(ctx.handle(), size, precision, &h_));
        clear();
    }
    // over DoFs the caller already holds in device memory, without a copy
    // (AdaptiveDiscreteFunction(name, space, const DofType *data),
    // dune/fem/function/adaptivefunction/adaptivefunction.hh:81-87); the
    // memory is not released with this object
    DiscreteFunction(const std::string &name, Context &ctx, int64_t size, void *devicePtr, int precision = FEM_F64)
        : name_(name), ctx_(&ctx)
    {
        check(fem_vec_wrap(ctx.handle(), size, precision, devicePtr, &h_));
    }
    DiscreteFunction(const DiscreteFunction &other)
        : name_(other.name_), ctx_(other.ctx_)
    {
        check(fem_vec_create(ctx_->handle(), other.size(), other.precision(), &h_));
        assign(other);
    }
    DiscreteFunction(DiscreteFunction &&other) noexcept
        : name_(std::move(other.name_)), ctx_(other.ctx_), h_(other.h_)
    {
        other.h_ = nullptr;
    }
    DiscreteFunction &operator=(const DiscreteFunction &other)
    {
        if (this != &other) assign(other);
        return *this;
    }
    ~DiscreteFunction() { if (h_) fem_vec_destroy(h_); }

    const std::string &name() const { return name_; }
    int64_t size() const { return fem_vec_size(h_); }
    // the device DoF array (leakPointer(), adaptivefunction.hh:130)
    void *leakPointer() const { return fem_vec_device_ptr(h_); }
    int precision() const { return fem_vec_precision(h_); }

    void clear() { check(fem_vec_fill(h_, 0.0)); }
    void assign(const DiscreteFunction &g) { check(fem_vec_copy(h_, g.h_)); }
    void axpy(double s, const DiscreteFunction &g) { check(fem_vec_axpy(h_, s, g.h_)); }
    DiscreteFunction &operator+=(const DiscreteFunction &g) { axpy(1.0, g); return *this; }
    DiscreteFunction &operator-=(const DiscreteFunction &g) { axpy(-1.0, g); return *this; }
    DiscreteFunction &operator*=(double s) { check(fem_vec_scal(h_, s)); return *this; }
    DiscreteFunction &operator/=(double s) { return (*this) *= 1.0 / s; }
    double scalarProductDofs(const DiscreteFunction &g) const
    {
        double d = 0;
        check(fem_vec_dot(h_, g.h_, &d));
        return d;
    }
    double normSquaredDofs() const
    {
        double d = 0;
        check(fem_vec_norm2sq(h_, &d));
        return d;
    }
    // write / read in the binary stream format of DiscreteFunctionDefault
    // (discretefunction_inline.hh:213-305) to / from a file; read takes the
    // stored name, as the reference does
    void write(const std::string &path, int spaceId = FEM_SPACE_LEGENDRE_DG) const
    {
        check(fem_vec_write(h_, path.c_str(), name_.c_str(), spaceId));
    }
    void read(const std::string &path, int spaceId = FEM_SPACE_LEGENDRE_DG)
    {
        std::vector<char> nm(4096, 0);
        check(fem_vec_read(h_, path.c_str(), spaceId, nm.data(), (int64_t)nm.size()));
        name_ = nm.data();
    }
    // host <-> device copies of the dof vector
    void upload(const std::vector<double> &dofs)
    {
        if ((int64_t)dofs.size() != size()) throw Exception(FEM_ERR_ARG, "upload: size mismatch");
        check(fem_vec_upload(h_, dofs.data()));
    }
    std::vector<double> download() const
    {
        std::vector<double> v((size_t)size());
        check(fem_vec_download(h_, v.data()));
        return v;
    }
    fem_vec *handle() const { return h_; }
    Context &context() const { return *ctx_; }

private:
    std::string name_;
    Context *ctx_;
    fem_vec *h_ = nullptr;
};

class GalerkinOperator {
public:
    using DomainFunctionType = DiscreteFunction;
    using RangeFunctionType = DiscreteFunction;
    using JacobianOperatorType = GalerkinOperator;

    GalerkinOperator(Context &ctx, const fem_op_desc &desc) : ctx_(&ctx), desc_(desc)
    {
        check(fem_op_create(ctx.handle(), &desc, &h_));
    }
    ~GalerkinOperator() { if (h_) fem_op_destroy(h_); }
    GalerkinOperator(const GalerkinOperator &) = delete;
    GalerkinOperator &operator=(const GalerkinOperator &) = delete;

    // w = A u (w overwritten)
    void operator()(const DiscreteFunction &u, DiscreteFunction &w) const { check(fem_op_apply(h_, u.handle(), w.handle())); }
    bool nonlinear() const { return false; }
    // the linearisation of the affine operator is its linear part: this operator
    void jacobian(const DiscreteFunction &, const JacobianOperatorType *&jOp) const { jOp = this; }

    void setKappaQp(const std::vector<double> &kvol, const std::vector<double> &kx, const std::vector<double> &ky,
                    const std::vector<double> &kz = {})
    {
        check(fem_op_set_kappa_qp(h_, kvol.data(), kx.data(), ky.data(), kz.empty() ? nullptr : kz.data()));
    }
    void setKappaCell(const std::vector<double> &kcell) { check(fem_op_set_kappa_cell(h_, kcell.data())); }
    // F = int f phi from f at the volume quadrature points
    void assembleRHS(const std::vector<double> &fq, DiscreteFunction &F) const
    {
        check(fem_op_assemble_rhs(h_, fq.data(), F.handle()));
    }
    void diagonal(DiscreteFunction &d) const { check(fem_op_diagonal(h_, d.handle())); }
    // DG: the element blocks [E][nb][nb] (an FP64 function of E nb^2 entries)
    void blockDiagonal(DiscreteFunction &blocks) const { check(fem_op_block_diagonal(h_, blocks.handle())); }
    // u = g on the boundary.  DG: g at the face quadrature points per normal
    // direction (gx, gy, gz in the kface layout); Q2: nodal values in gx
    void setBoundaryData(const std::vector<double> &gx, const std::vector<double> &gy = {},
                         const std::vector<double> &gz = {})
    {
        check(fem_op_set_boundary_data(h_, gx.data(), gy.empty() ? nullptr : gy.data(),
                                       gz.empty() ? nullptr : gz.data()));
    }
    // Q2: x_D = g_D (+ add_D), the reference's setConstraints / addConstraints
    void constrain(DiscreteFunction &x, const DiscreteFunction *add = nullptr) const
    {
        check(fem_op_constrain(h_, x.handle(), add ? add->handle() : nullptr));
    }
    int64_t size() const { return fem_op_ndofs(h_); }
    int precision() const { return desc_.precision; }
    // a dof vector matching this operator (FP64 for FEM_F64, FP32 otherwise)
    DiscreteFunction function(const std::string &name) const
    {
        return DiscreteFunction(name, *ctx_, size(), desc_.precision == FEM_F64 ? FEM_F64 : FEM_F32);
    }
    const fem_op_desc &desc() const { return desc_; }
    fem_op *handle() const { return h_; }

private:
    Context *ctx_;
    fem_op_desc desc_;
    fem_op *h_ = nullptr;
};

class Preconditioner {
public:
    Preconditioner(GalerkinOperator &op, int kind) { check(fem_precond_create(op.handle(), kind, &h_)); }
    ~Preconditioner() { if (h_) fem_precond_destroy(h_); }
    Preconditioner(const Preconditioner &) = delete;
    Preconditioner &operator=(const Preconditioner &) = delete;
    // y = M^-1 x
    void operator()(const DiscreteFunction &x, DiscreteFunction &y) const
    {
        check(fem_precond_apply(h_, x.handle(), y.handle()));
    }
    fem_precond *handle() const { return h_; }

private:
    fem_precond *h_ = nullptr;
};

class SolverParameter {
public:
    SolverParameter() { fem_solver_default_params(&p_); }
    explicit SolverParameter(const fem_solver_params &p) : p_(p) {}
    // "fem.solver.method" = "cg" etc.
    SolverParameter &set(const std::string &key, const std::string &value)
    {
        check(fem_solver_params_set(&p_, key.c_str(), value.c_str()));
        return *this;
    }
    int solverMethod() const { return p_.method; }
    double tolerance() const { return p_.tolerance; }
    int maxIterations() const { return p_.maxiterations; }
    int errorMeasure() const { return p_.errormeasure; }
    int gmresRestart() const { return p_.restart; }
    bool verbose() const { return p_.verbose != 0; }
    void setMaxIterations(int it) { p_.maxiterations = it; }
    void setTolerance(double tol) { p_.tolerance = tol; }
    void setMethod(int m) { p_.method = m; }
    fem_solver_params &raw() { return p_; }
    const fem_solver_params &raw() const { return p_; }

private:
    fem_solver_params p_;
};

struct SolverInfo {
    bool converged = false;
    int linearIterations = 0;
    int nonlinearIterations = 0;     // FP64 refinement steps of a mixed solve
    std::vector<double> timing;
    double trueRelativeResidual = 0;  // ||b - A x|| / ||b|| in FP64
    bool fallback = false;
};

class InverseOperator {
public:
    using DomainFunctionType = DiscreteFunction;
    using RangeFunctionType = DiscreteFunction;
    using OperatorType = GalerkinOperator;
    using PreconditionerType = Preconditioner;

    InverseOperator(Context &ctx, const SolverParameter &param) : ctx_(&ctx), param_(param) { create(); }
    virtual ~InverseOperator() { if (h_) fem_solver_destroy(h_); }
    InverseOperator(const InverseOperator &) = delete;
    InverseOperator &operator=(const InverseOperator &) = delete;

    void bind(const OperatorType &op) { bind_(&op, nullptr, nullptr, nullptr); }
    void bind(const OperatorType &op, const PreconditionerType &pre) { bind_(&op, &pre, nullptr, nullptr); }
    // FP64 iterative refinement around inner solves with a low-precision operator
    void bindRefinement(const OperatorType &op, const OperatorType &innerOp, const PreconditionerType *innerPre)
    {
        bind_(&op, nullptr, &innerOp, innerPre);
    }
    void unbind()
    {
        check(fem_solver_unbind(h_));
        op_ = inner_ = nullptr;
        pre_ = innerPre_ = nullptr;
    }
    // x = op^-1 b, x is the initial guess on entry
    virtual void operator()(const DomainFunctionType &b, RangeFunctionType &x) const
    {
        fem_solver_info i{};
        check(fem_solver_solve(h_, b.handle(), x.handle(), &i));
        info_.converged = i.converged != 0;
        info_.linearIterations = i.iterations;
        info_.nonlinearIterations = i.outer_iterations;
        info_.timing = { i.time_s };
        info_.trueRelativeResidual = i.true_rel_residual;
        info_.fallback = i.fallback != 0;
    }
    int iterations() const { return info_.linearIterations; }
    SolverInfo info() const { return info_; }
    void setMaxIterations(int it)
    {
        param_.setMaxIterations(it);
        recreate();
    }
    void setMaxLinearIterations(int it) { setMaxIterations(it); }
    void setParameters(const SolverParameter &p)
    {
        const int method = param_.solverMethod();
        param_ = p;
        if (fixed_method_) param_.setMethod(method);
        recreate();
    }
    SolverParameter &parameter() { return param_; }
    bool verbose() const { return param_.verbose(); }

protected:
    InverseOperator(Context &ctx, SolverParameter param, int method) : ctx_(&ctx), param_(std::move(param)), fixed_method_(true)
    {
        param_.setMethod(method);
        create();
    }

private:
    void create() { check(fem_solver_create(ctx_->handle(), &param_.raw(), &h_)); }
    void recreate()
    {
        fem_solver *old = h_;
        h_ = nullptr;
        create();
        fem_solver_destroy(old);
        if (op_) bind_(op_, pre_, inner_, innerPre_);
    }
    void bind_(const OperatorType *op, const PreconditionerType *pre, const OperatorType *inner,
               const PreconditionerType *innerPre)
    {
        check(fem_solver_bind(h_, op->handle(), pre ? pre->handle() : nullptr, inner ? inner->handle() : nullptr,
                              innerPre ? innerPre->handle() : nullptr));
        op_ = op;
        pre_ = pre;
        inner_ = inner;
        innerPre_ = innerPre;
    }

    Context *ctx_;
    SolverParameter param_;
    bool fixed_method_ = false;
    fem_solver *h_ = nullptr;
    const OperatorType *op_ = nullptr, *inner_ = nullptr;
    const PreconditionerType *pre_ = nullptr, *innerPre_ = nullptr;
    mutable SolverInfo info_;
};

// KrylovInverseOperator<DF, cg | gmres> and a flexible GMRES
class CgInverseOperator : public InverseOperator {
public:
    explicit CgInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_CG) {}
};
class GmresInverseOperator : public InverseOperator {
public:
    explicit GmresInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_GMRES) {}
};
class FgmresInverseOperator : public InverseOperator {
public:
    explicit FgmresInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_FGMRES) {}
};

// R(u) = A u - F: a Galerkin operator with its source term, as the
// reference's GalerkinOperator evaluates it (galerkin.hh:333-360); its
// jacobian is the linear operator A (nonlinear() == false)
class AffineGalerkinOperator {
public:
    using DomainFunctionType = DiscreteFunction;
    using RangeFunctionType = DiscreteFunction;
    using JacobianOperatorType = GalerkinOperator;

    AffineGalerkinOperator(const GalerkinOperator &A, const DiscreteFunction &F) : A_(&A), F_(&F) {}
    void operator()(const DiscreteFunction &u, DiscreteFunction &w) const
    {
        (*A_)(u, w);
        w -= *F_;
    }
    bool nonlinear() const { return false; }
    const JacobianOperatorType &jacobian(const DiscreteFunction &) const { return *A_; }

private:
    const GalerkinOperator *A_;
    const DiscreteFunction *F_;
};

// NewtonInverseOperator<JacobianOperator, LInvOp>::operator()(u, w): solves
// R(w) = u (dune/fem/solver/newtoninverseoperator.hh:688-802): residual
// R(w) - u, dw = DR^-1 residual, w -= dw; for an operator with
// nonlinear() == false one step, else until |residual| <= tolerance
template <class LinearInverseOperator>
class NewtonInverseOperator {
public:
    NewtonInverseOperator(Context &ctx, const SolverParameter &linear = SolverParameter(), double tolerance = 1e-6,
                          int maxIterations = 100, int maxLinearIterations = 1 << 30)
        : jInv_(ctx, linear), tolerance_(tolerance), maxIterations_(maxIterations),
          maxLinearIterations_(maxLinearIterations) {}

    void bind(const AffineGalerkinOperator &op) { op_ = &op; }
    void unbind() { op_ = nullptr; }

    void operator()(const DiscreteFunction &u, DiscreteFunction &w) const
    {
        if (!op_) throw Exception(FEM_ERR_STATE, "NewtonInverseOperator: bind an operator first");
        DiscreteFunction residual(u), dw(w);
        iterations_ = 0;
        linearIterations_ = 0;
        (*op_)(w, residual);
        residual -= u;
        delta_ = std::sqrt(residual.normSquaredDofs());
        const bool nonlinear = op_->nonlinear();
        nonlinear_ = nonlinear;
        while (true) {
            jInv_.bind(op_->jacobian(w));
            const int left = maxLinearIterations_ - linearIterations_;
            if (left <= 0) break;
            jInv_.setMaxIterations(left);
            dw.clear();
            jInv_(residual, dw);
            if (jInv_.iterations() < 0) {   // the linear solve did not converge
                linearIterations_ = jInv_.iterations();
                break;
            }
            linearIterations_ += jInv_.iterations();
            w -= dw;
            (*op_)(w, residual);
            residual -= u;
            delta_ = std::sqrt(residual.normSquaredDofs());
            ++iterations_;
            if (!nonlinear || delta_ <= tolerance_ || iterations_ >= maxIterations_) break;
        }
    }
    int iterations() const { return iterations_; }
    int linearIterations() const { return linearIterations_; }
    double residualNorm() const { return delta_; }
    // the linear solves converged and, for a nonlinear operator, |residual| <= tolerance
    bool converged() const { return linearIterations_ >= 0 && iterations_ > 0 && (!nonlinear_ || delta_ <= tolerance_); }

private:
    const AffineGalerkinOperator *op_ = nullptr;
    mutable LinearInverseOperator jInv_;
    double tolerance_;
    int maxIterations_, maxLinearIterations_;
    mutable int iterations_ = 0, linearIterations_ = 0;
    mutable double delta_ = 0;
    mutable bool nonlinear_ = false;
};

// FemScheme::solve / _solve (dune/fem/schemes/femscheme.hh:195-253) for the
// affine operator R(u) = A u - F (boundary data inside F and, for Q2, in the
// Dirichlet rows): setConstraints and addConstraints on the solution, then
// the Newton step of NewtonInverseOperator over R with the linear inverse
// operator (one step: R is affine)
template <class LinearInverseOperator>
class FemScheme {
public:
    FemScheme(const GalerkinOperator &A, const DiscreteFunction &F, LinearInverseOperator &inv)
        : A_(&A), R_(A, F), inv_(&inv) {}
    SolverInfo solve(const DiscreteFunction &rhs, DiscreteFunction &solution) const
    {
        A_->constrain(solution, &rhs);          // sol = g + rhs on the boundary
        return newton(&rhs, solution);
    }
    SolverInfo solve(DiscreteFunction &solution) const
    {
        A_->constrain(solution);                // sol = g on the boundary
        return newton(nullptr, solution);
    }

private:
    SolverInfo newton(const DiscreteFunction *rhs, DiscreteFunction &w) const
    {
        DiscreteFunction residual(w), dw(w);
        R_(w, residual);
        if (rhs) residual -= *rhs;
        inv_->bind(*A_);
        dw.clear();
        (*inv_)(residual, dw);
        SolverInfo info = inv_->info();
        w -= dw;
        inv_->unbind();
        info.nonlinearIterations = info.converged ? 1 : 0;
        return info;
    }
    const GalerkinOperator *A_;
    AffineGalerkinOperator R_;
    LinearInverseOperator *inv_;
};

// Gauss-Legendre rule with m points on [0,1] (the reference's GaussPts
// rule, dune/fem/quadrature/gausspoints_implementation.hh)
inline void gaussPoints(int m, std::vector<double> &x, std::vector<double> &w)
{
    x.resize(m);
    w.resize(m);
    check(fem_gauss_points(m, x.data(), w.data()));
}

}  // namespace dunefem_amd

#endif
