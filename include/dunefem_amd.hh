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
 *                      (operator.hh:64); jacobian(u, jOp), matrix-free
 *                      (operator/common/differentiableoperator.hh:54)
 *   SolverParameter    fem.solver.* keys (dune/fem/solver/parameter.hh:71-250)
 *   SolverInfo         {converged, linearIterations, nonlinearIterations,
 *                       timing} (dune/fem/solver/inverseoperatorinterface.hh:15-30)
 *   InverseOperator    bind(op), bind(op, precond), unbind(), operator()(b, x)
 *                      with x the initial guess, iterations() < 0 when not
 *                      converged, info(), setMaxIterations, setParameters
 *                      (inverseoperatorinterface.hh:66-166); CgInverseOperator,
 *                      BicgstabInverseOperator, GmresInverseOperator
 *                      (krylovinverseoperators.hh:280-317) and
 *                      FgmresInverseOperator
 *   AffineGalerkinOperator, NewtonParameter, NewtonInverseOperator, FemScheme
 *                      R(u) = S[u] - F and the Newton solve of R(w) = u
 *                      (dune/fem/solver/newtoninverseoperator.hh:101-802,
 *                       dune/fem/schemes/femscheme.hh:195-253)
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
    // the Jacobian operator object of `op` (NewtonInverseOperator's jOp):
    // a linear operator with op's description and coefficient data
    struct JacobianOf {};
    GalerkinOperator(const GalerkinOperator &op, JacobianOf) : ctx_(op.ctx_), desc_(op.desc_)
    {
        check(fem_op_jacobian_create(op.h_, &h_));
    }
    ~GalerkinOperator() { if (h_) fem_op_destroy(h_); }
    GalerkinOperator(const GalerkinOperator &) = delete;
    GalerkinOperator &operator=(const GalerkinOperator &) = delete;

    // w = A u (w overwritten)
    void operator()(const DiscreteFunction &u, DiscreteFunction &w) const { check(fem_op_apply(h_, u.handle(), w.handle())); }
    bool nonlinear() const
    {
        const int r = fem_op_nonlinear(h_);
        if (r < 0) check(r);
        return r != 0;
    }
    // jOp = S'(u), jOp made by GalerkinOperator(op, JacobianOf{})
    void jacobian(const DiscreteFunction &u, JacobianOperatorType &jOp) const
    {
        check(fem_op_jacobian(h_, u.handle(), jOp.handle()));
    }
    // DG: reaction int (cvol + coef u^pw) u v beyond c: cvol at the volume
    // quadrature points [E][m^3] (empty clears), the law coef u^pw (coef 0 clears)
    void setReactionQp(const std::vector<double> &cvol)
    {
        check(fem_op_set_reaction_qp(h_, cvol.empty() ? nullptr : cvol.data()));
    }
    void setReactionLaw(double coef, int pw) { check(fem_op_set_reaction_law(h_, coef, pw)); }
    // a compiled pointwise reaction r(u, x) (a library written by
    // dunefem_amd.forms); empty path unloads it
    void setReactionPlugin(const std::string &path)
    {
        check(fem_op_set_reaction_plugin(h_, path.empty() ? nullptr : path.c_str()));
    }
    // u at the volume quadrature points [E][m^3] (x fastest)
    void evalQp(const DiscreteFunction &u, DiscreteFunction &uq) const { check(fem_op_eval_qp(h_, u.handle(), uq.handle())); }

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
    // the Krylov method a subclass fixes (-1: fem.solver.method decides)
    static constexpr int fixedMethod = -1;
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
    Context &ctx() const { return *ctx_; }

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
    static constexpr int fixedMethod = FEM_SOLVER_CG;
    explicit CgInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_CG) {}
};
class GmresInverseOperator : public InverseOperator {
public:
    static constexpr int fixedMethod = FEM_SOLVER_GMRES;
    explicit GmresInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_GMRES) {}
};
class FgmresInverseOperator : public InverseOperator {
public:
    static constexpr int fixedMethod = FEM_SOLVER_FGMRES;
    explicit FgmresInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_FGMRES) {}
};
class BicgstabInverseOperator : public InverseOperator {
public:
    static constexpr int fixedMethod = FEM_SOLVER_BICGSTAB;
    explicit BicgstabInverseOperator(Context &ctx, const SolverParameter &p = SolverParameter())
        : InverseOperator(ctx, p, FEM_SOLVER_BICGSTAB) {}
};
// the reference's other names for these (krylovinverseoperators.hh:302-317)
using ParDGGeneralizedMinResInverseOperator = GmresInverseOperator;
using ParDGBiCGStabInverseOperator = BicgstabInverseOperator;
using OEMCGOp = CgInverseOperator;
using OEMBICGSTABOp = BicgstabInverseOperator;
using OEMGMRESOp = GmresInverseOperator;

// R(u) = S[u] - F: a Galerkin operator with its source term, as the
// reference's GalerkinOperator evaluates it (galerkin.hh:333-360); nonlinear
// and differentiable as S is
class AffineGalerkinOperator {
public:
    using DomainFunctionType = DiscreteFunction;
    using RangeFunctionType = DiscreteFunction;
    using JacobianOperatorType = GalerkinOperator;

    AffineGalerkinOperator(const GalerkinOperator &S, const DiscreteFunction &F) : S_(&S), F_(&F) {}
    void operator()(const DiscreteFunction &u, DiscreteFunction &w) const
    {
        (*S_)(u, w);
        w -= *F_;
    }
    bool nonlinear() const { return S_->nonlinear(); }
    void jacobian(const DiscreteFunction &u, JacobianOperatorType &jOp) const { S_->jacobian(u, jOp); }
    const GalerkinOperator &op() const { return *S_; }
    const DiscreteFunction &source() const { return *F_; }

private:
    const GalerkinOperator *S_;
    const DiscreteFunction *F_;
};

// NewtonParameter (newtoninverseoperator.hh:101-380): fem.solver.nonlinear.*
// keys for Newton, the other fem.solver.* keys (and nonlinear.linear.*) for
// its linear solver
class NewtonParameter {
public:
    NewtonParameter() { fem_newton_default_params(&p_); }
    // the linear solver's parameters with Newton's defaults
    explicit NewtonParameter(const SolverParameter &linear) : linear_(linear) { fem_newton_default_params(&p_); }
    NewtonParameter &set(const std::string &key, const std::string &value)
    {
        if (newtonKey(key)) check(fem_newton_params_set(&p_, key.c_str(), value.c_str()));
        else linear_.set(key, value);
        return *this;
    }
    double tolerance() const { return p_.tolerance; }
    void setTolerance(double tol) { p_.tolerance = tol; }
    int maxIterations() const { return p_.maxiterations; }
    void setMaxIterations(int it) { p_.maxiterations = it; }
    int maxLinearIterations() const { return p_.maxlineariterations >= 0 ? p_.maxlineariterations : linear_.maxIterations(); }
    void setMaxLinearIterations(int it) { p_.maxlineariterations = it; }
    int lineSearch() const { return p_.linesearch; }
    int forcing() const { return p_.forcing; }
    bool simplified() const { return p_.simplified != 0; }
    bool verbose() const { return p_.verbose != 0; }
    SolverParameter &linear() { return linear_; }
    const SolverParameter &linear() const { return linear_; }
    fem_newton_params &raw() { return p_; }
    const fem_newton_params &raw() const { return p_; }

private:
    static bool newtonKey(std::string k)
    {
        const std::string fs = "fem.solver.";
        if (k.compare(0, fs.size(), fs) == 0) k = k.substr(fs.size());
        for (const char *pre : { "nonlinear.", "newton." }) {
            const std::string ps(pre);
            if (k.compare(0, ps.size(), ps) == 0) {
                const std::string rest = k.substr(ps.size());
                return rest.compare(0, 7, "linear.") != 0 || rest == "linear.tolerance.strategy";
            }
        }
        return false;
    }
    fem_newton_params p_;
    SolverParameter linear_;
};

// NewtonInverseOperator<JacobianOperator, LInvOp>::operator()(u, w): solves
// S[w] = u (newtoninverseoperator.hh:688-802) on the device, with the
// Jacobian S'(w) applied matrix-free.  The linear solver is the library's,
// with LinearInverseOperator's Krylov method (CgInverseOperator: cg, ...,
// InverseOperator: fem.solver.method).  A linear operator takes one step and
// counts no nonlinear iteration; failed() returns the NewtonFailure code
// (:389-400, :568-584).
template <class LinearInverseOperator = InverseOperator>
class NewtonInverseOperator {
public:
    using DomainFunctionType = DiscreteFunction;
    using RangeFunctionType = DiscreteFunction;

    explicit NewtonInverseOperator(Context &ctx, const NewtonParameter &param = NewtonParameter())
        : ctx_(&ctx), param_(param) { create(); }
    // the reference's (epsilon, parameter) constructor
    NewtonInverseOperator(Context &ctx, double epsilon, const NewtonParameter &param = NewtonParameter())
        : ctx_(&ctx), param_(param)
    {
        param_.setTolerance(epsilon);
        create();
    }
    // a linear solver's parameters with Newton's defaults
    NewtonInverseOperator(Context &ctx, const SolverParameter &linear) : ctx_(&ctx), param_(linear) { create(); }
    ~NewtonInverseOperator() { if (h_) fem_newton_destroy(h_); }
    NewtonInverseOperator(const NewtonInverseOperator &) = delete;
    NewtonInverseOperator &operator=(const NewtonInverseOperator &) = delete;

    void bind(const GalerkinOperator &op) { bind_(&op, nullptr, nullptr); }
    void bind(const GalerkinOperator &op, const Preconditioner &pre) { bind_(&op, &pre, nullptr); }
    // R(w) = S[w] - F = u is S[w] = u + F
    void bind(const AffineGalerkinOperator &op) { bind_(&op.op(), nullptr, &op.source()); }
    void bind(const AffineGalerkinOperator &op, const Preconditioner &pre) { bind_(&op.op(), &pre, &op.source()); }
    void unbind()
    {
        check(fem_newton_unbind(h_));
        op_ = nullptr;
        pre_ = nullptr;
        F_ = nullptr;
    }

    void operator()(const DiscreteFunction &u, DiscreteFunction &w) const
    {
        if (!op_) throw Exception(FEM_ERR_STATE, "NewtonInverseOperator: bind an operator first");
        if (F_) {
            DiscreteFunction rhs(u);
            rhs += *F_;
            check(fem_newton_solve(h_, rhs.handle(), w.handle(), &info_));
        } else {
            check(fem_newton_solve(h_, u.handle(), w.handle(), &info_));
        }
    }
    int iterations() const { return info_.iterations; }
    int linearIterations() const { return info_.linear_iterations; }
    double residual() const { return info_.residual; }
    int failed() const { return info_.failure; }
    bool converged() const { return info_.converged != 0; }
    std::vector<double> timing() const { return { info_.timing[0], info_.timing[1], info_.timing[2] }; }
    SolverInfo info() const
    {
        SolverInfo i;
        i.converged = converged();
        i.linearIterations = linearIterations();
        i.nonlinearIterations = iterations();
        i.timing = timing();
        return i;
    }
    void setMaxIterations(int it)
    {
        param_.setMaxIterations(it);
        recreate();
    }
    void setMaxLinearIterations(int it)
    {
        param_.setMaxLinearIterations(it);
        recreate();
    }
    NewtonParameter &parameter() { return param_; }

private:
    void create()
    {
        if (LinearInverseOperator::fixedMethod >= 0) param_.linear().setMethod(LinearInverseOperator::fixedMethod);
        check(fem_newton_create(ctx_->handle(), &param_.raw(), &param_.linear().raw(), &h_));
    }
    void recreate()
    {
        fem_newton *old = h_;
        h_ = nullptr;
        create();
        fem_newton_destroy(old);
        if (op_) bind_(op_, pre_, F_);
    }
    void bind_(const GalerkinOperator *op, const Preconditioner *pre, const DiscreteFunction *F)
    {
        check(fem_newton_bind(h_, op->handle(), pre ? pre->handle() : nullptr));
        op_ = op;
        pre_ = pre;
        F_ = F;
    }
    Context *ctx_;
    NewtonParameter param_;
    fem_newton *h_ = nullptr;
    const GalerkinOperator *op_ = nullptr;
    const Preconditioner *pre_ = nullptr;
    const DiscreteFunction *F_ = nullptr;
    mutable fem_newton_info info_{};
};

// FemScheme::solve / _solve (dune/fem/schemes/femscheme.hh:195-253) for
// R(u) = S[u] - F (boundary data inside F and, for Q2, in the Dirichlet
// rows): setConstraints and addConstraints on the solution, then
// NewtonInverseOperator over R with the linear inverse operator's
// parameters (one step when S is linear)
template <class LinearInverseOperator>
class FemScheme {
public:
    FemScheme(const GalerkinOperator &S, const DiscreteFunction &F, LinearInverseOperator &inv,
              const NewtonParameter &nonlinear = NewtonParameter())
        : S_(&S), R_(S, F), newton_(inv.ctx(), withLinear(nonlinear, inv.parameter())) {}
    SolverInfo solve(const DiscreteFunction &rhs, DiscreteFunction &solution) const
    {
        S_->constrain(solution, &rhs);          // sol = g + rhs on the boundary
        newton_.bind(R_);
        newton_(rhs, solution);
        newton_.unbind();
        return newton_.info();
    }
    SolverInfo solve(DiscreteFunction &solution) const
    {
        DiscreteFunction zero(solution);
        zero.clear();
        return solve(zero, solution);
    }

private:
    static NewtonParameter withLinear(NewtonParameter p, const SolverParameter &linear)
    {
        p.linear() = linear;
        return p;
    }
    const GalerkinOperator *S_;
    AffineGalerkinOperator R_;
    mutable NewtonInverseOperator<LinearInverseOperator> newton_;
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
