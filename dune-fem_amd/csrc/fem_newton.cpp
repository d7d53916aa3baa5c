// Newton's method over the matrix-free operators.
//
// The loop is that of NewtonInverseOperator::operator()
// (dune/fem/solver/newtoninverseoperator.hh:688-802): residual S[w] - u, the
// Jacobian S'(w) bound to the linear inverse operator, dw from
// S'(w) dw = S[w] - u starting at dw = 0, w -= dw, stop when |S[w] - u|_2 <
// tolerance or a failure criterion holds (:568-584).  A linear operator takes
// one step and counts no nonlinear iteration (:786-790).  The Jacobian of a
// nonlinear operator is itself matrix-free: the linear operator with the
// pointwise reaction coefficient of the linearised law at w, evaluated on
// the device (Operator::jacobian).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>

#include "fem_blas.hpp"
#include "fem_newton.hpp"

namespace femamd {

double EisenstatWalker::next(double res)
{
    // eta_k = gamma (|F_k| / |F_k-1|)^2, not below gamma eta_k-1^2 when that
    // is at least 0.1, capped by eta_max, and not below half the Newton
    // tolerance relative to |F_k| (no oversolving near the end)
    // (Eisenstat and Walker, SIAM J. Sci. Comput. 17 (1996), choice 2;
    // newtoninverseoperator.hh:64-81)
    double eta = eta_max;
    if (prev_eta >= 0) {
        const double ea = gamma * res * res / (prev_res * prev_res);
        const double safe = gamma * prev_eta * prev_eta;
        const double ec = safe < 0.1 ? std::min(ea, eta_max) : std::min(eta_max, std::max(ea, safe));
        eta = std::min(eta_max, std::max(ec, 0.5 * tol / res));
    }
    prev_res = res;
    prev_eta = eta;
    return eta;
}

void Newton::init(Context *c, const fem_newton_params &p, const fem_solver_params &linear)
{
    ctx = c;
    FEM_CHECK(p.tolerance > 0, FEM_ERR_ARG, "newton: tolerance > 0");
    FEM_CHECK(p.maxiterations >= 0 && p.maxlinesearchiterations >= 0, FEM_ERR_ARG, "newton: iteration limits >= 0");
    FEM_CHECK(p.linesearch == FEM_LINESEARCH_NONE || p.linesearch == FEM_LINESEARCH_SIMPLE, FEM_ERR_ARG,
              "newton: linesearch none or simple");
    FEM_CHECK(p.forcing == FEM_FORCING_NONE || p.forcing == FEM_FORCING_EISENSTATWALKER, FEM_ERR_ARG,
              "newton: forcing none or eisenstatwalker");
    // the forcing term is a relative reduction of the linear residual
    // (newtoninverseoperator.hh:173-181)
    FEM_CHECK(p.forcing != FEM_FORCING_EISENSTATWALKER || linear.errormeasure == FEM_ERRMEASURE_REDUCTION,
              FEM_ERR_ARG, "newton: Eisenstat-Walker forcing needs the linear errormeasure residualreduction");
    FEM_CHECK(p.forcing != FEM_FORCING_EISENSTATWALKER ||
                  (p.eisenstatwalker_etamax > 0 && p.eisenstatwalker_etamax < 1 && p.eisenstatwalker_gamma > 0),
              FEM_ERR_ARG, "newton: Eisenstat-Walker etamax in (0, 1), gamma > 0");
    prm = p;
    lin.init(c, linear);
}

void Newton::bind(Operator *o, Precond *p)
{
    FEM_CHECK(o, FEM_ERR_ARG, "newton bind: null operator");
    FEM_CHECK(!p || p->op->N == o->N, FEM_ERR_ARG, "newton bind: preconditioner does not match the operator");
    op = o;
    pre = p;
    const size_t bytes = (size_t)o->N * (o->vec_prec == FEM_F64 ? 8 : 4);
    if (res.bytes != bytes) res.alloc(bytes);
    if (dw.bytes != bytes) dw.alloc(bytes);
}

void Newton::unbind()
{
    lin.unbind();
    op = nullptr;
    pre = nullptr;
}

int Newton::max_linear() const
{
    return linear_limit;
}

// NewtonInverseOperator::failed (newtoninverseoperator.hh:568-584)
int Newton::failed() const
{
    if (!(delta < std::numeric_limits<double>::max()) || std::isnan(delta)) return FEM_NEWTON_INVALID_RESIDUAL;
    if (iterations >= prm.maxiterations) return FEM_NEWTON_TOO_MANY_ITERATIONS;
    if (linear_iterations >= max_linear()) return FEM_NEWTON_TOO_MANY_LINEAR_ITERATIONS;
    if (linear_iterations < 0) return FEM_NEWTON_LINEAR_SOLVER_FAILED;
    if (!step_completed) return FEM_NEWTON_LINESEARCH_FAILED;
    return FEM_NEWTON_SUCCESS;
}

void Newton::residual_vec(const Vec *u, const Vec &w, Vec &r) const
{
    op->apply(w, r);
    if (u) vec_axpy(r, -1.0, *u);
}

double Newton::residual(const Vec *u, const Vec &w, Vec &r) const
{
    residual_vec(u, w, r);
    return std::sqrt(vec_dot(r, r));
}

void Newton::update_linear_tolerance()
{
    if (prm.forcing == FEM_FORCING_EISENSTATWALKER) lin.prm.tolerance = ew.next(delta);
}

// the step-halving line search of NewtonInverseOperator::lineSearch
// (newtoninverseoperator.hh:588-627): w = w_old - dw on entry, halve the
// step (w += factor dw) while the residual norm has not decreased.  Returns
// -1 on failure, 1 when the full step decreased the residual, else 0.
int Newton::line_search(const Vec *u, Vec &w, const Vec &d, Vec &r)
{
    const double old = delta;
    delta = std::sqrt(vec_dot(r, r));
    if (prm.linesearch == FEM_LINESEARCH_NONE) return 0;
    if (failed() == FEM_NEWTON_INVALID_RESIDUAL) {
        const double t = vec_dot(d, d);
        if (!(t < std::numeric_limits<double>::max() && !std::isnan(t))) delta = 2 * old;
    }
    double factor = 1.0;
    const int full = delta < old ? 1 : 0;
    int it = 0;
    while (delta >= old) {
        const double prev = delta;
        factor *= 0.5;
        if (std::abs(delta - old) < 1e-5 * delta) return -1;
        vec_axpy(w, factor, d);
        delta = residual(u, w, r);
        if (std::abs(delta - prev) < 1e-15) return -1;
        if (failed() == FEM_NEWTON_INVALID_RESIDUAL) delta = 2 * old;
        if (++it >= prm.maxlinesearchiterations) return -1;
    }
    return full;
}

void Newton::solve(const Vec *u, Vec &w, fem_newton_info &info)
{
    FEM_CHECK(op, FEM_ERR_STATE, "newton: solve before bind");
    FEM_CHECK(w.n == op->N && w.prec == op->vec_prec, FEM_ERR_ARG, "newton: solution vector size or precision");
    FEM_CHECK(!u || (u->n == op->N && u->prec == op->vec_prec), FEM_ERR_ARG, "newton: right-hand side size or precision");
    FEM_CHECK(!u || u->data != w.data, FEM_ERR_ARG, "newton: u and w must be distinct");
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
    FEM_HIP(hipStreamSynchronize(ctx->stream));
    const auto t_all = clk::now();
    double t_jac = 0;
    Vec r, d;
    r.ctx = d.ctx = ctx;
    r.n = d.n = op->N;
    r.prec = d.prec = op->vec_prec;
    r.data = res.p;
    d.data = dw.p;
    r.owner = d.owner = false;

    const bool nonlinear = op->nonlinear();
    if (nonlinear) {
        // S'(w) starts as a copy of S's linear part and coefficient data,
        // taken here so that data set on S after bind is seen
        jop.reset(new Operator());
        op->make_jacobian(*jop);
    }
    // a linear solver with FP64 refinement solves each Newton system by FP32
    // inner solves on an FP32 twin of S'(w) (the true residual and the
    // update stay FP64); a preconditioner given at bind serves the inner
    // solves
    const bool ir = lin.prm.refinement != 0;
    if (ir) {
        FEM_CHECK(op->is_dg() && op->desc.precision == FEM_F64, FEM_ERR_UNSUPPORTED,
                  "newton with refinement: FP64 DG operators");
        jop32.reset(new Operator());
        op->make_jacobian(*jop32, FEM_F32);
    } else {
        jop32.reset();
    }
    // the linear solver's tolerance and iteration limit change per step:
    // restored, and the solver unbound, however the solve ends
    struct Restore {
        Solver &s;
        double tol;
        int maxit;
        ~Restore()
        {
            s.prm.tolerance = tol;
            s.prm.maxiterations = maxit;
            try { s.unbind(); } catch (...) {}
        }
    } restore{ lin, lin.prm.tolerance, lin.prm.maxiterations };
    // the total over all steps (newtoninverseoperator.hh:246-250), fixed
    // before the per-step limits below change the linear solver's own
    linear_limit = prm.maxlineariterations >= 0 ? prm.maxlineariterations : lin.prm.maxiterations;
    ew = EisenstatWalker{ prm.eisenstatwalker_etamax, prm.eisenstatwalker_gamma, prm.tolerance };
    step_completed = true;
    iterations = 0;
    linear_iterations = 0;
    delta = residual(u, w, r);   // S[w] - u
    update_linear_tolerance();
    const bool verbose = prm.verbose && nonlinear;
    if (verbose) std::printf("Newton iteration 0: |residual| = %.6e", delta);

    bool compute_jacobian = true;
    Operator *J = nonlinear ? jop.get() : op;
    while (true) {
        if (verbose) std::printf("\n");
        const auto tj = clk::now();
        if (compute_jacobian && nonlinear) {
            op->jacobian(w, *J);
            if (ir) op->jacobian(w, *jop32);
            compute_jacobian = !prm.simplified;
        }
        // bind S'(w) (and the preconditioner it builds from
        // fem.solver.preconditioning.method, or the one given)
        if (ir) lin.bind(J, nullptr, jop32.get(), pre);
        else lin.bind(J, pre, nullptr, nullptr);
        FEM_HIP(hipStreamSynchronize(ctx->stream));
        t_jac += secs(tj);
        const int left = max_linear() - linear_iterations;
        if (left <= 0) break;
        lin.prm.maxiterations = left;
        vec_fill(d, 0.0);
        fem_solver_info li;
        lin.solve(r, d, li);   // dw = S'(w)^-1 (S[w] - u)
        if (li.iterations < 0) {
            linear_iterations = li.iterations;
            break;
        }
        linear_iterations += li.iterations;
        vec_axpy(w, -1.0, d);   // w -= dw
        if (!nonlinear) break;   // a linear operator: one step
        residual_vec(u, w, r);   // S[w] - u; line_search takes its norm
        const int ls = line_search(u, w, d, r);
        step_completed = ls >= 0;
        update_linear_tolerance();
        ++iterations;
        if (verbose) std::printf("Newton iteration %d: |residual| = %.6e", iterations, delta);
        if (delta < prm.tolerance || failed() != FEM_NEWTON_SUCCESS) break;
    }
    if (verbose) std::printf("\nLinear iterations: %d\n", linear_iterations);
    FEM_HIP(hipStreamSynchronize(ctx->stream));
    timing[0] = secs(t_all);
    timing[1] = t_jac;
    timing[2] = timing[0] - t_jac;
    info.failure = failed();
    info.converged = info.failure == FEM_NEWTON_SUCCESS;
    info.iterations = iterations;
    info.linear_iterations = linear_iterations;
    info.residual = delta;
    for (int i = 0; i < 3; ++i) info.timing[i] = timing[i];
}

}  // namespace femamd
