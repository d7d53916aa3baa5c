// C++ driver through the host facade (include/dunefem_amd.hh), after the
// reference's dune/fem/solver/test/inverseoperatortest.cc:
//   Algorithm::apply (:100-205): L2 projection of f = prod sin(pi x_k) into
//   the 2D Lagrange Q2 space on the 16x16 unit-square grid refined once
//   (:243-248), solved with the inverse operators; pass when the L2 distance
//   to f is below eps = 5e-6 (:117).
// The mass operator here is the Q2 operator with kappa = 0, c = 1 and strong
// u = 0 (f vanishes on the boundary).  Further checks use the SIPG DG
// operator: w overwritten by operator(), linearity, preconditioned and mixed
// solves, negative iterations on non-convergence, error codes as exceptions.
// Prints one line per check and "PASS" / "FAIL" at the end; exit code 0 on pass.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "dunefem_amd.hh"

using namespace dunefem_amd;

static int failures = 0;

static void report(bool ok, const std::string &what, double value)
{
    std::printf("%s %-60s %.3e\n", ok ? "ok  " : "FAIL", what.c_str(), value);
    if (!ok) ++failures;
}

static double fexact(double x, double y) { return std::sin(M_PI * x) * std::sin(M_PI * y); }

// f at the m x m Gauss points of every cell, [E][m^2], qp index x fastest
static std::vector<double> volume_values(int n, int m)
{
    std::vector<double> gx, gw;
    gaussPoints(m, gx, gw);
    std::vector<double> fq((size_t)n * n * m * m);
    const double h = 1.0 / n;
    for (int ey = 0; ey < n; ++ey)
        for (int ex = 0; ex < n; ++ex)
            for (int qy = 0; qy < m; ++qy)
                for (int qx = 0; qx < m; ++qx)
                    fq[((size_t)(ey * n + ex) * m + qy) * m + qx] = fexact((ex + gx[qx]) * h, (ey + gx[qy]) * h);
    return fq;
}

// L2 distance between the Q2 function (nodal lattice (2n+1)^2, x fastest)
// and f, with a 5x5 Gauss rule per cell
static double l2distance(int n, const std::vector<double> &u)
{
    std::vector<double> gx, gw;
    gaussPoints(5, gx, gw);
    const int NX = 2 * n + 1;
    const double h = 1.0 / n;
    auto L = [](int a, double t) {
        return a == 0 ? 2 * (t - 0.5) * (t - 1) : a == 1 ? -4 * t * (t - 1) : 2 * t * (t - 0.5);
    };
    double s = 0;
    for (int ey = 0; ey < n; ++ey)
        for (int ex = 0; ex < n; ++ex)
            for (int qy = 0; qy < 5; ++qy)
                for (int qx = 0; qx < 5; ++qx) {
                    double v = 0;
                    for (int b = 0; b < 3; ++b)
                        for (int a = 0; a < 3; ++a)
                            v += u[(size_t)(2 * ey + b) * NX + 2 * ex + a] * L(a, gx[qx]) * L(b, gx[qy]);
                    const double d = v - fexact((ex + gx[qx]) * h, (ey + gx[qy]) * h);
                    s += gw[qx] * gw[qy] * h * h * d * d;
                }
    return std::sqrt(s);
}

template <class InverseOperatorType>
static void algorithm(Context &ctx, const std::string &designation, const SolverParameter *param)
{
    const double eps = 5e-6;
    const int n = 32;   // 16 x 16, refined once
    fem_op_desc d{};
    d.kind = FEM_OP_CG_Q2_2D;
    d.n[0] = d.n[1] = n;
    d.n[2] = 1;
    d.order = 2;
    d.c = 1.0;
    d.kappa_mode = FEM_KAPPA_CONST;
    d.kappa_const = 0.0;
    d.precision = FEM_F64;
    GalerkinOperator massOperator(ctx, d);
    DiscreteFunction u("u", ctx, massOperator.size());
    DiscreteFunction rhs("rhs", ctx, massOperator.size());
    u.clear();
    massOperator.assembleRHS(volume_values(n, 3), rhs);
    if (param) {
        InverseOperatorType inverseOperator(ctx, *param);
        inverseOperator.bind(massOperator);
        inverseOperator(rhs, u);
        const double dist = l2distance(n, u.download());
        report(dist < eps && inverseOperator.iterations() > 0, designation + " InvOp(myParam)", dist);
    } else {
        InverseOperatorType inverseOperator(ctx);
        inverseOperator.bind(massOperator);
        inverseOperator(rhs, u);
        double dist = l2distance(n, u.download());
        report(dist < eps && inverseOperator.iterations() > 0, designation + " inverseOp()", dist);
        // new parameters through the fem.solver keys
        SolverParameter p;
        p.set("fem.solver.method", "cg").set("fem.solver.tolerance", "1e-10");
        InverseOperatorType inverseOperatorA(ctx, p);
        inverseOperatorA.bind(massOperator);
        u.clear();
        inverseOperatorA(rhs, u);
        dist = l2distance(n, u.download());
        report(dist < eps && inverseOperatorA.info().converged, designation + " inverseOp(NewSParam)", dist);
    }
    // NewtonInverseOperator on the affine operator R(u) = M u - F with
    // right-hand side 0 (inverseoperatortest.cc: newtonInvOp(rhs, u) after
    // rhs.clear()); one Newton step for the linear operator, which counts no
    // nonlinear iteration (newtoninverseoperator.hh:786-790)
    AffineGalerkinOperator affineMassOperator(massOperator, rhs);
    DiscreteFunction zero("zero", ctx, massOperator.size());
    NewtonInverseOperator<InverseOperatorType> newtonInvOp(ctx, param ? *param : SolverParameter());
    newtonInvOp.bind(affineMassOperator);
    u.clear();
    newtonInvOp(zero, u);
    const double dist = l2distance(n, u.download());
    report(dist < eps && newtonInvOp.converged() && newtonInvOp.iterations() == 0 && newtonInvOp.linearIterations() > 0,
           designation + " NewtonInvOp", dist);
}

// a parameter value in full precision (std::to_string prints six decimals)
static std::string num(double v)
{
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

static fem_op_desc dg_desc(int n, int p, int precision, bool advection = true)
{
    fem_op_desc d{};
    d.kind = FEM_OP_DG_SIPG_3D;
    d.n[0] = d.n[1] = d.n[2] = n;
    d.order = p;
    d.beta = 10.0 * p * p;
    d.b[0] = advection ? 1.0 : 0.0;
    d.b[1] = advection ? 0.5 : 0.0;
    d.b[2] = advection ? 0.25 : 0.0;
    d.kappa_mode = FEM_KAPPA_CELL;
    d.precision = precision;
    return d;
}

static void dg_checks(Context &ctx)
{
    const int n = 8, p = 2;
    const int E = n * n * n;
    std::vector<double> kcell(E);
    for (int e = 0; e < E; ++e) kcell[e] = std::exp(std::sin(0.37 * e));
    GalerkinOperator op(ctx, dg_desc(n, p, FEM_F64));
    op.setKappaCell(kcell);
    const int64_t N = op.size();
    std::vector<double> uh(N), vh(N);
    for (int64_t i = 0; i < N; ++i) {
        uh[i] = std::sin(0.001 * i * i + 0.3);
        vh[i] = std::cos(0.7 * i);
    }
    DiscreteFunction u("u", ctx, N), v("v", ctx, N), w("w", ctx, N), w2("w2", ctx, N), t("t", ctx, N);
    u.upload(uh);
    v.upload(vh);
    // operator()(u, w) overwrites w
    op(u, w);
    std::vector<double> garbage(N, 7.0);
    w2.upload(garbage);
    op(u, w2);
    w2 -= w;
    report(w2.normSquaredDofs() == 0.0, "DG apply overwrites w", std::sqrt(w2.normSquaredDofs()));
    // a discrete function over existing device DoF storage is a view
    {
        DiscreteFunction uv("u view", ctx, N, u.leakPointer());
        op(uv, w2);
        w2 -= w;
        report(w2.normSquaredDofs() == 0.0, "DG apply on a view of u", std::sqrt(w2.normSquaredDofs()));
    }
    report(std::fabs(u.normSquaredDofs() - u.scalarProductDofs(u)) == 0.0 && u.normSquaredDofs() > 0,
           "u intact after its view is gone", u.normSquaredDofs());
    // linearity: A(2u + v) = 2 A u + A v
    t.assign(u);
    t *= 2.0;
    t += v;
    op(t, w2);
    op(v, t);
    w2 -= t;
    w2.axpy(-2.0, w);
    const double lin = std::sqrt(w2.normSquaredDofs() / w.normSquaredDofs());
    report(lin < 1e-13, "DG apply linear", lin);
    // Jacobi FGMRES on the advection-diffusion operator to 1e-8 relative
    Preconditioner bj(op, FEM_PRE_JACOBI);
    SolverParameter prm;
    prm.set("fem.solver.errormeasure", "relative").set("fem.solver.tolerance", "1e-8").set("fem.solver.gmres.restart", "30");
    FgmresInverseOperator fgmres(ctx, prm);
    fgmres.bind(op, bj);
    DiscreteFunction x("x", ctx, N);
    fgmres(v, x);
    SolverInfo info = fgmres.info();
    report(info.converged && info.trueRelativeResidual <= 1e-8 && fgmres.iterations() > 0,
           "FGMRES + Jacobi, advection-diffusion", info.trueRelativeResidual);
    // CG on the symmetric (diffusion only) operator after setParameters:
    // the method stays CG
    GalerkinOperator sym(ctx, dg_desc(n, p, FEM_F64, false));
    sym.setKappaCell(kcell);
    CgInverseOperator cg(ctx);
    SolverParameter prmcg = prm;
    prmcg.set("fem.solver.residualcheck", "true");   // stop on ||r||, not on <r, M^-1 r>
    cg.setParameters(prmcg);
    Preconditioner bjs(sym, FEM_PRE_BLOCK_JACOBI);
    cg.bind(sym, bjs);
    x.clear();
    cg(v, x);
    info = cg.info();
    report(info.converged && info.trueRelativeResidual <= 1.01e-8 && cg.parameter().solverMethod() == FEM_SOLVER_CG,
           "CG + block Jacobi, diffusion only", info.trueRelativeResidual);
    // non-convergence: negative iterations, no exception
    fgmres.setMaxIterations(2);
    x.clear();
    fgmres(v, x);
    report(fgmres.iterations() < 0 && !fgmres.info().converged, "maxiterations 2 -> iterations() < 0",
           (double)fgmres.iterations());
    // FP64 refinement around FP32 FGMRES + block Jacobi (diffusion only)
    GalerkinOperator op32(ctx, dg_desc(n, p, FEM_F32, false));
    op32.setKappaCell(kcell);
    Preconditioner bj32(op32, FEM_PRE_BLOCK_JACOBI);
    SolverParameter pr = prm;
    pr.raw().refinement = 1;
    pr.raw().inner_tolerance = 1e-3;
    pr.raw().inner_maxiterations = 1000;
    FgmresInverseOperator mixed(ctx, pr);
    mixed.bindRefinement(sym, op32, &bj32);
    x.clear();
    mixed(v, x);
    info = mixed.info();
    report(info.converged && info.trueRelativeResidual <= 1e-8 && info.nonlinearIterations > 0,
           "FP64 refinement around FP32 FGMRES", info.trueRelativeResidual);
    // errors become exceptions: a dof vector of the wrong size
    DiscreteFunction small("small", ctx, 10);
    bool thrown = false;
    int code = 0;
    try {
        op(small, w);
    } catch (const Exception &e) {
        thrown = true;
        code = e.code();
    }
    report(thrown && code == FEM_ERR_ARG, "apply on a wrong-size vector throws FEM_ERR_ARG", (double)code);
}

// FemScheme::solve with inhomogeneous Dirichlet data (femscheme.hh:247-253):
// -lap u + u = f with u* = 1 + x - 2 y + x y + x^2 y^2 in Q2, f = -lap u* + u*
// and u = u* on the boundary; the scheme returns u* at every node, and the
// preconditioner comes from fem.solver.preconditioning.method
static void scheme_check(Context &ctx)
{
    const int n = 10, NX = 2 * n + 1;
    auto us = [](double x, double y) { return 1 + x - 2 * y + x * y + x * x * y * y; };
    auto fs = [&](double x, double y) { return -(2 * y * y + 2 * x * x) + us(x, y); };
    fem_op_desc d{};
    d.kind = FEM_OP_CG_Q2_2D;
    d.n[0] = d.n[1] = n;
    d.n[2] = 1;
    d.order = 2;
    d.c = 1.0;
    d.kappa_mode = FEM_KAPPA_CONST;
    d.kappa_const = 1.0;
    d.precision = FEM_F64;
    GalerkinOperator A(ctx, d);
    std::vector<double> gx, gw;
    gaussPoints(3, gx, gw);
    std::vector<double> fq((size_t)n * n * 9), g((size_t)NX * NX);
    for (int ey = 0; ey < n; ++ey)
        for (int ex = 0; ex < n; ++ex)
            for (int qy = 0; qy < 3; ++qy)
                for (int qx = 0; qx < 3; ++qx)
                    fq[((size_t)(ey * n + ex) * 3 + qy) * 3 + qx] = fs((ex + gx[qx]) / n, (ey + gx[qy]) / n);
    for (int J = 0; J < NX; ++J)
        for (int I = 0; I < NX; ++I) g[(size_t)J * NX + I] = us(I / (2.0 * n), J / (2.0 * n));
    A.setBoundaryData(g);
    DiscreteFunction F = A.function("F"), u = A.function("u");
    A.assembleRHS(fq, F);
    SolverParameter p;
    p.set("fem.solver.tolerance", "1e-13").set("fem.solver.errormeasure", "relative")
        .set("fem.solver.preconditioning.method", "jacobi");
    CgInverseOperator inv(ctx, p);
    FemScheme<CgInverseOperator> scheme(A, F, inv);
    SolverInfo info = scheme.solve(u);
    const std::vector<double> uh = u.download();
    double err = 0, mx = 0;
    for (size_t i = 0; i < uh.size(); ++i) {
        err = std::max(err, std::fabs(uh[i] - g[i]));
        mx = std::max(mx, std::fabs(g[i]));
    }
    report(info.converged && info.nonlinearIterations == 0 && info.linearIterations > 0,
           "FemScheme::solve converged in one (linear) Newton step",
           (double)info.linearIterations);
    report(err <= 1e-9 * mx, "FemScheme::solve reproduces u* in Q2 with Dirichlet data", err / mx);
    // backup and restore of the solution (DiscreteFunction::write / read)
    const std::string path = "/tmp/dunefem_amd_inverseoperatortest_u.bin";
    u.write(path, FEM_SPACE_LAGRANGE);
    DiscreteFunction v = A.function("v");
    v.read(path, FEM_SPACE_LAGRANGE);
    const std::vector<double> vh = v.download();
    std::remove(path.c_str());
    report(vh == uh && v.name() == "u", "DiscreteFunction write / read restores DoFs and name", 0.0);
}

// Newton on -div(kappa grad u) + c3 u^3 = f (SIPG DG Q2, cellwise kappa):
// F = S[u*] for a given u*, then NewtonInverseOperator and FemScheme solve
// S[w] = F from w = 0 with matrix-free Jacobians and return u*
static void newton_check(Context &ctx)
{
    const int n = 6, p = 2, E = n * n * n;
    fem_op_desc d = dg_desc(n, p, FEM_F64, false);
    GalerkinOperator S(ctx, d);
    std::vector<double> kc(E);
    for (int e = 0; e < E; ++e) kc[e] = 0.5 + 1.5 * ((e * 37) % 17) / 16.0;
    S.setKappaCell(kc);
    S.setReactionLaw(10.0, 2);
    std::vector<double> us(S.size());
    for (size_t i = 0; i < us.size(); ++i) us[i] = std::sin(0.37 * (double)i) + 0.5 * std::cos(1.3 * (double)i);
    DiscreteFunction ustar = S.function("ustar"), F = S.function("F"), w = S.function("w");
    ustar.upload(us);
    S(ustar, F);
    const double fnorm = std::sqrt(F.normSquaredDofs());
    auto error = [&](const DiscreteFunction &x) {
        const std::vector<double> h = x.download();
        double e2 = 0, n2 = 0;
        for (size_t i = 0; i < h.size(); ++i) {
            e2 += (h[i] - us[i]) * (h[i] - us[i]);
            n2 += us[i] * us[i];
        }
        return std::sqrt(e2 / n2);
    };
    report(S.nonlinear(), "operator with a reaction law is nonlinear()", 0.0);
    GalerkinOperator J(S, GalerkinOperator::JacobianOf{});
    S.jacobian(ustar, J);
    report(!J.nonlinear(), "its Jacobian operator is linear", 0.0);

    NewtonParameter np;
    np.set("fem.solver.nonlinear.tolerance", num(1e-10 * fnorm))
        .set("fem.solver.nonlinear.linear.tolerance", "1e-12")
        .set("fem.solver.nonlinear.linear.errormeasure", "relative")
        .set("fem.solver.nonlinear.linear.preconditioning.method", "blockjacobi");
    NewtonInverseOperator<CgInverseOperator> newton(ctx, np);
    newton.bind(S);
    newton(F, w);
    report(newton.converged() && newton.failed() == FEM_NEWTON_SUCCESS && newton.iterations() >= 2 &&
               newton.iterations() <= 15 && newton.residual() < 1e-10 * fnorm,
           "NewtonInverseOperator solves S[w] = F (iterations)", (double)newton.iterations());
    report(error(w) < 1e-8, "NewtonInverseOperator returns u*", error(w));
    // the same Newton solve with the multigrid the linear solver builds on
    // each Jacobian (the reaction in its levels, updated in place per step)
    {
        NewtonParameter mp;
        mp.set("fem.solver.nonlinear.tolerance", num(1e-10 * fnorm))
            .set("fem.solver.nonlinear.linear.tolerance", "1e-12")
            .set("fem.solver.nonlinear.linear.errormeasure", "relative")
            .set("fem.solver.nonlinear.linear.preconditioning.method", "multigrid");
        NewtonInverseOperator<CgInverseOperator> mg(ctx, mp);
        mg.bind(S);
        DiscreteFunction wm = S.function("wm");
        mg(F, wm);
        report(mg.converged() && mg.iterations() == newton.iterations() &&
                   mg.linearIterations() * 2 < newton.linearIterations(),
               "Newton with the multigrid of each Jacobian (linear iterations)", (double)mg.linearIterations());
        report(error(wm) < 1e-8, "Newton with multigrid returns u*", error(wm));
    }
    // FemScheme over R(u) = S[u] - F
    CgInverseOperator inv(ctx, np.linear());
    NewtonParameter nl;
    nl.setTolerance(1e-10 * fnorm);
    FemScheme<CgInverseOperator> scheme(S, F, inv, nl);
    DiscreteFunction v = S.function("v");
    SolverInfo info = scheme.solve(v);
    report(info.converged && info.nonlinearIterations >= 2 && info.timing.size() == 3,
           "FemScheme::solve of the nonlinear problem (Newton iterations)", (double)info.nonlinearIterations);
    report(error(v) < 1e-8, "FemScheme::solve returns u*", error(v));
}

// Newton on the continuous Q2 Lagrange operator (2D, strong u = 0): the
// reaction law 20 u^2 u, a manufactured u* zero on the boundary, CG with the
// Jacobi preconditioner the solver builds on each Jacobian
static void newton_q2_check(Context &ctx)
{
    const int n = 24;
    fem_op_desc d{};
    d.kind = FEM_OP_CG_Q2_2D;
    d.n[0] = d.n[1] = n;
    d.n[2] = 1;
    d.order = 2;
    d.c = 0.5;
    d.kappa_mode = FEM_KAPPA_CONST;
    d.kappa_const = 1.0;
    d.precision = FEM_F64;
    GalerkinOperator S(ctx, d);
    S.setReactionLaw(20.0, 2);
    const int W = 2 * n + 1;
    std::vector<double> us(S.size(), 0.0);
    for (int j = 1; j < W - 1; ++j)
        for (int i = 1; i < W - 1; ++i) us[(size_t)j * W + i] = std::sin(0.37 * (i + 3 * j)) + 0.5;
    DiscreteFunction ustar = S.function("ustar"), F = S.function("F"), w = S.function("w");
    ustar.upload(us);
    S(ustar, F);
    const double fnorm = std::sqrt(F.normSquaredDofs());
    NewtonParameter np;
    np.set("fem.solver.nonlinear.tolerance", num(1e-10 * fnorm))
        .set("fem.solver.nonlinear.maxiterations", "30")
        .set("fem.solver.nonlinear.linear.tolerance", "1e-12")
        .set("fem.solver.nonlinear.linear.errormeasure", "relative")
        .set("fem.solver.nonlinear.linear.preconditioning.method", "jacobi");
    NewtonInverseOperator<CgInverseOperator> newton(ctx, np);
    newton.bind(S);
    newton(F, w);
    const std::vector<double> h = w.download();
    double e2 = 0, n2 = 0;
    for (size_t i = 0; i < h.size(); ++i) {
        e2 += (h[i] - us[i]) * (h[i] - us[i]);
        n2 += us[i] * us[i];
    }
    report(newton.converged() && newton.iterations() >= 2 && newton.iterations() <= 15,
           "Newton on the Q2 Lagrange operator (iterations)", (double)newton.iterations());
    report(std::sqrt(e2 / n2) < 1e-8, "Newton on Q2 returns u*", std::sqrt(e2 / n2));
}

int main()
{
    try {
        Context ctx(0);
        algorithm<CgInverseOperator>(ctx, " === CgInverseOperator === ", nullptr);
        algorithm<GmresInverseOperator>(ctx, " === GmresInverseOperator === ", nullptr);
        algorithm<FgmresInverseOperator>(ctx, " === FgmresInverseOperator === ", nullptr);
        algorithm<BicgstabInverseOperator>(ctx, " === BicgstabInverseOperator === ", nullptr);
        SolverParameter p;
        p.set("fem.solver.errormeasure", "relative").set("fem.solver.tolerance", "1e-12");
        algorithm<CgInverseOperator>(ctx, " === CgInverseOperator + SolverParameter === ", &p);
        algorithm<GmresInverseOperator>(ctx, " === GmresInverseOperator + SolverParameter === ", &p);
        dg_checks(ctx);
        scheme_check(ctx);
        newton_check(ctx);
        newton_q2_check(ctx);
    } catch (const Exception &e) {
        std::printf("FAIL unexpected exception %d: %s\n", e.code(), e.what());
        ++failures;
    }
    std::printf("%s\n", failures ? "FAIL" : "PASS");
    return failures ? 1 : 0;
}
