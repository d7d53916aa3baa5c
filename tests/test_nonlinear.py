"""Pointwise/nonlinear reaction, matrix-free Jacobians and Newton.

The nonlinear problem is -div(kappa grad u) + c3 u^3 = f in the SIPG DG form,
the reaction evaluated at the volume quadrature points as the reference's
GalerkinOperator evaluates every integrand (dune/fem/schemes/galerkin.hh:333-360):
S[u] = A u + int c3 u^2 u v, its Jacobian J(u) w = A w + int 3 c3 u^2 w v.
Newton follows NewtonInverseOperator::operator()
(dune/fem/solver/newtoninverseoperator.hh:688-802).

CPU tests pin the oracle's reaction term and quadrature-point evaluation:
  * the Legendre basis at the Gauss points against numpy's Legendre series
    (closed form, x fastest);
  * mass orthonormality: sum_q W_q u(x_q)^2 = |u|^2 per element;
  * cvol = const reproduces the constant reaction c (dense and
    sum-factorised oracle paths);
  * with kappa = 0 and u = s_e on each element, S[u] = c3 s_e^3 |K| on the
    constant mode (closed form);
  * the Jacobian form against a central difference of S, which a wrong
    factor (1 instead of 3) fails.
GPU tests compare the HIP path with the oracle (FP64 within 1e-12, FP32
within 2e-5) and check Newton against a dense Newton on the oracle.
"""
import sys

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

import pyoracle as ora

TOL64 = 1e-12
TOL32 = 2e-5


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cubic_cvol(op, u, c3, factor=1.0):
    """factor * c3 * u^2 at the volume quadrature points (oracle evaluation)"""
    uq = op.eval_qp(u)
    return factor * c3 * uq * uq


def oracle_S(op, u, c3):
    """S[u] = A u + int c3 u^3 v on the oracle"""
    op.set_reaction_qp(cubic_cvol(op, u, c3))
    w = op.apply(u)
    op.set_reaction_qp(None)
    return w


def oracle_J(op, u, c3, w, factor=3.0):
    """J(u) w = A w + int factor c3 u^2 w v on the oracle"""
    op.set_reaction_qp(cubic_cvol(op, u, c3, factor))
    y = op.apply(w)
    op.set_reaction_qp(None)
    return y


# ---- CPU: the oracle ------------------------------------------------------------------
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_oracle_eval_qp_closed_form(p):
    # mode (a, b, c) of element 0 alone: u(x_q) = P_a(x) P_b(y) P_c(z), with
    # P_n(x) = sqrt(2n+1) L_n(2x - 1) the orthonormal Legendre basis on [0, 1]
    m = p + 1
    op = ora.DGOperator((2, 1, 1), p, 1.0)
    x, _ = ora.gauss(m)
    rng = np.random.default_rng(p)
    for _ in range(3):
        a, b, c = rng.integers(0, m, size=3)
        u = np.zeros(op.N)
        u[a * m * m + b * m + c] = 1.0          # DoF (a, b, c): x mode a slowest
        uq = op.eval_qp(u)
        P = [np.sqrt(2 * k + 1) * npleg.legval(2 * np.asarray(x) - 1, [0] * k + [1]) for k in (a, b, c)]
        exact = np.einsum("z,y,x->zyx", P[2], P[1], P[0]).ravel()   # qz slowest, qx fastest
        assert np.abs(uq[0] - exact).max() < 1e-13
        assert np.abs(uq[1]).max() == 0.0


@pytest.mark.parametrize("p", [1, 3])
def test_oracle_eval_qp_mass_orthonormal(p):
    m = p + 1
    op = ora.DGOperator((3, 2, 2), p, 1.0)
    _, w = ora.gauss(m)
    W = np.einsum("z,y,x->zyx", w, w, w).ravel()
    u = np.random.default_rng(1).normal(size=op.N)
    uq = op.eval_qp(u)
    E = op.N // m ** 3
    lhs = (uq * uq) @ W
    rhs = (u.reshape(E, m ** 3) ** 2).sum(axis=1)
    assert np.abs(lhs - rhs).max() < 1e-12 * rhs.max()


@pytest.mark.parametrize("p", [1, 2, 4])
def test_oracle_constant_cvol_equals_c(p):
    n, rng = (3, 2, 4), np.random.default_rng(p)
    kc = rng.uniform(0.5, 2.0, size=n[0] * n[1] * n[2])
    b = (0.3, -0.2, 0.1)
    with_c = ora.DGOperator(n, p, 10.0 * p * p, b, 1.7, kappa=("cell", kc))
    with_cv = ora.DGOperator(n, p, 10.0 * p * p, b, 0.0, kappa=("cell", kc))
    with_cv.set_reaction_qp(np.full(with_cv.N, 1.7))
    u = rng.normal(size=with_c.N)
    ref = with_c.apply(u)
    assert rel(with_cv.apply(u), ref) < 1e-14
    assert rel(with_cv.apply_sf(u), with_c.apply_sf(u)) < 1e-14


@pytest.mark.parametrize("p", [1, 2])
def test_oracle_cubic_reaction_closed_form(p):
    # kappa = 0, b = 0: only the reaction remains; u = s_e on element e gives
    # S[u]_e = c3 s_e^3 |K| on the constant mode and 0 on the others
    n, m, c3 = (2, 3, 2), p + 1, 2.5
    op = ora.DGOperator(n, p, 0.0, kappa=0.0)
    E = n[0] * n[1] * n[2]
    s = np.random.default_rng(2).uniform(-2, 2, size=E)
    u = np.zeros((E, m ** 3))
    u[:, 0] = s
    w = oracle_S(op, u.ravel(), c3).reshape(E, m ** 3)
    vol = 1.0 / (n[0] * n[1] * n[2])
    assert np.abs(w[:, 0] - c3 * s ** 3 * vol).max() < 1e-14
    assert np.abs(w[:, 1:]).max() < 1e-14


def test_oracle_jacobian_matches_central_difference():
    # S(u + e w) - S(u - e w) = 2 e J(u) w + 2 e^3 int c3 w^3 v: the central
    # difference meets J within O(e^2); the form with factor 1 misses by O(1)
    n, p, c3 = (3, 2, 2), 2, 4.0
    kc = np.random.default_rng(3).uniform(0.5, 2.0, size=12)
    op = ora.DGOperator(n, p, 10.0 * p * p, kappa=("cell", kc))
    rng = np.random.default_rng(4)
    u, w = rng.normal(size=op.N), rng.normal(size=op.N)
    eps = 1e-4
    fd = (oracle_S(op, u + eps * w, c3) - oracle_S(op, u - eps * w, c3)) / (2 * eps)
    assert rel(fd, oracle_J(op, u, c3, w)) < 1e-7
    assert rel(fd, oracle_J(op, u, c3, w, factor=1.0)) > 1e-2


# ---- GPU --------------------------------------------------------------------------------
def _dg(ctx, fa, n, p, kc, b=(0.0, 0.0, 0.0), c=0.0, precision=None):
    precision = fa.F64 if precision is None else precision
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, b, c, fa.KAPPA_CELL, precision=precision)
    op.set_kappa_cell(kc)
    return op


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_gpu_reaction_qp_and_eval_qp_match_oracle(ctx, p):
    import dunefem_amd as fa
    n, rng = (5, 3, 4), np.random.default_rng(10 + p)
    kc = rng.uniform(0.5, 2.0, size=60)
    b = (0.7, -0.3, 0.2)
    oop = ora.DGOperator(n, p, 10.0 * p * p, b, 0.3, kappa=("cell", kc))
    cv = rng.uniform(0.0, 3.0, size=oop.N)
    oop.set_reaction_qp(cv)
    u = rng.normal(size=oop.N)
    ref = oop.apply(u)
    for prec, tol in ((fa.F64, TOL64), (fa.F32, TOL32)):
        op = _dg(ctx, fa, n, p, kc, b, 0.3, prec)
        op.set_reaction_qp(cv)
        U, W = op.vector(u), op.vector()
        op(U, W)
        err = rel(W.download(), ref)
        print("p", p, "prec", prec, "apply with cvol rel err", err, flush=True)
        assert err < tol
        uq = op.eval_qp(U).download()
        assert rel(uq, oop.eval_qp(u).ravel()) < (1e-13 if prec == fa.F64 else 1e-6)
        op.set_reaction_qp(None)
        op(U, W)
        oop0 = ora.DGOperator(n, p, 10.0 * p * p, b, 0.3, kappa=("cell", kc))
        assert rel(W.download(), oop0.apply(u)) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("p,pw", [(1, 2), (2, 1), (3, 2), (4, 3)])
def test_gpu_nonlinear_apply_and_jacobian_match_oracle(ctx, p, pw):
    import dunefem_amd as fa
    n, rng, coef = (4, 3, 3), np.random.default_rng(20 + p), 1.5
    kc = rng.uniform(0.5, 2.0, size=36)
    oop = ora.DGOperator(n, p, 10.0 * p * p, kappa=("cell", kc))
    u, w = rng.normal(size=oop.N), rng.normal(size=oop.N)
    uq = oop.eval_qp(u)
    oop.set_reaction_qp(coef * uq ** pw)
    S_ref = oop.apply(u)
    oop.set_reaction_qp((pw + 1) * coef * uq ** pw)
    J_ref = oop.apply(w)
    op = _dg(ctx, fa, n, p, kc)
    op.set_reaction_law(coef, pw)
    assert op.nonlinear()
    U, Wv, Y = op.vector(u), op.vector(w), op.vector()
    op(U, Y)
    errS = rel(Y.download(), S_ref)
    jop = op.jacobian(U)
    assert not jop.nonlinear()
    jop(Wv, Y)
    errJ = rel(Y.download(), J_ref)
    print("p", p, "pw", pw, "S rel err", errS, "J rel err", errJ, flush=True)
    assert errS < TOL64 and errJ < TOL64
    # the Jacobian of the device apply against its own central difference
    eps = 1e-4
    Sp, Sm = op.vector(), op.vector()
    op(op.vector(u + eps * w), Sp)
    op(op.vector(u - eps * w), Sm)
    fd = (Sp.download() - Sm.download()) / (2 * eps)
    assert rel(fd, J_ref) < 1e-6
    # a nonlinear operator has no diagonal, blocks or linear solve of its own
    with pytest.raises(fa.FemError):
        op.diagonal()
    with pytest.raises(fa.FemError):
        fa.InverseOperator(ctx, method=fa.CG).bind(op)
    # the Jacobian's diagonal and CSR matrix include the reaction
    d = jop.diagonal().download()
    A = fa.SparseRowMatrix(jop)
    assert rel(A.diagonal().download(), d) < 1e-12
    x = op.vector(rng.normal(size=op.N))
    y1, y2 = op.vector(), op.vector()
    A(x, y1)
    jop(x, y2)
    assert rel(y1.download(), y2.download()) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("p,pw", [(1, 2), (2, 2), (3, 1), (4, 2)])
def test_gpu_reaction_with_pointwise_kappa_matches_oracle(ctx, p, pw):
    # the reaction variant of the pointwise-kappa apply (kappa at volume and
    # face points, cvol and the law): S[u] and S'(u) w against the oracle,
    # FP64 and FP32, with advection
    import dunefem_amd as fa
    from dunefem_amd import workloads as wl
    n, rng, coef, b = (4, 3, 5), np.random.default_rng(60 + p), 2.0, (0.5, -0.25, 0.3)
    P = wl.problem1(n=n, p=p, seed=60 + p, with_u=False)
    kq = ("qp", P["kvol"], P["kfaces"])
    cv = rng.uniform(0.0, 2.0, size=n[0] * n[1] * n[2] * (p + 1) ** 3)
    oop = ora.DGOperator(n, p, 10.0 * p * p, b, 0.2, kappa=kq)
    u, w = rng.normal(size=oop.N), rng.normal(size=oop.N)
    uq = oop.eval_qp(u).ravel()
    oop.set_reaction_qp(cv + coef * uq ** pw)
    S_ref = oop.apply(u)
    oop.set_reaction_qp(cv + (pw + 1) * coef * uq ** pw)
    J_ref = oop.apply(w)
    for prec, tol in ((fa.F64, TOL64), (fa.F32, TOL32)):
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, b, 0.2, fa.KAPPA_QP, precision=prec)
        op.set_kappa_qp(P["kvol"], P["kfaces"])
        op.set_reaction_qp(cv)
        op.set_reaction_law(coef, pw)
        U, Wv, Y = op.vector(u), op.vector(w), op.vector()
        op(U, Y)
        errS = rel(Y.download(), S_ref)
        op.jacobian(U)(Wv, Y)
        errJ = rel(Y.download(), J_ref)
        print("pointwise kappa p", p, "prec", prec, "S rel err", errS, "J rel err", errJ, flush=True)
        assert errS < tol and errJ < tol


@pytest.mark.gpu
def test_gpu_block_jacobi_exact_for_element_constant_reaction(ctx):
    # an element-constant cvol keeps the element block a sum of Kronecker
    # terms: the block-Jacobi inverse (eigen data shifted by the reaction)
    # inverts the probed element blocks exactly
    import dunefem_amd as fa
    n, p, rng = (3, 4, 2), 3, np.random.default_rng(5)
    E, m = 24, 4
    op = _dg(ctx, fa, n, p, rng.uniform(0.5, 2.0, size=E), c=0.4)
    op.set_reaction_qp(np.repeat(rng.uniform(0, 50, size=E), m ** 3))
    B = op.block_diagonal()
    v = rng.normal(size=(E, m ** 3))
    Bv = np.einsum("eij,ej->ei", B, v)
    P = fa.Preconditioner(op, fa.PRE_BLOCK_JACOBI)
    y = op.vector()
    P(op.vector(Bv.ravel()), y)
    assert rel(y.download(), v.ravel()) < 1e-10


def oracle_newton(oop, c3, u_rhs, w0, tol, maxit=30):
    """dense Newton on the oracle: the Jacobian assembled column by column"""
    w = w0.copy()
    N = oop.N
    its = 0
    r = oracle_S(oop, w, c3) - u_rhs
    while np.linalg.norm(r) >= tol and its < maxit:
        oop.set_reaction_qp(cubic_cvol(oop, w, c3, 3.0))
        J = np.empty((N, N))
        e = np.zeros(N)
        for j in range(N):
            e[j] = 1.0
            J[:, j] = oop.apply(e)
            e[j] = 0.0
        oop.set_reaction_qp(None)
        w -= np.linalg.solve(J, r)
        r = oracle_S(oop, w, c3) - u_rhs
        its += 1
    return w, its, np.linalg.norm(r)


@pytest.mark.gpu
def test_gpu_newton_matches_oracle_newton(ctx):
    import dunefem_amd as fa
    n, p, c3 = (3, 3, 3), 2, 5.0
    rng = np.random.default_rng(7)
    kc = rng.uniform(0.5, 2.0, size=27)
    oop = ora.DGOperator(n, p, 10.0 * p * p, kappa=("cell", kc))
    ustar = rng.normal(size=oop.N)
    F = oracle_S(oop, ustar, c3)
    tol = 1e-9 * np.linalg.norm(F)
    w_ref, its_ref, res_ref = oracle_newton(oop, c3, F, np.zeros(oop.N), tol)
    assert rel(w_ref, ustar) < 1e-8
    op = _dg(ctx, fa, n, p, kc)
    op.set_reaction_law(c3, 2)
    inv = fa.NewtonInverseOperator(ctx, {"nonlinear.tolerance": tol},
                                   linear=dict(method=fa.CG, tolerance=1e-13, errormeasure=fa.RELATIVE,
                                               preconditioning=fa.PRE_BLOCK_JACOBI))
    inv.bind(op)
    W = op.vector()
    inv(op.vector(F), W)
    info = inv.info()
    print("newton", info, "oracle its", its_ref, "res", res_ref, flush=True)
    assert info["converged"] and info["failure"] == "Success"
    assert info["iterations"] == its_ref
    assert info["residual"] < tol
    assert rel(W.download(), w_ref) < 1e-9
    assert len(info["timing"]) == 3 and info["timing"][0] >= info["timing"][1] >= 0


@pytest.mark.gpu
def test_gpu_newton_linear_operator_takes_one_step(ctx):
    # a linear operator: one step and no nonlinear iteration counted
    # (newtoninverseoperator.hh:786-790)
    import dunefem_amd as fa
    n, p = (4, 4, 4), 2
    rng = np.random.default_rng(8)
    op = _dg(ctx, fa, n, p, rng.uniform(0.5, 2.0, size=64))
    b = op.vector(rng.normal(size=op.N))
    inv = fa.NewtonInverseOperator(ctx, {"fem.solver.nonlinear.tolerance": 1e-8},
                                   linear=dict(method=fa.CG, tolerance=1e-12, errormeasure=fa.RELATIVE,
                                               preconditioning=fa.PRE_BLOCK_JACOBI))
    inv.bind(op)
    W = op.vector()
    inv(b, W)
    info = inv.info()
    assert info["converged"] and info["iterations"] == 0 and info["linear_iterations"] > 0
    R = op.vector()
    op(W, R)
    R.axpy(-1.0, b)
    assert np.sqrt(R.normSquaredDofs()) <= 1e-10 * np.sqrt(b.normSquaredDofs())


@pytest.mark.gpu
def test_gpu_newton_options_and_failures(ctx):
    import dunefem_amd as fa
    n, p, c3 = (4, 4, 4), 2, 20.0
    rng = np.random.default_rng(9)
    kc = rng.uniform(0.5, 2.0, size=64)
    op = _dg(ctx, fa, n, p, kc)
    op.set_reaction_law(c3, 2)
    ustar = op.vector(2.0 * rng.normal(size=op.N))
    F = op.vector()
    op(ustar, F)
    tol = 1e-9 * np.sqrt(F.normSquaredDofs())
    lin = dict(method=fa.CG, tolerance=1e-12, errormeasure=fa.RELATIVE, preconditioning=fa.PRE_BLOCK_JACOBI)

    def run(params, linear=lin, start=None):
        inv = fa.NewtonInverseOperator(ctx, dict(params, **{"nonlinear.tolerance": tol}), linear=linear)
        inv.bind(op)
        W = op.vector(start if start is not None else np.zeros(op.N))
        inv(F, W)
        inv.unbind()
        return inv.info(), W.download()

    plain, w1 = run({})
    print("plain", plain, flush=True)
    assert plain["converged"] and rel(w1, ustar.download()) < 1e-8
    # the linear solver's maxiterations bounds the total over all steps
    # (newtoninverseoperator.hh:246-250): a limit just above the total the
    # solve needs is enough
    tight, _ = run({}, linear=dict(lin, maxiterations=plain["linear_iterations"] + 5))
    assert tight["converged"] and tight["linear_iterations"] == plain["linear_iterations"]
    ls, w2 = run({"nonlinear.linesearch": "simple"}, start=10.0 * ustar.download())
    print("line search from 10 u*", ls, flush=True)
    assert ls["converged"] and rel(w2, ustar.download()) < 1e-8
    # Eisenstat-Walker: inexact linear solves, fewer linear iterations in total
    ew, w3 = run({"nonlinear.forcing": "eisenstatwalker"},
                 linear=dict(method=fa.CG, preconditioning=fa.PRE_BLOCK_JACOBI))
    print("eisenstat-walker", ew, flush=True)
    assert ew["converged"] and ew["residual"] < tol and rel(w3, ustar.download()) < 1e-5
    assert ew["linear_iterations"] < plain["linear_iterations"]
    # simplified Newton (the Jacobian of the first step only) from a start
    # near u*: linear convergence, more steps than Newton from the same start
    near = ustar.download() * (1.0 + 0.01 * rng.normal(size=op.N))
    full, _ = run({}, start=near)
    simp, w4 = run({"nonlinear.simplified": "true", "nonlinear.maxiterations": 200}, start=near)
    print("from near u*: newton", full, "simplified", simp, flush=True)
    assert full["converged"] and simp["converged"] and simp["iterations"] > full["iterations"]
    assert rel(w4, ustar.download()) < 1e-7
    # failure codes (newtoninverseoperator.hh:568-584)
    few, _ = run({"nonlinear.maxiterations": 1})
    assert not few["converged"] and few["failure"] == "TooManyIterations" and few["iterations"] == 1
    lim, _ = run({"nonlinear.maxlineariterations": 3})
    assert not lim["converged"] and lim["failure"] in ("TooManyLinearIterations", "LinearSolverFailed")
    with pytest.raises(fa.FemError):
        fa.NewtonInverseOperator(ctx, {"nonlinear.forcing": "eisenstatwalker"},
                                 linear=dict(errormeasure=fa.ABSOLUTE))


@pytest.mark.gpu
def test_gpu_newton_with_fp32_refinement_solves(ctx):
    # each Newton system solved by FP64 refinement around FP32 inner CG on an
    # FP32 twin of the Jacobian: the same Newton steps as with FP64 solves,
    # the same solution; with Eisenstat-Walker forcing (residualreduction in
    # the refinement loop) too, and with a multigrid of the diffusion part
    # (FP32) as the inner preconditioner given at bind
    import dunefem_amd as fa
    n, p, c3 = (6, 6, 6), 2, 10.0
    rng = np.random.default_rng(12)
    kc = rng.uniform(0.5, 2.0, size=216)
    op = _dg(ctx, fa, n, p, kc)
    op.set_reaction_law(c3, 2)
    ustar = op.vector(rng.normal(size=op.N))
    F = op.vector()
    op(ustar, F)
    tol = 1e-9 * np.sqrt(F.normSquaredDofs())

    def run(nl, linear, pre=None):
        inv = fa.NewtonInverseOperator(ctx, dict(nl, **{"nonlinear.tolerance": tol}), linear=linear)
        inv.bind(op, pre)
        W = op.vector()
        inv(F, W)
        inv.unbind()
        return inv.info(), W.download()

    exact = dict(method=fa.CG, tolerance=1e-12, errormeasure=fa.RELATIVE, preconditioning=fa.PRE_BLOCK_JACOBI)
    ref, w_ref = run({}, exact)
    ir, w_ir = run({}, dict(exact, refinement=1, inner_tolerance=1e-4, inner_maxiterations=2000))
    print("newton fp64 solves", ref, "with fp32 refinement", ir, flush=True)
    assert ref["converged"] and ir["converged"] and ir["iterations"] == ref["iterations"]
    assert rel(w_ir, ustar.download()) < 1e-8 and rel(w_ref, ustar.download()) < 1e-8
    ew, w_ew = run({"nonlinear.forcing": "eisenstatwalker"},
                   dict(method=fa.CG, refinement=1, inner_tolerance=1e-3, inner_maxiterations=2000,
                        preconditioning=fa.PRE_BLOCK_JACOBI))
    print("eisenstat-walker with fp32 refinement", ew, flush=True)
    assert ew["converged"] and ew["residual"] < tol and rel(w_ew, ustar.download()) < 1e-6
    lin32 = _dg(ctx, fa, n, p, kc, precision=fa.F32)
    mg32 = fa.Preconditioner(lin32, fa.PRE_MULTIGRID)
    mg, w_mg = run({"nonlinear.forcing": "eisenstatwalker"},
                   dict(method=fa.CG, refinement=1, inner_tolerance=1e-3, inner_maxiterations=2000), mg32)
    print("eisenstat-walker, fp32 refinement, fp32 multigrid of the diffusion part", mg, flush=True)
    assert mg["converged"] and mg["residual"] < tol and rel(w_mg, ustar.download()) < 1e-6


@pytest.mark.gpu
def test_gpu_multigrid_with_reaction(ctx):
    # the V-cycle of an operator with c and a pointwise reaction (a Jacobian's
    # form): CG + multigrid reaches the solution CG + block Jacobi reaches,
    # in far fewer iterations; and Newton with the multigrid rebuilt on each
    # Jacobian (preconditioning.method multigrid) takes the steps Newton with
    # block Jacobi takes
    import dunefem_amd as fa
    n, p = (12, 12, 12), 2
    rng = np.random.default_rng(14)
    kc = 10.0 ** rng.uniform(-1, 1, size=12 ** 3)
    op = _dg(ctx, fa, n, p, kc, c=0.5)
    op.set_reaction_qp(rng.uniform(0.0, 200.0, size=op.N))
    b = op.vector(rng.normal(size=op.N))
    sols = {}
    for name, pre, maxit in (("bj", fa.PRE_BLOCK_JACOBI, 5000), ("mg", fa.PRE_MULTIGRID, 500)):
        inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-11, errormeasure=fa.RELATIVE, maxiterations=maxit)
        inv.bind(op, fa.Preconditioner(op, pre))
        X = op.vector()
        inv(b, X)
        sols[name] = (inv.info(), X.download())
        sys.stderr.write("%s %s\n" % (name, inv.info()))
        sys.stderr.flush()
    (ibj, xbj), (img, xmg) = sols["bj"], sols["mg"]
    assert ibj["converged"] and img["converged"] and img["true_rel_residual"] < 1e-10
    assert img["iterations"] * 5 < ibj["iterations"]
    assert rel(xmg, xbj) < 1e-8
    # Newton on the cubic law with the multigrid of each Jacobian
    S = _dg(ctx, fa, n, p, kc)
    S.set_reaction_law(20.0, 2)
    ustar = S.vector(rng.normal(size=S.N))
    F = S.vector()
    S(ustar, F)
    tol = 1e-9 * np.sqrt(F.normSquaredDofs())
    res = {}
    for name, pre in (("bj", fa.PRE_BLOCK_JACOBI), ("mg", fa.PRE_MULTIGRID)):
        # bounded: at most 30 Newton steps and 60000 linear iterations in all
        inv = fa.NewtonInverseOperator(ctx, {"nonlinear.tolerance": tol, "nonlinear.maxiterations": 30},
                                       linear=dict(method=fa.CG, tolerance=1e-11, errormeasure=fa.RELATIVE,
                                                   maxiterations=60000, preconditioning=pre))
        inv.bind(S)
        W = S.vector()
        inv(F, W)
        res[name] = (inv.info(), W.download())
        sys.stderr.write("newton %s %s\n" % (name, inv.info()))
        sys.stderr.flush()
    assert res["mg"][0]["converged"] and res["mg"][0]["iterations"] == res["bj"][0]["iterations"]
    assert res["mg"][0]["linear_iterations"] * 4 < res["bj"][0]["linear_iterations"]
    assert rel(res["mg"][1], ustar.download()) < 1e-8


@pytest.mark.gpu
def test_gpu_scheme_solves_nonlinear_problem(ctx):
    # the scheme path: S[u] = A u + c3 u^3 with the affine part F of a
    # manufactured u*: scheme.solve from u = 0 returns u*, the info dict has
    # the reference keys and counts the Newton iterations
    import dunefem_amd as fa
    n, p, c3 = (4, 3, 5), 3, 8.0
    rng = np.random.default_rng(11)
    op = _dg(ctx, fa, n, p, rng.uniform(0.5, 2.0, size=60))
    op.set_reaction_law(c3, 2)
    ustar = rng.normal(size=op.N)
    F = op.vector()
    op(op.vector(ustar), F)
    scheme = fa.Scheme(op, F, parameters={"fem.solver.nonlinear.tolerance": 1e-10 * np.sqrt(F.normSquaredDofs()),
                                          "fem.solver.nonlinear.linear.method": "cg",
                                          "fem.solver.nonlinear.linear.preconditioning.method": "blockjacobi"},
                       tolerance=1e-12, errormeasure=fa.RELATIVE)
    U = op.vector()
    info = scheme.solve(U)
    print("nonlinear scheme solve", info, flush=True)
    assert set(info) == {"converged", "iterations", "linear_iterations", "timing"}
    assert info["converged"] and 2 <= info["iterations"] <= 12
    assert rel(U.download(), ustar) < 1e-8
