"""The pointwise and nonlinear reaction on the continuous Q2 Lagrange operator
(2D, strong u = 0 on the boundary), and Newton over it.

S[u] = A u + int (cvol + coef u^pw) u v (+ int r(u, x) v for a compiled
reaction) on the interior rows, S[u]_D = u_D on the Dirichlet rows
(DirichletWrapperOperator, dune/fem/schemes/dirichletwrapper.hh:101-105);
the Jacobian S'(u) is the linear operator with cvol_J = cvol + (pw+1) coef
u^pw (+ dr/du), matrix-free, as for the DG operators.

CPU tests pin the oracle's new Q2 functions by closed forms: the quadrature-
point evaluation of a Q2-representable function against the function at the
points, a constant cvol against the constant c, and v^T A u = int (1 + x)
b^2 = 1/600 for the interior bubble b = x(1-x) y(1-y) with kappa = 0.  GPU
tests compare the HIP path with the oracle (FP64 within 1e-12, FP32 within
2e-5) and run Newton on manufactured problems.
"""
import numpy as np
import pytest

import pyoracle as ora
from dunefem_amd import forms

N2 = (7, 5)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def nodes(n):
    """x, y of the (2 nx + 1) x (2 ny + 1) Q2 node lattice, x fastest"""
    xs, ys = np.linspace(0, 1, 2 * n[0] + 1), np.linspace(0, 1, 2 * n[1] + 1)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    return X.ravel(), Y.ravel()


def qp_xy(n):
    X, Y, _ = forms.qp_coordinates_2d(n, 3, ora.gauss(3)[0])
    return X, Y


def test_oracle_q2_eval_qp_closed_form():
    # a function in Q2 on every element is reproduced exactly at the points
    n = N2
    xn, yn = nodes(n)
    f = lambda x, y: 1 + 2 * x - 3 * y + x * y - x * x * y + 0.5 * x * x * y * y
    op = ora.CGOperator(n)
    X, Y = qp_xy(n)
    assert np.abs(op.eval_qp(f(xn, yn)) - f(X, Y)).max() < 1e-13


def test_oracle_q2_constant_cvol_equals_c():
    n = N2
    rng = np.random.default_rng(0)
    u = rng.normal(size=(2 * n[0] + 1) * (2 * n[1] + 1))
    op = ora.CGOperator(n, c=0.5)
    op.set_reaction_qp(np.full(n[0] * n[1] * 9, 2.0))
    assert rel(op.apply(u), ora.CGOperator(n, c=2.5).apply(u)) < 1e-15
    op.set_reaction_qp(np.full(n[0] * n[1] * 9, 2.0))
    assert rel(op.diag(), ora.CGOperator(n, c=2.5).diag()) < 1e-15


def test_oracle_q2_reaction_closed_form():
    # kappa = 0, cvol = 1 + x: v^T A u = int (1 + x) b^2 = (1/30 + 1/60) / 30
    # for the bubble b (zero on the boundary, in Q2 on every element)
    n = (4, 6)
    xn, yn = nodes(n)
    b = xn * (1 - xn) * yn * (1 - yn)
    op = ora.CGOperator(n, c=0.0, kappa=0.0)
    X, _ = qp_xy(n)
    op.set_reaction_qp(1.0 + X)
    assert abs(b @ op.apply(b) - 1.0 / 600.0) < 1e-15


# ---- GPU --------------------------------------------------------------------------------
def _q2(ctx, fa, n, c=0.3, precision=None):
    return fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, n, 2, c=c,
                               precision=fa.F64 if precision is None else precision)


def _interior(n, u):
    u = u.copy()
    W = 2 * n[0] + 1
    I, J = np.arange(u.size) % W, np.arange(u.size) // W
    u[(I == 0) | (J == 0) | (I == 2 * n[0]) | (J == 2 * n[1])] = 0.0
    return u


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gpu_q2_reaction_law_and_jacobian_match_oracle(ctx, prec):
    import dunefem_amd as fa
    precision, tol = (fa.F64, 1e-12) if prec == "f64" else (fa.F32, 2e-5)
    n, rng = (9, 6), np.random.default_rng(21)
    E = n[0] * n[1]
    cv = rng.uniform(0.0, 5.0, size=E * 9)
    op = _q2(ctx, fa, n, precision=precision)
    op.set_reaction_qp(cv)
    oop = ora.CGOperator(n, c=0.3)
    oop.set_reaction_qp(cv)
    u, w = rng.normal(size=op.N), rng.normal(size=op.N)
    U, Wv, Y = op.vector(u), op.vector(w), op.vector()
    # eval_qp, the linear apply with cvol, its diagonal
    Q = op.eval_qp(U)
    assert rel(Q.download(), oop.eval_qp(u).ravel()) < tol
    op(U, Y)
    assert rel(Y.download(), oop.apply(u)) < tol
    D = op.vector()
    op.diagonal(D)
    assert rel(D.download(), oop.diag()) < tol
    # the law 4 u^2 (cubic): S[u] and S'(u) w
    op.set_reaction_law(4.0, 2)
    assert op.nonlinear()
    op(U, Y)
    uq = oop.eval_qp(u).ravel()
    oS = ora.CGOperator(n, c=0.3)
    oS.set_reaction_qp(cv + 4.0 * uq ** 2)
    errS = rel(Y.download(), oS.apply(u))
    J = op.jacobian(U)
    J(Wv, Y)
    oJ = ora.CGOperator(n, c=0.3)
    oJ.set_reaction_qp(cv + 12.0 * uq ** 2)
    errJ = rel(Y.download(), oJ.apply(w))
    print(prec, "S rel err", errS, "J rel err", errJ, flush=True)
    assert errS < tol and errJ < tol
    with pytest.raises(fa.FemError):   # a nonlinear operator has no diagonal
        op.diagonal(D)


@pytest.mark.gpu
@pytest.mark.parametrize("r", [forms.PREBUILT[2], forms.PREBUILT[3]])
def test_gpu_q2_compiled_reaction_matches_oracle(ctx, r):
    # r(u, x, y, 0) at the Q2 points: S = A u + int r v, J = A + dr/du
    import dunefem_amd as fa
    n, rng = (8, 7), np.random.default_rng(22)
    op = _q2(ctx, fa, n)
    op.set_reaction(r, build=False)
    u, w = rng.normal(size=op.N), rng.normal(size=op.N)
    oop = ora.CGOperator(n, c=0.3)
    X, Y = qp_xy(n)
    uq = oop.eval_qp(u)
    rq, drq = forms.evaluate(r, uq, X, Y, np.zeros_like(X))
    S_ref = oop.apply(u) + _interior(n, oop.rhs(rq))
    oop.set_reaction_qp(drq)
    J_ref = oop.apply(w)
    U, Wv, Yv = op.vector(u), op.vector(w), op.vector()
    op(U, Yv)
    errS = rel(Yv.download(), S_ref)
    op.jacobian(U)(Wv, Yv)
    errJ = rel(Yv.download(), J_ref)
    print("q2 r", r, "S rel err", errS, "J rel err", errJ, flush=True)
    assert errS < 1e-12 and errJ < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("reaction", ["law", "compiled"])
def test_gpu_q2_newton_returns_manufactured_solution(ctx, reaction):
    # -div grad u + 0.3 u + 30 u^3 = F with F = S[u*] for u* zero on the
    # boundary: Newton from 0 with CG + Jacobi returns u*; the compiled
    # u^3 + (1 + x y z) u (z = 0 in 2D) likewise
    import dunefem_amd as fa
    n, rng = (40, 32), np.random.default_rng(23)
    op = _q2(ctx, fa, n)
    if reaction == "law":
        op.set_reaction_law(30.0, 2)
    else:
        op.set_reaction(forms.PREBUILT[2], build=False)
    ustar = op.vector(_interior(n, rng.normal(size=op.N)))
    F = op.vector()
    op(ustar, F)
    tol = 1e-10 * np.sqrt(F.normSquaredDofs())
    inv = fa.NewtonInverseOperator(ctx, {"nonlinear.tolerance": tol},
                                   linear=dict(method=fa.CG, tolerance=1e-12, errormeasure=fa.RELATIVE,
                                               maxiterations=20000, preconditioning=fa.PRE_JACOBI))
    inv.bind(op)
    W = op.vector()
    inv(F, W)
    info = inv.info()
    print("q2 newton", reaction, info, flush=True)
    assert info["converged"] and 2 <= info["iterations"] <= 15
    assert rel(W.download(), ustar.download()) < 1e-8
