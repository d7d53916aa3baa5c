"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py
from the oracle): the oracle must keep reproducing them, the GPU path must
match them, and the quadrature reproduces the reference's quad_test known
answers (quadrature/test/quad_test.cc:246-278)."""
import json
import os

import numpy as np
import pytest

import pyoracle as ora

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def test_quad_test_known_answers_fixture():
    k = json.load(open(os.path.join(G, "quad_test_kats.json")))
    x, w = ora.gauss(3)   # order 5: exact for the cubic hex integrand
    f1 = lambda a: 0.25 - a
    f2 = lambda a, b: (0.25 - a) * (0.1 + 0.7 * a + 0.8 * b)
    f3 = lambda a, b, c: (0.25 - a) * (0.1 + 0.7 * a + 0.8 * b) * (a - 2.0 * b + 0.9 * c)
    r = range(3)
    assert abs(sum(w[i] * f1(x[i]) for i in r) - k["line"]) < 1e-12
    assert abs(sum(w[i] * w[j] * f2(x[i], x[j]) for i in r for j in r) - k["quad"]) < 1e-10
    assert abs(sum(w[i] * w[j] * w[l] * f3(x[i], x[j], x[l]) for i in r for j in r for l in r) - k["hex"]) < 1e-10


def test_oracle_reproduces_fixtures():
    d = load("dg_sipg_q2_2x2x2_cell.npz")
    op = ora.DGOperator(tuple(d["n"]), int(d["p"]), float(d["beta"]), tuple(d["b"]), float(d["c"]),
                        kappa=("cell", d["kcell"]))
    assert rel(op.apply(d["u"]), d["w"]) < 1e-14
    assert rel(op.rhs(d["fq"]), d["F"]) < 1e-14
    assert rel(np.einsum("eii->ei", op.block_diag()).ravel(), d["diag"]) < 1e-14
    d = load("dg_sipg_q3_3x2x2_qp.npz")
    op = ora.DGOperator(tuple(d["n"]), int(d["p"]), float(d["beta"]), tuple(d["b"]), float(d["c"]),
                        kappa=("qp", d["kvol"], (d["kx"], d["ky"], d["kz"])))
    assert rel(op.apply(d["u"]), d["w"]) < 1e-14
    d = load("cg_q2_4x4_qp.npz")
    cop = ora.CGOperator(tuple(d["n"]), c=float(d["c"]), kappa=d["kvol"])
    assert rel(cop.apply(d["u"]), d["w"]) < 1e-14
    assert rel(cop.rhs(d["fq"]), d["F"]) < 1e-14
    assert rel(cop.diag(), d["diag"]) < 1e-14
    d = load("cg_q2_8x8_jacobi_solve.npz")
    cop = ora.CGOperator(tuple(d["n"]), c=float(d["c"]))
    x, its = ora.cg(cop, d["F"], pre=ora.jacobi(1.0 / cop.diag()), tol=1e-10, maxit=5000, criteria=1)
    assert its == int(d["iterations"]) and rel(x, d["x"]) < 1e-13


@pytest.mark.gpu
def test_gpu_matches_fixtures(ctx):
    import dunefem_amd as fa
    d = load("dg_sipg_q2_2x2x2_cell.npz")
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, tuple(d["n"]), int(d["p"]), float(d["beta"]), tuple(d["b"]),
                             float(d["c"]), fa.KAPPA_CELL)
    op.set_kappa_cell(d["kcell"])
    U, W = op.vector(d["u"]), op.vector()
    op(U, W)
    assert rel(W.download(), d["w"]) < 1e-12
    assert rel(op.rhs(d["fq"]).download(), d["F"]) < 1e-12
    assert rel(op.diagonal().download(), d["diag"]) < 1e-12
    d = load("dg_sipg_q3_3x2x2_qp.npz")
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, tuple(d["n"]), int(d["p"]), float(d["beta"]), tuple(d["b"]),
                             float(d["c"]), fa.KAPPA_QP)
    op.set_kappa_qp(d["kvol"], (d["kx"], d["ky"], d["kz"]))
    U, W = op.vector(d["u"]), op.vector()
    op(U, W)
    assert rel(W.download(), d["w"]) < 1e-12
    d = load("cg_q2_4x4_qp.npz")
    cop = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, tuple(d["n"]), 2, c=float(d["c"]), kappa_mode=fa.KAPPA_QP)
    cop.set_kappa_qp(d["kvol"])
    U, W = cop.vector(d["u"]), cop.vector()
    cop(U, W)
    assert rel(W.download(), d["w"]) < 1e-12
    assert rel(cop.rhs(d["fq"]).download(), d["F"]) < 1e-12
    assert rel(cop.diagonal().download(), d["diag"]) < 1e-12
    d = load("cg_q2_8x8_jacobi_solve.npz")
    cop = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, tuple(d["n"]), 2, c=float(d["c"]))
    inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-10, errormeasure=fa.RELATIVE, maxiterations=5000)
    inv.bind(cop, fa.Preconditioner(cop, fa.PRE_JACOBI))
    X = cop.vector()
    inv(cop.vector(d["F"]), X)
    info = inv.info()
    assert info["converged"] and info["iterations"] == int(d["iterations"])
    assert rel(X.download(), d["x"]) < 1e-10
