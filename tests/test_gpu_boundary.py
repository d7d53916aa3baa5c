"""The drop-in boundary beyond apply and solve: the element blocks of the
linearisation, the solver's bind-time validation, and the preconditioner a
solver builds for itself from fem.solver.preconditioning.method."""
import numpy as np
import pytest

import dunefem_amd as fa
import pyoracle as ora
from dunefem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def qp_kappa(rng, n, m):
    E = n[0] * n[1] * n[2]
    kvol = rng.uniform(0.2, 5.0, size=(E, m ** 3))
    kf = [rng.uniform(0.2, 5.0, size=((n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2)), m * m))
          for d in range(3)]
    return kvol, kf


@pytest.mark.parametrize("p,kmode,b", [(1, "cell", (0, 0, 0)), (2, "qp", (1.0, -0.5, 0.25)),
                                       (3, "cell", (0.3, 0.2, -1.0)), (4, "qp", (0, 0, 0))])
def test_block_diagonal_matches_oracle(ctx, p, kmode, b):
    rng = np.random.default_rng(10 + p)
    n = (3, 2, 4)
    E = 24
    m = p + 1
    if kmode == "cell":
        kc = rng.uniform(0.1, 10.0, size=E)
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, b, 0.0, fa.KAPPA_CELL)
        op.set_kappa_cell(kc)
        oop = ora.DGOperator(n, p, 10.0 * p * p, b, 0.0, kappa=("cell", kc))
    else:
        kvol, kf = qp_kappa(rng, n, m)
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, b, 0.0, fa.KAPPA_QP)
        op.set_kappa_qp(kvol, kf)
        oop = ora.DGOperator(n, p, 10.0 * p * p, b, 0.0, kappa=("qp", kvol, tuple(kf)))
    blocks = op.block_diagonal()
    ref = oop.block_diag()
    assert blocks.shape == ref.shape
    assert np.abs(blocks - ref).max() <= 1e-12 * np.abs(ref).max()


def test_block_diagonal_fp32_operator(ctx):
    rng = np.random.default_rng(3)
    n, p = (2, 3, 2), 2
    kc = rng.uniform(0.5, 2.0, size=12)
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 40.0, (0, 0, 0), 0.0, fa.KAPPA_CELL, precision=fa.F32)
    op.set_kappa_cell(kc)
    ref = ora.DGOperator(n, p, 40.0, (0, 0, 0), 0.0, kappa=("cell", kc)).block_diag()
    assert np.abs(op.block_diagonal() - ref).max() <= 1e-5 * np.abs(ref).max()


def dg_cell(ctx, n, p, kc, prec=fa.F64):
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, (0, 0, 0), 0.0, fa.KAPPA_CELL, precision=prec)
    op.set_kappa_cell(kc)
    return op


def test_bind_rejects_fp64_inner_operator_and_learned_with_cg(ctx):
    rng = np.random.default_rng(4)
    n, p = (4, 4, 4), 2
    kc = rng.uniform(0.5, 2.0, size=64)
    op64 = dg_cell(ctx, n, p, kc)
    op64b = dg_cell(ctx, n, p, kc)
    inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, refinement=1)
    with pytest.raises(fa.FemError, match="inner operator must be FP32"):
        inv.bind(op64, None, op64b, None)
    op32 = dg_cell(ctx, n, p, kc, fa.F32)
    small32 = dg_cell(ctx, (2, 2, 2), p, kc[:8], fa.F32)
    with pytest.raises(fa.FemError, match="inner preconditioner"):
        inv.bind(op64, None, op32, fa.Preconditioner(small32, fa.PRE_BLOCK_JACOBI))
    inv.bind(op64, None, op32, fa.Preconditioner(op32, fa.PRE_BLOCK_JACOBI))   # valid
    net = fa.CNN(ctx, n)
    net.set_coefficient(kc, 40.0)
    lp = fa.LearnedPreconditioner(op64, net, coarse_steps=3)
    for method in (fa.CG, fa.GMRES):
        inv2 = fa.InverseOperator(ctx, method=method, tolerance=1e-8)
        with pytest.raises(fa.FemError, match="nonlinear"):
            inv2.bind(op64, lp)
    fa.InverseOperator(ctx, method=fa.FGMRES, tolerance=1e-8).bind(op64, lp)


def test_solver_builds_preconditioner_from_parameter_key(ctx):
    # a DG solve through "fem.solver.preconditioning.method": "blockjacobi"
    # takes the same iterations as one bound to an explicit block-Jacobi object
    rng = np.random.default_rng(6)
    n, p = (6, 5, 4), 2
    kc = 10.0 ** (2 * rng.integers(0, 2, size=120))
    op = dg_cell(ctx, n, p, kc)
    f = rng.normal(size=op.N)
    B = op.vector(f)
    its, res = {}, {}
    for name, pre, keys in (("object", fa.Preconditioner(op, fa.PRE_BLOCK_JACOBI), {}),
                            ("key", None, {"fem.solver.preconditioning.method": "blockjacobi"})):
        inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-9, errormeasure=fa.RELATIVE, maxiterations=3000,
                                 **keys)
        inv.bind(op, pre) if pre else inv.bind(op)
        X = op.vector()
        inv(B, X)
        info = inv.info()
        # CG stops on sqrt(<r, M^-1 r>); with kappa in {1, 100} the true
        # residual sits up to two orders above the tolerance
        assert info["converged"] and info["true_rel_residual"] < 1e-6, name
        its[name], res[name] = info["iterations"], info["true_rel_residual"]
    assert its["object"] == its["key"], its
    assert abs(res["object"] - res["key"]) <= 1e-6 * res["object"], res
    # the continuous Q2 problem through "jacobi", against the oracle's CG + Jacobi
    P = wl.problem0(n=32)
    gop = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, (32, 32), 2, c=1.0)
    oop = ora.CGOperator((32, 32), c=1.0)
    F = oop.rhs(P["fq"])
    _, ito = ora.cg(oop, F, pre=ora.jacobi(1.0 / oop.diag()), tol=1e-10, maxit=5000, criteria=1)
    inv = fa.InverseOperator(ctx, **{"fem.solver.method": "cg", "fem.solver.tolerance": "1e-10",
                                     "fem.solver.errormeasure": "relative",
                                     "fem.solver.preconditioning.method": "jacobi"})
    inv.bind(gop)
    X = gop.vector()
    inv(gop.vector(F), X)
    assert abs(inv.iterations() - ito) <= 1 and inv.info()["true_rel_residual"] < 1e-9


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_sipg_boundary_data_consistency_on_gpu(ctx, p):
    # the closed-form identity A u*_h = F(f) + F_bnd(g) for a polynomial u*,
    # pointwise kappa = 1 + x + 2y + 3z, advection and reaction (through the
    # C ABI; the oracle satisfies the same identity), and the GPU load vector
    # with boundary data against the oracle's
    from manufactured import sipg_problem
    n, b, c = (4, 3, 5), (0.7, -0.4, 0.3), 0.5
    M = sipg_problem(n, p, b, c, seed=p)
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, b, c, fa.KAPPA_QP)
    op.set_kappa_qp(M["kvol"], M["kfaces"])
    op.set_boundary_data(M["gfaces"])
    U, W = op.vector(M["u"]), op.vector()
    op(U, W)
    F = op.rhs(M["fq"]).download()
    lhs = W.download()
    err = np.abs(lhs - F).max() / np.abs(F).max()
    print("p", p, "A u* - F rel", err, flush=True)
    assert err <= 1e-11
    oop = ora.DGOperator(n, p, 10.0 * p * p, b, c, kappa=("qp", M["kvol"], tuple(M["kfaces"])))
    Fo = oop.rhs_bnd(M["gfaces"], oop.rhs(M["fq"]))
    assert np.abs(F - Fo).max() <= 1e-12 * np.abs(Fo).max()


def test_q2_scheme_solve_with_dirichlet_data_reproduces_polynomial(ctx):
    # continuous Q2, -div(kappa grad u) + u = f, kappa = 1 + x + 2y, u = u* on
    # the boundary with u* in Q2: the scheme solve (constraints, one Newton
    # step with CG + Jacobi) returns u* at every node; a linear operator counts
    # no nonlinear iteration (newtoninverseoperator.hh:786-790)
    from manufactured import q2_problem
    n = (12, 9)
    M = q2_problem(n, c=1.0, seed=3)
    op = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, n, 2, c=1.0, kappa_mode=fa.KAPPA_QP)
    op.set_kappa_qp(M["kvol"])
    op.set_boundary_data(M["nodes"])
    F = op.rhs(M["fq"])
    scheme = fa.Scheme(op, F, fa.Preconditioner(op, fa.PRE_JACOBI), method=fa.CG, tolerance=1e-13,
                       errormeasure=fa.RELATIVE, maxiterations=5000)
    U = op.vector()
    info = scheme.solve(U)
    print("Q2 scheme solve:", info, flush=True)
    assert set(info) == {"converged", "iterations", "linear_iterations", "timing"}
    assert info["converged"] and info["iterations"] == 0 and info["linear_iterations"] > 0
    u = U.download()
    assert np.abs(u - M["nodes"]).max() <= 1e-9 * np.abs(M["nodes"]).max()
    # R(u*) = A u* - F vanishes, Dirichlet rows included (u_D - g_D)
    R = op.vector()
    scheme(op.vector(M["nodes"]), R)
    assert np.abs(R.download()).max() <= 1e-10 * np.abs(F.download()).max()
    # a right-hand side: u_D = g_D + rhs_D and R(u) = rhs
    rng = np.random.default_rng(4)
    rhs = op.vector(rng.normal(size=op.N))
    U2 = op.vector()
    info2 = scheme.solve(U2, rhs)
    assert info2["converged"]
    scheme(U2, R)
    R.axpy(-1.0, rhs)
    assert np.sqrt(R.normSquaredDofs()) <= 1e-9 * np.sqrt(rhs.normSquaredDofs() + F.normSquaredDofs())
