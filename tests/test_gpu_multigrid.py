"""The multigrid preconditioner (FEM_PRE_MULTIGRID) against its CPU
restatement (oracle/mg_oracle.py): one V-cycle M r on the same residual, in
FP64 to rounding and in FP32 to single precision; symmetry; and PCG with it
on the channel and log-normal workloads against the oracle's PCG."""
import numpy as np
import pytest

import dunefem_amd as fa
import mg_oracle as mgo
import pyoracle as ora
from dunefem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def dg(ctx, n, p, beta, kappa, prec=fa.F64):
    if isinstance(kappa, tuple):
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, beta, (0, 0, 0), 0.0, fa.KAPPA_CELL, precision=prec)
        op.set_kappa_cell(kappa[1])
    else:
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, beta, (0, 0, 0), 0.0, fa.KAPPA_CONST, kappa,
                                 precision=prec)
    return op


CASES = [
    # n, p, kappa kind: a single vertex level (dense), one and two h-levels,
    # odd vertex counts (non-uniform last coarse cell), every DG order
    ((3, 2, 4), 1, "cell"),
    ((4, 3, 5), 2, "cell"),
    ((5, 4, 3), 3, "const"),
    ((4, 4, 3), 4, "cell"),
    ((9, 7, 8), 2, "cell"),
    ((16, 12, 10), 1, "cell"),
    ((20, 20, 20), 1, "cell"),
]


@pytest.mark.parametrize("n,p,kind", CASES)
def test_multigrid_vcycle_matches_oracle(ctx, n, p, kind):
    rng = np.random.default_rng(hash((n, p)) % 2 ** 32)
    E = n[0] * n[1] * n[2]
    kappa = ("cell", 10.0 ** (2 * rng.integers(0, 2, size=E) + rng.uniform(-0.3, 0.3, size=E))) if kind == "cell" \
        else 1.7
    beta = 10.0 * p * p
    M = mgo.MGOracle(n, p, beta, kappa)
    op = dg(ctx, n, p, beta, kappa)
    pre = fa.Preconditioner(op, fa.PRE_MULTIGRID)
    r = rng.normal(size=op.N)
    zref = M.apply(r)
    R, Z = op.vector(r), op.vector()
    Z.fill(np.nan)
    pre(R, Z)
    z = Z.download()
    err = np.abs(z - zref).max() / np.abs(zref).max()
    print("multigrid", n, p, kind, "vertex levels", [L["A"].shape[0] for L in M.vx], "rel err", err, flush=True)
    assert err < 1e-12
    # FP32 preconditioner on an FP32 operator
    op32 = dg(ctx, n, p, beta, kappa, fa.F32)
    pre32 = fa.Preconditioner(op32, fa.PRE_MULTIGRID)
    R32, Z32 = op32.vector(r), op32.vector()
    pre32(R32, Z32)
    e32 = np.linalg.norm(Z32.download() - zref) / np.linalg.norm(zref)
    print("  fp32 rel l2 err", e32, flush=True)
    assert e32 < 1e-4


REACTION_CASES = [
    # the reaction in every level kind: pointwise cvol on a Q1 finest level
    # (probed into the vertex levels as it is), element means on coarse DG
    # levels, one and two vertex h-levels
    ((3, 2, 4), 1),
    ((4, 3, 5), 2),
    ((5, 4, 3), 3),
    ((4, 4, 3), 4),
    ((9, 7, 8), 2),
    ((20, 20, 20), 1),
]


@pytest.mark.parametrize("n,p", REACTION_CASES)
def test_multigrid_vcycle_with_reaction_matches_oracle(ctx, n, p):
    # c and a pointwise cvol (the form of a Jacobian S'(w)): the finest
    # smoother's blocks by the element's largest cvol, the coarse DG levels
    # with its element means, the bounded vertex Jacobi scaling
    rng = np.random.default_rng(7 * sum(n) + p)
    E = n[0] * n[1] * n[2]
    kc = 10.0 ** rng.uniform(-1, 1, size=E)
    cvol = 10.0 ** rng.uniform(0, 3, size=E * (p + 1) ** 3)
    beta, c = 10.0 * p * p, 0.3
    M = mgo.MGOracle(n, p, beta, ("cell", kc), c=c, cvol=cvol)
    zs = {}
    for prec in (fa.F64, fa.F32):
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, beta, (0, 0, 0), c, fa.KAPPA_CELL, precision=prec)
        op.set_kappa_cell(kc)
        op.set_reaction_qp(cvol)
        pre = fa.Preconditioner(op, fa.PRE_MULTIGRID)
        r = np.random.default_rng(3).normal(size=op.N)
        R, Z = op.vector(r), op.vector()
        Z.fill(np.nan)
        pre(R, Z)
        zs[prec] = Z.download()
    zref = M.apply(np.random.default_rng(3).normal(size=E * (p + 1) ** 3))
    err = np.abs(zs[fa.F64] - zref).max() / np.abs(zref).max()
    e32 = np.linalg.norm(zs[fa.F32] - zref) / np.linalg.norm(zref)
    print("multigrid with reaction", n, p, "vertex levels", [L["A"].shape[0] for L in M.vx], "rel err", err,
          "fp32 rel l2 err", e32, flush=True)
    assert err < 1e-12 and e32 < 1e-4


@pytest.mark.parametrize("pre", [fa.PRE_MULTIGRID, fa.PRE_BLOCK_JACOBI])
@pytest.mark.parametrize("prec", [fa.F64, fa.F32])
def test_rebinding_updates_the_solvers_preconditioner(ctx, pre, prec):
    # a solver that builds its preconditioner from
    # fem.solver.preconditioning.method updates it in place when the same
    # operator is bound again (the Newton steps do this): after new cvol and
    # kappa values the solve equals, bit for bit, that of a solver built
    # afresh; removing cvol changes the structure and rebuilds
    n, p = (9, 7, 8), 2
    rng = np.random.default_rng(17)
    E = n[0] * n[1] * n[2]
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 40.0, (0, 0, 0), 0.2, fa.KAPPA_CELL, precision=prec)
    op.set_kappa_cell(10.0 ** rng.uniform(-1, 1, size=E))
    op.set_reaction_qp(10.0 ** rng.uniform(0, 3, size=op.N))
    b = op.vector(rng.normal(size=op.N))
    kw = dict(method=fa.CG, tolerance=1e-6 if prec == fa.F32 else 1e-10, errormeasure=fa.RELATIVE,
              maxiterations=3000, preconditioning=pre)
    kept = fa.InverseOperator(ctx, **kw)
    kept.bind(op)
    X = op.vector()
    kept(b, X)
    changes = [("cvol", lambda: op.set_reaction_qp(10.0 ** rng.uniform(1, 4, size=op.N))),
               ("kappa", lambda: op.set_kappa_cell(10.0 ** rng.uniform(-2, 2, size=E))),
               ("cvol removed", lambda: op.set_reaction_qp(None))]
    for name, change in changes:
        change()
        kept.bind(op)
        X.clear()
        kept(b, X)
        fresh = fa.InverseOperator(ctx, **kw)
        fresh.bind(op)
        Y = op.vector()
        fresh(b, Y)
        ik, jf = kept.info(), fresh.info()
        print(name, "kept", ik["iterations"], "fresh", jf["iterations"], flush=True)
        assert ik["converged"] and ik["iterations"] == jf["iterations"]
        assert np.array_equal(X.download(), Y.download())


def test_multigrid_is_symmetric_positive(ctx):
    P = wl.problem3(n=12)
    op = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]))
    pre = fa.Preconditioner(op, fa.PRE_MULTIGRID)
    rng = np.random.default_rng(2)
    R1, R2, Z1, Z2 = op.vector(rng.normal(size=op.N)), op.vector(rng.normal(size=op.N)), op.vector(), op.vector()
    pre(R1, Z1)
    pre(R2, Z2)
    a, b = Z1.scalarProductDofs(R2), R1.scalarProductDofs(Z2)
    assert abs(a - b) <= 1e-12 * abs(a)
    assert Z1.scalarProductDofs(R1) > 0


def test_multigrid_pcg_channel_problem_matches_oracle(ctx):
    # the channel workload (kappa = 10^(4H)) at 12^3: GPU CG + multigrid against
    # the oracle's CG with the CPU V-cycle, same stopping rule
    P = wl.problem3(n=12)
    kappa = ("cell", P["kcell"])
    oop = ora.DGOperator(P["n"], P["p"], P["beta"], kappa=kappa)
    M = mgo.MGOracle(P["n"], P["p"], P["beta"], kappa)
    F = oop.rhs(P["fq"])
    xo, ito = ora.cg(oop, F, pre=ora.function_pre(M.apply, oop.N), tol=1e-8, maxit=500, criteria=1, nthreads=8)
    op = dg(ctx, P["n"], P["p"], P["beta"], kappa)
    inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, errormeasure=fa.RELATIVE, maxiterations=500)
    inv.bind(op, fa.Preconditioner(op, fa.PRE_MULTIGRID))
    X = op.vector()
    inv(op.vector(F), X)
    info = inv.info()
    print("12^3 channel CG+MG: gpu its", info["iterations"], "oracle its", ito, "true rel res",
          info["true_rel_residual"], flush=True)
    assert info["converged"] and abs(info["iterations"] - ito) <= 1
    assert np.abs(X.download() - xo).max() <= 1e-6 * np.abs(xo).max()


@pytest.mark.parametrize("prob", ["problem3", "problem4"])
def test_multigrid_solves_full_size_workloads(ctx, prob):
    # the configured sizes: 48^3 Q2 channels (2.99 M DoFs) and 96^3 Q4
    # log-normal (110.6 M DoFs); plain FP64 CG with the multigrid
    # preconditioner to a 1e-8 FP64 true residual
    P = getattr(wl, prob)()
    op = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]))
    F = op.rhs(P["fq"])
    del P["fq"]
    inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, errormeasure=fa.RELATIVE, maxiterations=1000,
                             residual_check=1)
    inv.bind(op, fa.Preconditioner(op, fa.PRE_MULTIGRID))
    X = op.vector()
    inv(F, X)
    info = inv.info()
    print(prob, "full size CG+MG:", info, flush=True)
    # residual_check: the solve also holds b - A x to the tolerance
    assert info["converged"] and info["true_rel_residual"] <= 1e-8


def test_cg_step_graph_replays_match_eager_solve(ctx):
    # one CG + multigrid step captured as a HIP graph and replayed for every
    # iteration (host synchronisation between chunks of replays): the same
    # kernels in the same order, so iterations and iterate equal the eager
    # solve bitwise; a second solve re-captures on the same solver
    P = wl.problem3(n=16)
    op = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]))
    pre = fa.Preconditioner(op, fa.PRE_MULTIGRID)
    F = op.rhs(P["fq"])
    res = {}
    for graph in (0, 1):
        inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-9, errormeasure=fa.RELATIVE, maxiterations=500,
                                 graph=graph)
        inv.bind(op, pre)
        out = []
        for rep in range(2):
            X = op.vector()
            inv(F, X)
            out.append((inv.info()["iterations"], X.download()))
        res[graph] = out
    for rep in range(2):
        assert res[0][rep][0] == res[1][rep][0] > 0
        assert np.array_equal(res[0][rep][1], res[1][rep][1])


def test_low_precision_multigrid_inside_refinement(ctx):
    # the log-normal Q4 workload at 24^3 (1.7 M DoFs): FP64 refinement around
    # FGMRES(30) with FP32 inner operator and a V-cycle on the FP32 and on
    # the BF16-MFMA operator. Both reach the FP64 1e-8 true residual. The
    # BF16 V-cycle forms its restricted residual in FP32; its smoothing
    # sweeps still add BF16 rounding noise, and it needs more iterations
    # (measured 127 against 51; at 96^3 it stalls, which the bound below
    # does not cover)
    P = wl.problem4(n=24)
    op64 = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]))
    op32 = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]), prec=fa.F32)
    F = op64.rhs(P["fq"])
    its = {}
    for name, prec in (("fp32", fa.F32), ("bf16", fa.BF16MFMA)):
        opm = op32 if prec == fa.F32 else dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]), prec=prec)
        inv = fa.InverseOperator(ctx, method=fa.FGMRES, tolerance=1e-8, errormeasure=fa.RELATIVE,
                                 maxiterations=600, restart=30, residual_check=1, refinement=1,
                                 inner_tolerance=1e-3, inner_maxiterations=200)
        inv.bind(op64, None, op32, fa.Preconditioner(opm, fa.PRE_MULTIGRID))
        X = op64.vector()
        inv(F, X)
        info = inv.info()
        print("Q4 24^3", name, "multigrid in refinement:", info, flush=True)
        assert info["converged"] and info["true_rel_residual"] <= 1e-8
        its[name] = info["iterations"]
    assert its["bf16"] <= 3 * its["fp32"], its


def test_fp32_multigrid_serves_fp64_krylov(ctx):
    # an FP32 V-cycle bound to the FP64 operator (rounded in, widened out):
    # FP64 CG and FGMRES keep their FP64 vectors and stopping test and reach
    # the FP64 1e-8 residual in about the FP64 V-cycle's iterations
    P = wl.problem3(n=16)
    op64 = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]))
    op32 = dg(ctx, P["n"], P["p"], P["beta"], ("cell", P["kcell"]), prec=fa.F32)
    F = op64.rhs(P["fq"])
    mg64, mg32 = fa.Preconditioner(op64, fa.PRE_MULTIGRID), fa.Preconditioner(op32, fa.PRE_MULTIGRID)
    its = {}
    for method in (fa.CG, fa.FGMRES):
        for name, pre in (("fp64", mg64), ("fp32", mg32)):
            inv = fa.InverseOperator(ctx, method=method, tolerance=1e-8, errormeasure=fa.RELATIVE,
                                     maxiterations=500, restart=30, residual_check=1)
            inv.bind(op64, pre)
            X = op64.vector()
            inv(F, X)
            info = inv.info()
            print("16^3 channels", method, name, "V-cycle:", info, flush=True)
            assert info["converged"] and info["true_rel_residual"] <= 1.01e-8
            its[method, name] = info["iterations"]
        assert its[method, "fp32"] <= its[method, "fp64"] + max(2, its[method, "fp64"] // 10), its
    # with refinement the outer preconditioner must match the outer operator
    inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, refinement=1)
    with pytest.raises(fa.FemError):
        inv.bind(op64, mg32, op32)
