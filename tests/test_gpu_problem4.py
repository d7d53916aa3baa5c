"""The combined workload: 3D SIPG DG Q4 diffusion with a log-normal cellwise
kappa (sigma 1.5, correlation length 0.1) and a 32-mode forcing.

At reduced size (8^3 hexes, 64,000 DoFs) the FP64, FP32 and BF16-MFMA
applies and the FP64 CG + block-Jacobi solve are compared with the oracle;
at the configured size (96^3, 110.6 M DoFs) the FP64 apply is checked by
linearity and symmetry, with the GPU alone."""
import numpy as np
import pytest

import dunefem_amd as fa
import pyoracle as ora
from dunefem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def dg(ctx, P, prec=fa.F64):
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, fa.KAPPA_CELL,
                             precision=prec)
    op.set_kappa_cell(P["kcell"])
    return op


def test_problem4_reduced_applies_match_oracle(ctx):
    P = wl.problem4(n=8)
    oop = ora.DGOperator(P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, kappa=("cell", P["kcell"]))
    u = np.random.default_rng(4).normal(size=oop.N)
    wref = oop.apply(u, nthreads=8)
    errs = {}
    for name, prec in (("fp64", fa.F64), ("fp32", fa.F32), ("bf16", fa.BF16MFMA)):
        op = dg(ctx, P, prec)
        U, W = op.vector(u), op.vector()
        op(U, W)
        w = W.download()
        errs[name] = (np.abs(w - wref).max() / np.abs(wref).max(), np.linalg.norm(w - wref) / np.linalg.norm(wref))
    print("problem4 8^3 Q4 apply errors (max-rel, l2-rel):", errs, flush=True)
    assert errs["fp64"][0] < 1e-12
    assert errs["fp32"][0] < 2e-5
    assert errs["bf16"][1] < 1e-2
    F = dg(ctx, P).rhs(P["fq"]).download()
    Fo = oop.rhs(P["fq"])
    assert np.abs(F - Fo).max() < 1e-12 * np.abs(Fo).max()


def test_problem4_reduced_cg_block_jacobi_matches_oracle(ctx):
    P = wl.problem4(n=8)
    oop = ora.DGOperator(P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, kappa=("cell", P["kcell"]))
    F = oop.rhs(P["fq"])
    inv_blocks = np.linalg.inv(oop.block_diag())
    xo, ito = ora.cg(oop, F, pre=ora.block_jacobi(inv_blocks), tol=1e-8, maxit=5000, criteria=1, nthreads=8)
    op = dg(ctx, P)
    pre = fa.Preconditioner(op, fa.PRE_BLOCK_JACOBI)
    inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, errormeasure=fa.RELATIVE, maxiterations=5000)
    inv.bind(op, pre)
    X = op.vector()
    inv(op.vector(F), X)
    info = inv.info()
    print("problem4 8^3 CG+BJ: gpu its", info["iterations"], "oracle its", ito, "true rel res",
          info["true_rel_residual"], flush=True)
    assert info["converged"] and abs(info["iterations"] - ito) <= 2
    x = X.download()
    assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()
    # the CG rule stops on sqrt(<r, M^-1 r>), so the true residual is judged
    # against the oracle's own at the same stopping rule
    r = F - oop.apply(x, nthreads=8)
    ro = F - oop.apply(xo, nthreads=8)
    assert np.linalg.norm(r) <= 1.5 * np.linalg.norm(ro) + 1e-12 * np.linalg.norm(F)


def test_problem4_full_size_apply_properties(ctx):
    # 96^3 Q4 (110.6 M DoFs, 0.88 GB per FP64 vector): linearity and symmetry
    P = wl.problem4(n=96)
    del P["fq"]
    op = dg(ctx, P)
    assert op.N == 110_592_000
    rng = np.random.default_rng(9)
    u = rng.normal(size=op.N)
    v = rng.normal(size=op.N)
    U, V = op.vector(u), op.vector(v)
    S = op.vector(2.0 * u - 3.0 * v)
    del u, v
    W1, W2, W3 = op.vector(), op.vector(), op.vector()
    op(U, W1)
    op(V, W2)
    op(S, W3)
    W3.axpy(-2.0, W1)
    W3.axpy(3.0, W2)
    assert np.sqrt(W3.normSquaredDofs()) < 1e-12 * np.sqrt(W1.normSquaredDofs())
    a, b = W1.scalarProductDofs(V), U.scalarProductDofs(W2)
    print("96^3 Q4 <Au,v> vs <u,Av>:", a, b, flush=True)
    assert abs(a - b) < 1e-11 * abs(a)
    # SPD: <A u, u> > 0
    assert W1.scalarProductDofs(U) > 0
