"""Batched Q2 solves (many uneven instances per launch) against the oracle."""
import numpy as np
import pytest

import dunefem_amd as fa
import pyoracle as ora
from dunefem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def small_batch(seed=0, count=7):
    rng = np.random.default_rng(seed)
    sides = [(int(rng.integers(1, 40)), int(rng.integers(1, 40))) for _ in range(count)]
    kv = [np.exp(rng.normal(0, 1, size=(s[0] * s[1], 9))) for s in sides]
    fq = [rng.normal(size=(s[0] * s[1], 9)) for s in sides]
    return sides, kv, fq


def test_batch_apply_and_rhs_match_oracle_per_instance(ctx):
    sides, kv, fq = small_batch()
    B = fa.Batch(ctx, sides, c=0.0)
    B.set_kappa_qp(np.concatenate(kv))
    rng = np.random.default_rng(1)
    u = rng.normal(size=B.N)
    U, W = B.vector(u), B.vector()
    B(U, W)
    w = W.download()
    F = B.rhs(np.concatenate(fq)).download()
    for i, s in enumerate(sides):
        lo, hi = B.offsets[i], B.offsets[i + 1]
        op = ora.CGOperator(s, c=0.0, kappa=kv[i])
        ref = op.apply(u[lo:hi])
        assert np.abs(w[lo:hi] - ref).max() <= 1e-12 * np.abs(ref).max(), i
        fr = op.rhs(fq[i])
        assert np.abs(F[lo:hi] - fr).max() <= 1e-12 * np.abs(fr).max(), i
    # FP32 batch apply
    U32, W32 = B.vector(u, fa.F32), B.vector(precision=fa.F32)
    B(U32, W32)
    assert np.abs(W32.download() - w).max() < 1e-4 * np.abs(w).max()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_mixed_solve_reaches_1e8_per_instance(ctx, use_graph):
    sides, kv, fq = small_batch(seed=3, count=9)
    B = fa.Batch(ctx, sides, c=0.0)
    B.set_kappa_qp(np.concatenate(kv))
    (tx, ty), ms = B.autotune()
    assert (tx, ty) in [(15, 15), (7, 7), (15, 7), (31, 7), (7, 31)] and all(m > 0 for m in ms)
    F = B.rhs(np.concatenate(fq))
    X = B.vector()
    info = B.solve(F, X, tol=1e-8, inner_tol=1e-3, use_graph=use_graph)
    print("batch info", {k: v for k, v in info.items() if k not in ("iterations", "rel_residual")}, flush=True)
    assert info["converged"] == len(sides)
    assert info["max_true_rel_residual"] <= 1e-8
    x, Fh = X.download(), F.download()
    for i, s in enumerate(sides):
        lo, hi = B.offsets[i], B.offsets[i + 1]
        op = ora.CGOperator(s, c=0.0, kappa=kv[i])
        r = Fh[lo:hi] - op.apply(x[lo:hi])
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(Fh[lo:hi]), i
        xo, _ = ora.cg(op, Fh[lo:hi], pre=ora.jacobi(1.0 / op.diag()), tol=1e-12, maxit=20000, criteria=1)
        assert np.abs(x[lo:hi] - xo).max() <= 1e-6 * np.abs(xo).max(), i


def test_problem2_generator_batch_subset(ctx):
    # 8 instances of the 256-instance workload (log-normal kappa, corr. length 0.1)
    inst = wl.problem2_instances(count=8, smin=16, smax=48)
    B = fa.Batch(ctx, [d["n"] for d in inst])
    B.set_kappa_qp(np.concatenate([d["kvol"] for d in inst]))
    F = B.rhs(np.concatenate([d["fq"] for d in inst]))
    X = B.vector()
    info = B.solve(F, X, tol=1e-8)
    assert info["converged"] == 8 and info["max_true_rel_residual"] <= 1e-8


def test_batch_graph_replays_after_resolve_and_new_coefficient(ctx):
    # the captured inner step holds buffer addresses: a second graph solve and
    # a solve after new coefficient data must both run on live, current data
    sides, kv, fq = small_batch(seed=5, count=6)
    B = fa.Batch(ctx, sides, c=0.0)
    F_host = np.concatenate(fq)

    def check(kvs, X, F):
        x, Fh = X.download(), F.download()
        for i, s in enumerate(sides):
            lo, hi = B.offsets[i], B.offsets[i + 1]
            op = ora.CGOperator(s, c=0.0, kappa=kvs[i])
            r = Fh[lo:hi] - op.apply(x[lo:hi])
            assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(Fh[lo:hi]), i

    B.set_kappa_qp(np.concatenate(kv))
    F = B.rhs(F_host)
    for rep in range(2):
        X = B.vector()
        info = B.solve(F, X, tol=1e-8, inner_tol=1e-3, use_graph=True)
        assert info["converged"] == len(sides), rep
        check(kv, X, F)
    rng = np.random.default_rng(11)
    kv2 = [np.exp(rng.normal(0, 1.5, size=k.shape)) for k in kv]
    B.set_kappa_qp(np.concatenate(kv2))
    X = B.vector()
    info = B.solve(F, X, tol=1e-8, inner_tol=1e-3, use_graph=True)
    assert info["converged"] == len(sides)
    check(kv2, X, F)


def test_problem2_full_config_256_instances(ctx):
    # the batched workload at its configured size: 256 instances, sides
    # U{32..256}, log-normal kappa (sigma 1, correlation length 0.1); the
    # FP64 true residual of a sample of instances is recomputed on the host
    # by the oracle's operator
    inst = wl.problem2_instances(count=256)
    B = fa.Batch(ctx, [d["n"] for d in inst], c=0.0)
    B.set_kappa_qp(np.concatenate([d["kvol"] for d in inst]))
    B.autotune()
    F = B.rhs(np.concatenate([d["fq"] for d in inst]))
    X = B.vector()
    info = B.solve(F, X, tol=1e-8)
    print("C2 full batch:", {k: v for k, v in info.items() if k not in ("iterations", "rel_residual")}, flush=True)
    assert info["converged"] == 256 and info["max_true_rel_residual"] <= 1e-8
    x, Fh = X.download(), F.download()
    sample = list(range(0, 256, 37)) + [int(np.argmax([d["n"][0] * d["n"][1] for d in inst]))]
    for i in sample:
        lo, hi = B.offsets[i], B.offsets[i + 1]
        op = ora.CGOperator(inst[i]["n"], c=0.0, kappa=inst[i]["kvol"])
        r = Fh[lo:hi] - op.apply(x[lo:hi])
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(Fh[lo:hi]), i
