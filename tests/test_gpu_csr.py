"""The assembled Jacobian (fem_op_assemble_csr, the counterpart of
DifferentiableGalerkinOperator::jacobian into SparseRowMatrix): against the
dense matrix of the oracle's operator (columns A e_j) on small meshes, and
at the configured sizes through the reference's own invariant
(dune/fempy/test/testoperator.py:40-73: the assembled Jacobian applied to u
equals the operator applied to u) and the diagonal against fem_op_diagonal."""
import numpy as np
import pytest

import dunefem_amd as fa
import pyoracle as ora
from dunefem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def dense_of(apply, N):
    A = np.zeros((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[:] = 0
        e[j] = 1
        A[:, j] = apply(e)
    return A


@pytest.mark.parametrize("n,p,kmode,b", [((2, 3, 2), 1, "qp", (1.0, -0.5, 0.25)), ((3, 2, 2), 2, "cell", (0, 0, 0)),
                                         ((2, 2, 3), 3, "qp", (0.3, 0.2, -0.7))])
def test_dg_csr_equals_dense_oracle_matrix(ctx, n, p, kmode, b):
    rng = np.random.default_rng(p)
    E, m = n[0] * n[1] * n[2], p + 1
    if kmode == "cell":
        kappa = ("cell", rng.uniform(0.1, 10.0, size=E))
    else:
        kvol = rng.uniform(0.2, 5.0, size=(E, m ** 3))
        kf = tuple(rng.uniform(0.2, 5.0, size=((n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2)), m * m))
                   for d in range(3))
        kappa = ("qp", kvol, kf)
    oop = ora.DGOperator(n, p, 10.0 * p * p, b, 0.3, kappa=kappa)
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, b, 0.3,
                             fa.KAPPA_CELL if kmode == "cell" else fa.KAPPA_QP)
    if kmode == "cell":
        op.set_kappa_cell(kappa[1])
    else:
        op.set_kappa_qp(kappa[1], kappa[2])
    A = fa.SparseRowMatrix(op)
    S = A.to_scipy()
    D = dense_of(oop.apply, oop.N)
    assert np.abs(S.toarray() - D).max() <= 1e-12 * np.abs(D).max()
    # pattern: element + face neighbours, columns increasing in each row
    nb = m ** 3
    for r in range(0, oop.N, nb):
        cols = S.indices[S.indptr[r]:S.indptr[r + 1]]
        assert np.all(np.diff(cols) > 0)
    assert S.nnz == A.nnz == sum(nb * nb * (1 + sum((c > 0) + (c < n[d] - 1) for d, c in enumerate(
        (e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1]))))) for e in range(E))


def test_q2_csr_equals_dense_oracle_matrix(ctx):
    n = (3, 4)
    kv = np.random.default_rng(2).uniform(0.3, 3.0, size=(12, 9))
    oop = ora.CGOperator(n, c=0.7, kappa=kv)
    op = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, n, 2, c=0.7, kappa_mode=fa.KAPPA_QP)
    op.set_kappa_qp(kv)
    S = fa.SparseRowMatrix(op).to_scipy()
    D = dense_of(oop.apply, oop.N)
    assert np.abs(S.toarray() - D).max() <= 1e-12 * np.abs(D).max()
    # Dirichlet rows are unit rows
    W = 2 * n[0] + 1
    for r in range(oop.N):
        I, J = r % W, r // W
        if I in (0, 2 * n[0]) or J in (0, 2 * n[1]):
            row = S.getrow(r).toarray().ravel()
            assert row[r] == 1.0 and np.count_nonzero(row) == 1


@pytest.mark.parametrize("prob", ["problem3", "problem1"])
def test_csr_apply_equals_operator_at_full_size(ctx, prob):
    # 48^3 Q2 (0.56 G nonzeros) and 64^3 Q3 with pointwise kappa and
    # advection (7.5 G nonzeros, 90 GB): A u (assembled) = op(u) to rounding
    if prob == "problem3":
        P = wl.problem3()
        del P["fq"]
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, fa.KAPPA_CELL)
        op.set_kappa_cell(P["kcell"])
        u = np.random.default_rng(3).normal(size=op.N)
    else:
        P = wl.problem1()
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], P["b"], 0.0, fa.KAPPA_QP)
        op.set_kappa_qp(P["kvol"], P["kfaces"])
        u = P["u"]
    A = fa.SparseRowMatrix(op)
    print(prob, "CSR rows", A.rows, "nnz", A.nnz, flush=True)
    U, W1, W2 = op.vector(u), op.vector(), op.vector()
    op(U, W1)
    A(U, W2)
    W2.axpy(-1.0, W1)
    err = np.sqrt(W2.normSquaredDofs() / W1.normSquaredDofs())
    print(prob, "||A_csr u - op(u)|| / ||op(u)||", err, flush=True)
    assert err < 1e-13
    d1 = A.diagonal().download()
    d2 = op.diagonal().download()
    assert np.abs(d1 - d2).max() <= 1e-12 * np.abs(d2).max()
