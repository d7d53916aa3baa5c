"""Pins for the CPU oracle (no GPU).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a transposed index) fails it: reference known answers
(quad_test.cc), closed forms (Gauss exactness, Legendre coefficients, penalty
energies), invariants (symmetry, upwind energy identity, consistency on an
exactly representable solution) and dense linear algebra for the solvers.
"""
import math

import numpy as np
import pytest

import pyoracle as ora


# ---- quadrature (dune/fem/quadrature/test/quad_test.cc:204-283) -------------
def test_gauss_exactness_and_weights():
    for m in range(1, 11):
        x, w = ora.gauss(m)
        assert np.all(np.diff(x) > 0)
        assert abs(w.sum() - 1.0) < 1e-14
        for deg in range(2 * m):
            assert abs(np.dot(w, x ** deg) - 1.0 / (deg + 1)) < 1e-14, (m, deg)
        # not exact beyond 2m-1 (a rule with too many points would pass the above)
        assert abs(np.dot(w, x ** (2 * m)) - 1.0 / (2 * m + 1)) > 1e-13


def test_gauss_matches_legendre_roots():
    for m in range(1, 9):
        x, w = ora.gauss(m)
        t, wt = np.polynomial.legendre.leggauss(m)
        assert np.allclose(x, (t + 1) / 2, atol=1e-15, rtol=0)
        assert np.allclose(w, wt / 2, atol=1e-15, rtol=0)


def test_quadrature_order_selection():
    # CubeQuadrature: smallest m with order 2m-1 >= requested (femquadratures_inline.hh:59-67)
    assert [ora.gauss_m_for_order(o) for o in range(0, 10)] == [1, 1, 2, 2, 3, 3, 4, 4, 5, 5]
    for p in range(1, 6):   # volume 2p and face 2p+1 both give p+1 points
        assert ora.gauss_m_for_order(2 * p) == p + 1 == ora.gauss_m_for_order(2 * p + 1)


@pytest.mark.parametrize("order", [1, 3, 5, 7])
def test_quad_test_known_answers(order):
    # quad_test.hh Func integrals: line -0.25, quad -0.2708333333, hex -0.03854166666
    m = ora.gauss_m_for_order(order)
    x, w = ora.gauss(m)
    f1 = lambda a: 0.25 - a
    f2 = lambda a, b: (0.25 - a) * (0.1 + 0.7 * a + 0.8 * b)
    f3 = lambda a, b, c: (0.25 - a) * (0.1 + 0.7 * a + 0.8 * b) * (a - 2.0 * b + 0.9 * c)
    r1 = sum(w[i] * f1(x[i]) for i in range(m))
    r2 = sum(w[i] * w[j] * f2(x[i], x[j]) for i in range(m) for j in range(m))
    r3 = sum(w[i] * w[j] * w[k] * f3(x[i], x[j], x[k]) for i in range(m) for j in range(m) for k in range(m))
    assert abs(r1 + 0.25) < 1e-12
    if order >= 2:
        assert abs(r2 + 0.2708333333) < 1e-10
    if order >= 3:
        assert abs(r3 + 0.03854166666) < 1e-10


# ---- shape functions ----------------------------------------------------------
def test_legendre_closed_form_and_orthonormality():
    # coefficients of P_n(2x-1): (-1)^(n+k) C(n,k) C(n+k,k) x^k (the reference's
    # LegendrePolynomials::factor table), scaled by sqrt(2n+1)
    xs = np.linspace(0, 1, 7)
    for n in range(0, 9):
        for x in xs:
            val = sum((-1) ** (n + k) * math.comb(n, k) * math.comb(n + k, k) * x ** k for k in range(n + 1))
            der = sum((-1) ** (n + k) * math.comb(n, k) * math.comb(n + k, k) * k * x ** (k - 1)
                      for k in range(1, n + 1))
            s = math.sqrt(2 * n + 1)
            assert abs(ora.legendre(n, x) - s * val) < 1e-11 * max(1, abs(s * val))
            assert abs(ora.legendre_d(n, x) - s * der) < 1e-10 * max(1, abs(s * der))
    x, w = ora.gauss(10)
    G = np.array([[sum(w[q] * ora.legendre(a, x[q]) * ora.legendre(b, x[q]) for q in range(10))
                   for b in range(8)] for a in range(8)])
    assert np.allclose(G, np.eye(8), atol=1e-13)


def test_lagrange_q2_properties():
    nodes = [0.0, 0.5, 1.0]
    for i in range(3):
        for j, xj in enumerate(nodes):
            assert abs(ora.lagrange2(i, xj) - (1.0 if i == j else 0.0)) < 1e-15
    for x in np.linspace(0, 1, 11):
        assert abs(sum(ora.lagrange2(i, x) for i in range(3)) - 1) < 1e-14
        assert abs(sum(ora.lagrange2_d(i, x) for i in range(3))) < 1e-13
        h = 1e-6
        for i in range(3):
            fd = (ora.lagrange2(i, x + h) - ora.lagrange2(i, x - h)) / (2 * h)
            assert abs(fd - ora.lagrange2_d(i, x)) < 1e-8


# ---- DG SIPG operator -----------------------------------------------------------
def assemble(op):
    N = op.N
    A = np.empty((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[:] = 0
        e[j] = 1
        A[:, j] = op.apply(e)
    return A


def random_qp_kappa(rng, n, p):
    m = p + 1
    E = n[0] * n[1] * n[2]
    kvol = rng.uniform(0.5, 2.0, size=(E, m ** 3))
    kf = []
    for d in range(3):
        nf = (n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2))
        kf.append(rng.uniform(0.5, 2.0, size=(nf, m * m)))
    return ("qp", kvol, tuple(kf))


@pytest.mark.parametrize("kind", ["const", "qp", "cell"])
def test_dg_symmetric_without_advection(kind):
    rng = np.random.default_rng(1)
    n, p = (2, 2, 3), 1
    kappa = {"const": 1.7, "qp": random_qp_kappa(rng, n, p),
             "cell": ("cell", rng.uniform(0.1, 10, size=12))}[kind]
    op = ora.DGOperator(n, p, beta=10 * p * p, kappa=kappa)
    A = assemble(op)
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    # coercive with beta = 10 p^2
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


def test_dg_advection_breaks_symmetry_by_upwind_part():
    rng = np.random.default_rng(2)
    n, p = (2, 2, 2), 1
    op = ora.DGOperator(n, p, beta=10, b=(1.0, 0.5, 0.25), kappa=1.0)
    op0 = ora.DGOperator(n, p, beta=10, kappa=1.0)
    A, A0 = assemble(op), assemble(op0)
    assert np.abs(A - A.T).max() > 1e-3
    # the diffusion part is unchanged by b
    opb = ora.DGOperator(n, p, beta=10, b=(1.0, 0.5, 0.25), kappa=0.0)
    assert np.abs(A - A0 - assemble(opb)).max() < 1e-12 * np.abs(A).max()


def _leg(n, x):
    c = np.zeros(n + 1)
    c[n] = 1
    return math.sqrt(2 * n + 1) * np.polynomial.legendre.legval(2 * x - 1, c)


def _trace(u_e, p, d, s, pts):
    """value of an element's modal polynomial on its face (d, s) at tangential points
    (independent numpy evaluation, not the oracle's tables)"""
    k = p + 1
    U = u_e.reshape(k, k, k)
    Pd = np.array([_leg(a, float(s)) for a in range(k)])
    F = np.tensordot(Pd, U, axes=([0], [d]))   # remaining dims in increasing order
    V = np.array([[_leg(a, x) for a in range(k)] for x in pts])  # (npts, k)
    return V @ F @ V.T   # [t1 point, t2 point]


def test_dg_upwind_energy_identity():
    # kappa = 0, div b = 0:  <A u, u> = 1/2 sum_F int |b.n| [u]^2 + 1/2 int_dOmega |b.n| u^2
    rng = np.random.default_rng(3)
    n, p = (2, 3, 2), 2
    b = (1.0, -0.5, 0.25)
    op = ora.DGOperator(n, p, beta=10, b=b, kappa=0.0)
    u = rng.normal(size=op.N)
    lhs = op.apply(u) @ u
    k = p + 1
    x, w = np.polynomial.legendre.leggauss(k + 2)
    x, w = (x + 1) / 2, w / 2
    W2 = np.outer(w, w)
    h = [1.0 / v for v in n]
    nx, ny, nz = n
    rhs = 0.0
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                e = ix + nx * (iy + ny * iz)
                c = (ix, iy, iz)
                ue = u[e * k ** 3:(e + 1) * k ** 3]
                for d in range(3):
                    t = [tt for tt in range(3) if tt != d]
                    area = h[t[0]] * h[t[1]]
                    for s in (0, 1):
                        nb = list(c)
                        nb[d] += 1 if s else -1
                        tin = _trace(ue, p, d, s, x)
                        if 0 <= nb[d] < n[d]:
                            if s == 0:
                                continue   # each interior face once
                            en = nb[0] + nx * (nb[1] + ny * nb[2])
                            tout = _trace(u[en * k ** 3:(en + 1) * k ** 3], p, d, 0, x)
                            rhs += 0.5 * abs(b[d]) * area * np.sum(W2 * (tin - tout) ** 2)
                        else:
                            rhs += 0.5 * abs(b[d]) * area * np.sum(W2 * tin ** 2)
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)


def test_dg_penalty_closed_forms():
    n, p = (3, 3, 3), 1
    k = p + 1
    op = ora.DGOperator(n, p, beta=10.0, kappa=1.0)
    one = np.zeros(op.N)
    one[::k ** 3] = 1.0   # the constant function 1 (phi_0 = 1)
    # a(1,1) = sum over boundary faces beta/h |F| = 6 beta / h on the unit cube
    assert abs(op.apply(one) @ one - 6 * 10.0 * 3) < 1e-10
    # indicator of the centre element: 6 faces, kappa beta/h |F| each
    ind = np.zeros(op.N)
    e = 1 + 3 * (1 + 3 * 1)
    ind[e * k ** 3] = 1.0
    assert abs(op.apply(ind) @ ind - 6 * 10.0 * (1 / 3)) < 1e-12
    # cellwise kappa: penalty uses avg(kappa) on interior faces
    kc = np.arange(1, 28, dtype=float)
    opc = ora.DGOperator(n, p, beta=10.0, kappa=("cell", kc))
    exp = sum(0.5 * (kc[e] + kc[e + s]) for s in (1, -1, 3, -3, 9, -9)) * 10.0 * (1 / 3)
    assert abs(opc.apply(ind) @ ind - exp) < 1e-10 * exp


def _dg_project_separable(funcs, n, p):
    """L2 projection of prod_d f_d(x_d) onto the orthonormal Legendre space"""
    k = p + 1
    x, w = np.polynomial.legendre.leggauss(k + 3)
    x, w = (x + 1) / 2, w / 2
    P = []
    for d in range(3):
        h = 1.0 / n[d]
        rows = []
        for i in range(n[d]):
            xs = (i + x) * h
            rows.append([np.sum(w * funcs[d](xs) * np.array([_leg(a, t) for t in x])) for a in range(k)])
        P.append(np.array(rows))
    nx, ny, nz = n
    U = np.einsum("xa,yb,zc->zyxabc", P[0], P[1], P[2])
    return U.reshape(-1)


def test_dg_consistency_on_representable_solution():
    # u* = X(x)Y(y)Z(z), X = x(1-x): in the Q2 space, zero on the boundary.
    # SIPG + upwind is consistent, so A U* = F exactly (all integrals exact).
    n, p = (2, 3, 2), 2
    kap, b = 1.3, (1.0, 0.5, -0.25)
    op = ora.DGOperator(n, p, beta=10 * p * p, b=b, kappa=kap)
    X = lambda t: t * (1 - t)
    dX = lambda t: 1 - 2 * t
    d2X = lambda t: -2 + 0 * t
    U = _dg_project_separable([X, X, X], n, p)
    m = p + 1
    xg, _ = ora.gauss(m)
    nx, ny, nz = n
    fq = np.empty((nx * ny * nz, m ** 3))
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                e = ix + nx * (iy + ny * iz)
                for q in range(m ** 3):
                    qx, qy, qz = q % m, (q // m) % m, q // (m * m)
                    xx, yy, zz = (ix + xg[qx]) / nx, (iy + xg[qy]) / ny, (iz + xg[qz]) / nz
                    lap = d2X(xx) * X(yy) * X(zz) + X(xx) * d2X(yy) * X(zz) + X(xx) * X(yy) * d2X(zz)
                    grad = (dX(xx) * X(yy) * X(zz), X(xx) * dX(yy) * X(zz), X(xx) * X(yy) * dX(zz))
                    fq[e, q] = -kap * lap + sum(b[i] * grad[i] for i in range(3))
    F = op.rhs(fq)
    R = op.apply(U) - F
    assert np.abs(R).max() < 1e-13 * np.abs(F).max()


def test_dg_threaded_equals_serial():
    rng = np.random.default_rng(4)
    n, p = (4, 3, 5), 2
    op = ora.DGOperator(n, p, beta=40, b=(1, .5, .25), kappa=random_qp_kappa(rng, n, p))
    u = rng.normal(size=op.N)
    w1 = op.apply(u, nthreads=1)
    w4 = op.apply(u, nthreads=4)
    assert np.abs(w1 - w4).max() <= 1e-14 * np.abs(w1).max()


# ---- CG Q2 operator --------------------------------------------------------------
def test_cg_consistency_and_dirichlet_rows():
    # u* = x(1-x) y(1-y) is Q2-representable: A U* = F for -Lap u + u = f
    n = (3, 4)
    op = ora.CGOperator(n, c=1.0, kappa=1.0)
    NX, NY = 2 * n[0] + 1, 2 * n[1] + 1
    xs, ys = np.arange(NX) / (NX - 1), np.arange(NY) / (NY - 1)
    U = np.outer(ys * (1 - ys), xs * (1 - xs)).ravel()
    xg, _ = ora.gauss(3)
    fq = np.empty((n[0] * n[1], 9))
    for ey in range(n[1]):
        for ex in range(n[0]):
            for q in range(9):
                x = (ex + xg[q % 3]) / n[0]
                y = (ey + xg[q // 3]) / n[1]
                fq[ex + n[0] * ey, q] = 2 * y * (1 - y) + 2 * x * (1 - x) + x * (1 - x) * y * (1 - y)
    F = op.rhs(fq)
    assert np.abs(op.apply(U) - F).max() < 1e-14
    # Dirichlet rows act as identity
    v = np.random.default_rng(5).normal(size=op.N)
    w = op.apply(v)
    bnd = np.zeros((NY, NX), bool)
    bnd[0, :] = bnd[-1, :] = bnd[:, 0] = bnd[:, -1] = True
    assert np.array_equal(w[bnd.ravel()], v[bnd.ravel()])


def test_cg_symmetric_interior_and_diag():
    rng = np.random.default_rng(6)
    n = (3, 3)
    kv = rng.uniform(0.5, 2, size=(9, 9))
    op = ora.CGOperator(n, c=0.5, kappa=kv)
    A = assemble(op)
    NX = 7
    bnd = np.zeros((NX, NX), bool)
    bnd[0, :] = bnd[-1, :] = bnd[:, 0] = bnd[:, -1] = True
    I = ~bnd.ravel()
    Ai = A[np.ix_(I, I)]
    assert np.abs(Ai - Ai.T).max() < 1e-14 * np.abs(Ai).max()
    assert np.linalg.eigvalsh(Ai).min() > 0
    assert np.allclose(op.diag(), np.diag(A), rtol=1e-13, atol=0)


# ---- Krylov solvers --------------------------------------------------------------
def test_cg_matches_dense_solve_and_reference_semantics():
    n = (6, 5)
    op = ora.CGOperator(n, c=1.0, kappa=1.0)
    rng = np.random.default_rng(7)
    fq = rng.normal(size=(30, 9))
    F = op.rhs(fq)
    A = assemble(op)
    xref = np.linalg.solve(A, F)
    dinv = 1.0 / op.diag()
    for pre in (None, ora.jacobi(dinv)):
        x, it = ora.cg(op, F, pre=pre, tol=1e-12, maxit=500, criteria=1)
        assert it > 0
        assert np.abs(x - xref).max() < 1e-9 * np.abs(xref).max()
    # maxiterations reached -> negative count (cg.hh:116)
    x, it = ora.cg(op, F, pre=ora.jacobi(dinv), tol=1e-14, maxit=3, criteria=1)
    assert it == -3


@pytest.mark.parametrize("kind", ["gmres", "fgmres"])
def test_gmres_matches_dense_solve(kind):
    rng = np.random.default_rng(8)
    n, p = (2, 2, 2), 1
    op = ora.DGOperator(n, p, beta=10, b=(1, .5, .25), kappa=random_qp_kappa(rng, n, p))
    A = assemble(op)
    b = rng.normal(size=op.N)
    xref = np.linalg.solve(A, b)
    blocks = op.block_diag()
    inv = np.linalg.inv(blocks)
    f = ora.gmres if kind == "gmres" else ora.fgmres
    for pre in (None, ora.block_jacobi(inv)):
        x, it = f(op, b, pre=pre, tol=1e-12, maxit=500, criteria=1, restart=10)
        assert it > 0
        assert np.linalg.norm(A @ x - b) <= 1.01e-12 * np.linalg.norm(b)
        assert np.abs(x - xref).max() < 1e-8 * np.abs(xref).max()


def test_cg_energy_minimisation_property():
    # the preconditioned CG iterate after k steps from x0 = 0 minimises the
    # energy error ||x - x*||_A over K_k(M A, M b) = M K_k(A M, b); checked
    # against the brute-force Galerkin solution on an explicit Krylov basis
    # (SPD SIPG operator, with and without block Jacobi)
    rng = np.random.default_rng(12)
    n, p = (2, 2, 2), 1
    op = ora.DGOperator(n, p, beta=10, kappa=random_qp_kappa(rng, n, p))
    A = assemble(op)
    b = rng.normal(size=op.N)
    xs = np.linalg.solve(A, b)
    inv = np.linalg.inv(op.block_diag())
    E, nb, _ = inv.shape
    M = np.zeros((op.N, op.N))
    for e in range(E):
        M[e * nb:(e + 1) * nb, e * nb:(e + 1) * nb] = inv[e]
    for pre, Mm in ((None, np.eye(op.N)), (ora.block_jacobi(inv), M)):
        AM = A @ Mm
        for k in range(1, 9):
            x, it = ora.cg(op, b, pre=pre, tol=1e-300, maxit=k, criteria=1)
            assert it == -k
            K = np.empty((op.N, k))
            v = b / np.linalg.norm(b)
            for j in range(k):
                K[:, j] = v
                v = AM @ v
                v -= K[:, :j + 1] @ (K[:, :j + 1].T @ v)
                v -= K[:, :j + 1] @ (K[:, :j + 1].T @ v)
                v /= np.linalg.norm(v)
            Z = Mm @ K
            c = np.linalg.solve(Z.T @ A @ Z, Z.T @ b)
            xg = Z @ c
            ek, eg = x - xs, xg - xs
            assert abs(np.sqrt(ek @ A @ ek) - np.sqrt(eg @ A @ eg)) <= 1e-9 * np.sqrt(xs @ A @ xs), (k,)
            assert np.linalg.norm(x - xg) <= 1e-7 * np.linalg.norm(xg)


@pytest.mark.parametrize("kind", ["gmres", "fgmres"])
def test_gmres_minimal_residual_property(kind):
    # the right-preconditioned (F)GMRES iterate after k < restart steps from
    # x0 = 0 minimises ||b - A x|| over x in M K_k(A M, b); with a fixed linear
    # M both recurrences must reach the brute-force least-squares minimum over
    # an explicit Krylov basis (nonsymmetric operator: advection on)
    rng = np.random.default_rng(11)
    n, p = (2, 2, 2), 1
    op = ora.DGOperator(n, p, beta=10, b=(1, .5, .25), kappa=random_qp_kappa(rng, n, p))
    A = assemble(op)
    b = rng.normal(size=op.N)
    inv = np.linalg.inv(op.block_diag())
    E, nb, _ = inv.shape
    M = np.zeros((op.N, op.N))
    for e in range(E):
        M[e * nb:(e + 1) * nb, e * nb:(e + 1) * nb] = inv[e]
    f = ora.gmres if kind == "gmres" else ora.fgmres
    for pre, Mm in ((None, np.eye(op.N)), (ora.block_jacobi(inv), M)):
        AM = A @ Mm
        for k in range(1, 9):
            x, it = f(op, b, pre=pre, tol=1e-300, maxit=k, criteria=1, restart=20)
            assert it == -k
            K = np.empty((op.N, k))
            v = b / np.linalg.norm(b)
            for j in range(k):
                K[:, j] = v
                v = AM @ v
                v -= K[:, :j + 1] @ (K[:, :j + 1].T @ v)
                v -= K[:, :j + 1] @ (K[:, :j + 1].T @ v)
                v /= np.linalg.norm(v)
            c, *_ = np.linalg.lstsq(AM @ K, b, rcond=None)
            rmin = np.linalg.norm(b - AM @ (K @ c))
            rk = np.linalg.norm(b - A @ x)
            assert abs(rk - rmin) <= 1e-10 * np.linalg.norm(b), (kind, k, rk, rmin)
            assert np.linalg.norm(x - Mm @ (K @ c)) <= 1e-8 * np.linalg.norm(x)


def test_block_diag_matches_assembled_blocks():
    rng = np.random.default_rng(9)
    n, p = (2, 3, 2), 1
    op = ora.DGOperator(n, p, beta=10, kappa=("cell", rng.uniform(0.1, 10, size=12)))
    A = assemble(op)
    B = op.block_diag()
    nb = (p + 1) ** 3
    for e in range(12):
        assert np.allclose(B[e], A[e * nb:(e + 1) * nb, e * nb:(e + 1) * nb], rtol=0, atol=1e-14 * np.abs(A).max())


# ---- multigrid oracle (oracle/mg_oracle.py) -------------------------------------------
def test_mg_dg_levels_are_galerkin_products():
    # the embedded low Legendre modes: P^T A_p P equals the Q_p' operator with
    # the same penalty factor (exact quadrature of the Q_p' products)
    import mg_oracle as mgo
    rng = np.random.default_rng(30)
    n = (3, 2, 3)
    kc = rng.uniform(0.1, 10.0, size=18)
    for pf, pc in ((2, 1), (4, 2), (3, 1)):
        beta = 10.0 * pf * pf
        Af = mgo.assemble_dg(ora.DGOperator(n, pf, beta, kappa=("cell", kc)), n, pf)
        Ac = mgo.assemble_dg(ora.DGOperator(n, pc, beta, kappa=("cell", kc)), n, pc)
        P = mgo.embed_modes(18, pf + 1, pc + 1)
        G = (P.T @ Af @ P).toarray()
        assert np.abs(G - Ac.toarray()).max() <= 1e-13 * np.abs(G).max()


def test_mg_conforming_stencil_is_the_trilinear_stiffness():
    # kappa = 1: at an interior vertex the Galerkin operator of the conforming
    # trilinear space is the Q1 stiffness stencil h (8/3 centre, 0 faces,
    # -1/6 edges, -1/12 corners); the SIPG face terms vanish on continuous functions
    import mg_oracle as mgo
    n = (6, 6, 6)
    A1 = mgo.assemble_dg(ora.DGOperator(n, 1, 10.0, kappa=1.0), n, 1)
    Pc = mgo.conforming_P(n)
    A = (Pc.T @ A1 @ Pc).toarray()
    h = 1.0 / 6
    v = 3 + 7 * (3 + 7 * 3)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                w = v + dx + 7 * (dy + 7 * dz)
                nz = abs(dx) + abs(dy) + abs(dz)
                ref = h * {0: 8.0 / 3.0, 1: 0.0, 2: -1.0 / 6.0, 3: -1.0 / 12.0}[nz]
                assert abs(A[v, w] - ref) < 1e-13, (dx, dy, dz, A[v, w], ref)
    assert abs(A[v].sum()) < 1e-13   # no entries beyond distance 1
    assert np.allclose(A, A.T, atol=1e-13)


def test_mg_vcycle_is_symmetric_and_preconditions_cg():
    import mg_oracle as mgo
    from dunefem_amd import workloads as wl
    P = wl.problem3(n=8)
    kappa = ("cell", P["kcell"])
    M = mgo.MGOracle(P["n"], P["p"], P["beta"], kappa)
    assert len(M.vx) == 2 and M.vx[1]["A"].shape[0] == 125   # 9^3 vertices, then 5^3
    rng = np.random.default_rng(31)
    op = ora.DGOperator(P["n"], P["p"], P["beta"], kappa=kappa)
    r1, r2 = rng.normal(size=op.N), rng.normal(size=op.N)
    a, b = M.apply(r1) @ r2, r1 @ M.apply(r2)
    assert abs(a - b) <= 1e-12 * abs(a)
    assert M.apply(r1) @ r1 > 0
    F = op.rhs(P["fq"])
    pre = ora.function_pre(M.apply, op.N)
    _, it_mg = ora.cg(op, F, pre=pre, tol=1e-8, maxit=500, criteria=1)
    _, it_bj = ora.cg(op, F, pre=ora.block_jacobi(np.linalg.inv(op.block_diag())), tol=1e-8, maxit=5000,
                      criteria=1)
    print("8^3 channel problem: PCG iterations multigrid", it_mg, "block Jacobi", it_bj)
    assert 0 < it_mg < 40 and it_bj > 5 * it_mg


def _reaction_case(n, p, lo, hi, seed):
    rng = np.random.default_rng(seed)
    E = n[0] * n[1] * n[2]
    kc = 10.0 ** rng.uniform(-1, 1, size=E)
    cvol = 10.0 ** rng.uniform(lo, hi, size=E * (p + 1) ** 3)
    return rng, E, ("cell", kc), cvol


def test_mg_reaction_bound_block_is_exact_for_element_constant_cvol():
    # with cvol constant in each element the largest value is the mean and
    # the shifted diffusion block is the operator's exact element block (the
    # Legendre mass block is vol * I); element_means returns the constants
    import mg_oracle as mgo
    n, p = (3, 2, 2), 2
    rng, E, kappa, _ = _reaction_case(n, p, 0, 3, 41)
    cm = 10.0 ** rng.uniform(0, 3, size=E)
    cvol = np.repeat(cm, (p + 1) ** 3)
    assert np.abs(mgo.element_means(cvol, E, p + 1) - cm).max() < 1e-12 * cm.max()
    M = mgo.MGOracle(n, p, 10.0 * p * p, kappa, c=0.4, cvol=cvol)
    op = ora.DGOperator(n, p, 10.0 * p * p, (0, 0, 0), 0.4, kappa=kappa)
    op.set_reaction_qp(cvol)
    exact = op.block_diag()
    assert np.abs(np.linalg.inv(M.dg[0][2]) - exact).max() < 1e-11 * np.abs(exact).max()


def _sweep_radius(A, B):
    """largest eigenvalue of B^-1 A (A, B symmetric positive definite)"""
    import scipy.linalg as sl
    return sl.eigh(A, B, eigvals_only=True).max()


def test_mg_reaction_smoothers_stay_convergent():
    # a strongly varying cvol (1 .. 10^5 inside elements): with the element
    # blocks shifted by the largest cvol the damped block-Jacobi sweep is a
    # contraction (omega lambda_max(B^-1 A) < 2); shifted by the mean it
    # diverges.  On a reaction-dominated vertex level plain damped Jacobi
    # diverges and the Gershgorin-bounded scaling does not.
    import mg_oracle as mgo
    import scipy.linalg as sl
    n, p = (3, 2, 2), 2
    _, E, kappa, cvol = _reaction_case(n, p, 0, 5, 42)
    beta = 10.0 * p * p
    M = mgo.MGOracle(n, p, beta, kappa, cvol=cvol)
    op = ora.DGOperator(n, p, beta, kappa=kappa)
    op.set_reaction_qp(cvol)
    A = mgo.assemble_dg(op, n, p).toarray()
    Bbound = sl.block_diag(*np.linalg.inv(M.dg[0][2]))
    plain = ora.DGOperator(n, p, beta, kappa=kappa).block_diag()
    k3 = (p + 1) ** 3
    shift = mgo.element_means(cvol, E, p + 1) / E
    Bmean = sl.block_diag(*(plain + shift[:, None, None] * np.eye(k3)[None]))
    rb, rm = _sweep_radius(A, Bbound), _sweep_radius(A, Bmean)
    print("omega lambda_max: bound blocks", 0.8 * rb, "mean blocks", 0.8 * rm)
    assert 0.8 * rb < 2.0 < 0.8 * rm
    # vertex level of a reaction-dominated Q1 problem
    n1 = (8, 8, 8)
    _, E1, kappa1, cv1 = _reaction_case(n1, 1, 4, 6, 43)
    M1 = mgo.MGOracle(n1, 1, 10.0, kappa1, cvol=cv1)
    Av = M1.vx[0]["A"].toarray()
    lam_safe = sl.eigh(Av, np.diag(1.0 / M1.vx[0]["dinv"]), eigvals_only=True).max()
    lam_plain = sl.eigh(Av, np.diag(np.diag(Av)), eigvals_only=True).max()
    print("vertex level omega lambda_max: bounded", 0.7 * lam_safe, "plain diagonal", 0.7 * lam_plain)
    assert 0.7 * lam_safe < 2.0 < 0.7 * lam_plain


def test_mg_vcycle_with_reaction_is_spd_and_preconditions_cg():
    # symmetric positive definite even for cvol spanning 10^5 inside
    # elements (the dense V-cycle matrix on a small mesh); on 8^3 Q2 with
    # cvol spanning 10^4 PCG with it needs far fewer iterations than with
    # the exact element blocks.  (Where cvol varies by 10^5 inside the
    # elements of a tiny mesh the bound blocks over-damp and exact block
    # Jacobi wins; the V-cycle stays a valid preconditioner there.)
    import mg_oracle as mgo
    n, p = (3, 2, 2), 2
    _, E, kappa, cvol = _reaction_case(n, p, 0, 5, 44)
    beta = 10.0 * p * p
    M = mgo.MGOracle(n, p, beta, kappa, c=0.3, cvol=cvol)
    N = E * (p + 1) ** 3
    Md = np.stack([M.apply(np.eye(N)[i]) for i in range(N)], axis=1)
    assert np.abs(Md - Md.T).max() <= 1e-10 * np.abs(Md).max()
    assert np.linalg.eigvalsh(0.5 * (Md + Md.T)).min() > 0
    n = (8, 8, 8)
    rng, E, kappa, cvol = _reaction_case(n, p, 0, 4, 45)
    M = mgo.MGOracle(n, p, beta, kappa, c=0.3, cvol=cvol)
    op = ora.DGOperator(n, p, beta, (0, 0, 0), 0.3, kappa=kappa)
    op.set_reaction_qp(cvol)
    F = rng.normal(size=op.N)
    _, it_mg = ora.cg(op, F, pre=ora.function_pre(M.apply, op.N), tol=1e-10, maxit=500, criteria=1)
    _, it_bj = ora.cg(op, F, pre=ora.block_jacobi(np.linalg.inv(op.block_diag())), tol=1e-10, maxit=5000,
                      criteria=1)
    print("8^3 Q2 with reaction: PCG iterations V-cycle", it_mg, "exact block Jacobi", it_bj)
    assert 0 < it_mg and 3 * it_mg < it_bj


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_sipg_consistency_with_pointwise_kappa_and_boundary_data(p):
    # u* a polynomial of degree <= p per direction, kappa = 1 + x + 2y + 3z at
    # the volume and face quadrature points, advection and reaction, u = u*
    # on the boundary (Nitsche + inflow data): A u*_h = F(f) + F_bnd(g) to
    # rounding.  A transposed face-point index of kappa or g, a sign of the
    # Nitsche data terms or a wrong inflow side breaks the identity.
    from manufactured import sipg_problem
    n, b, c = (3, 2, 4), (0.7, -0.4, 0.3), 0.5
    M = sipg_problem(n, p, b, c, seed=p)
    op = ora.DGOperator(n, p, 10.0 * p * p, b, c, kappa=("qp", M["kvol"], tuple(M["kfaces"])))
    lhs = op.apply(M["u"])
    rhs = op.rhs_bnd(M["gfaces"], op.rhs(M["fq"]))
    assert np.abs(lhs - rhs).max() <= 1e-11 * np.abs(rhs).max()
    # the identity needs the boundary data: without it, it fails
    assert np.abs(lhs - op.rhs(M["fq"])).max() > 1e-3 * np.abs(rhs).max()


def test_bicgstab_oracle_reaches_the_dense_solution():
    # nonsymmetric SIPG with advection, with and without block Jacobi: the
    # BiCGStab restatement converges to the dense solution; its residual
    # estimate matches the true residual at the end (no preconditioner)
    rng = np.random.default_rng(40)
    n, p = (3, 3, 2), 2
    op = ora.DGOperator(n, p, 40.0, (1.0, -0.5, 0.25), 0.0, kappa=("cell", rng.uniform(0.5, 2.0, size=18)))
    A = assemble(op)
    b = rng.normal(size=op.N)
    xd = np.linalg.solve(A, b)
    x, it = ora.bicgstab(op, b, tol=1e-10, maxit=2000, criteria=1)
    assert it > 0 and np.abs(x - xd).max() <= 1e-7 * np.abs(xd).max()
    assert np.linalg.norm(b - A @ x) <= 2e-10 * np.linalg.norm(b)
    inv = np.linalg.inv(op.block_diag())
    x2, it2 = ora.bicgstab(op, b, pre=ora.block_jacobi(inv), tol=1e-10, maxit=2000, criteria=1)
    assert 0 < it2 < it and np.abs(x2 - xd).max() <= 1e-6 * np.abs(xd).max()
    # non-convergence: negative count
    _, it3 = ora.bicgstab(op, b, tol=1e-14, maxit=3, criteria=1)
    assert it3 == -3


@pytest.mark.parametrize("p,kmode,b,c", [(1, "qp", (1.0, 0.5, 0.25), 0.0), (2, "cell", (0, 0, 0), 0.3),
                                         (3, "qp", (1.0, -0.5, 0.25), 0.0), (4, "cell", (0.3, 0.2, -1.0), 1.0),
                                         (3, "const", (0, 0, 0), 0.0)])
def test_sum_factorised_cpu_apply_equals_dense_oracle(p, kmode, b, c):
    # the sum-factorised CPU apply (timing baseline) is the same operator as
    # the dense restatement of the reference algorithm
    rng = np.random.default_rng(50 + p)
    n = (4, 3, 5)
    E, m = 60, p + 1
    if kmode == "qp":
        kvol = rng.uniform(0.2, 5.0, size=(E, m ** 3))
        kf = tuple(rng.uniform(0.2, 5.0, size=((n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2)), m * m))
                   for d in range(3))
        kappa = ("qp", kvol, kf)
    elif kmode == "cell":
        kappa = ("cell", rng.uniform(0.1, 10.0, size=E))
    else:
        kappa = 1.7
    op = ora.DGOperator(n, p, 10.0 * p * p, b, c, kappa=kappa)
    u = rng.normal(size=op.N)
    w = op.apply(u)
    for nt in (1, 4):
        ws = op.apply_sf(u, nthreads=nt)
        assert np.abs(ws - w).max() <= 1e-12 * np.abs(w).max()
