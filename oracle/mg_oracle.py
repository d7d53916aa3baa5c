"""CPU restatement of the multigrid preconditioner (test infrastructure only:
imported by tests/ and never by the product path).

The algorithm is the one the library's multigrid implements: a symmetric
V-cycle over DG levels Q_p -> Q_{p/2} -> ... -> Q_1 (damped exact
element-block Jacobi), the conforming trilinear vertex level embedded in DG
Q1 (damped Jacobi) and geometric vertex levels with trilinear interpolation,
down to at most 512 vertices solved exactly.  The operators come from a
different route than on the device: the DG levels are the C oracle's own
operators (pyoracle.DGOperator), the DG Q1 operator is assembled as a sparse
matrix and every vertex operator is the sparse product P^T A P (the device
computes them by probing), and the coarsest solve is a dense LU solve.

No reference counterpart (dune-fem reaches multigrid only through assembled
external AMG, dune/fem/solver/parameter.hh:48-59); what pins it: the DG
Galerkin identity P^T A_p P = A_{p'} for the embedded Legendre modes, the
conforming Galerkin stencil against the closed-form trilinear stiffness
stencil, the symmetry of the V-cycle, and its convergence as a PCG
preconditioner (tests/test_oracle.py).

A reaction (the Jacobians of nonlinear operators): a constant c enters every
level as it is; a pointwise cvol at the volume quadrature points enters the
finest level as it is and the coarse DG levels by its Gauss-weighted element
means (then the vertex levels through P^T A P).  The finest level's smoother
blocks take cvol by its largest value in the element (an upper bound of the
true element block: the damped sweep stays convergent), and with any
reaction the vertex levels' Jacobi scaling is 1 / max(a_vv, omega/1.9
sum_w |a_vw|) (convergent by Gershgorin).  Pinned by: the bound block equals
the exact block for an element-constant cvol, the finest-level sweep and the
vertex sweeps stay contractions for a strongly varying cvol, and the V-cycle
stays symmetric positive definite there (tests/test_oracle.py)."""
import numpy as np
import scipy.sparse as sp

import pyoracle as ora

DENSE_MAX = 512
C1 = 0.5 / np.sqrt(3.0)   # Legendre coefficient of the 1D nodal functions


def assemble_dg(op, n, p):
    """Sparse matrix of an oracle DG operator by 27-colour probing of its
    apply (element colour (x mod 3, y mod 3, z mod 3))."""
    k = p + 1
    nb = k ** 3
    E = n[0] * n[1] * n[2]
    N = E * nb
    e = np.arange(E)
    ex, ey, ez = e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])
    col = (ex % 3) + 3 * ((ey % 3) + 3 * (ez % 3))
    offs = [(0, np.ones(E, bool))]
    for s, ok in ((-1, ex > 0), (1, ex < n[0] - 1), (-n[0], ey > 0), (n[0], ey < n[1] - 1),
                  (-n[0] * n[1], ez > 0), (n[0] * n[1], ez < n[2] - 1)):
        offs.append((s, ok))
    rows, cols, vals = [], [], []
    for c in range(27):
        sel = np.where(col == c)[0]
        if sel.size == 0:
            continue
        for j in range(nb):
            u = np.zeros(N)
            u[sel * nb + j] = 1.0
            w = op.apply(u, nthreads=8).reshape(E, nb)
            for s, ok in offs:
                src = sel[ok[sel]]
                dst = src + s
                blk = w[dst]                      # [len, nb]: column j of A[dst, src]
                r = (dst[:, None] * nb + np.arange(nb)[None, :]).ravel()
                rows.append(r)
                cols.append(np.repeat(src * nb + j, nb))
                vals.append(blk.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A.eliminate_zeros()
    return A


def conforming_P(n):
    """Trilinear vertex values -> DG Q1 Legendre modes (local index
    a 4 + b 2 + c, a the x mode), sparse [8E x (n+1)^3]."""
    T = np.array([[0.5, 0.5], [-C1, C1]])
    nvx, nvy = n[0] + 1, n[1] + 1
    E = n[0] * n[1] * n[2]
    e = np.arange(E)
    cx, cy, cz = e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])
    rows, cols, vals = [], [], []
    for a in range(2):
        for b in range(2):
            for c in range(2):
                for i in range(2):
                    for j in range(2):
                        for k in range(2):
                            rows.append(e * 8 + a * 4 + b * 2 + c)
                            cols.append((cx + i) + nvx * ((cy + j) + nvy * (cz + k)))
                            vals.append(np.full(E, T[a, i] * T[b, j] * T[c, k]))
    V = nvx * nvy * (n[2] + 1)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(8 * E, V))


def interp_1d(nf):
    """1D interpolation from the coarse vertices (fine positions 0, 2, 4, ...
    and the last fine vertex) to the nf fine vertices."""
    m = nf - 1
    pos = list(range(0, m + 1, 2))
    if pos[-1] != m:
        pos.append(m)
    nc = len(pos)
    R = np.zeros((nf, nc))
    for i in range(nf):
        J = int(np.searchsorted(pos, i))
        if pos[J] == i:
            R[i, J] = 1.0
        else:
            a, b = pos[J - 1], pos[J]
            R[i, J - 1] = (b - i) / (b - a)
            R[i, J] = (i - a) / (b - a)
    return sp.csr_matrix(R), nc


def embed_modes(E, kf, kc):
    idx = np.array([a * kf * kf + b * kf + c for a in range(kc) for b in range(kc) for c in range(kc)])
    rows = (np.arange(E)[:, None] * kf ** 3 + idx[None, :]).ravel()
    return sp.csr_matrix((np.ones(rows.size), (rows, np.arange(rows.size))), shape=(E * kf ** 3, E * kc ** 3))


def element_means(cvol, E, k):
    """Gauss-weighted means of a pointwise coefficient [E][k^3] (x fastest)
    over each element (the weights on [0, 1] sum to one)"""
    _, w = ora.gauss(k)
    W = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()   # [qz][qy][qx]
    return np.asarray(cvol, dtype=np.float64).reshape(E, k ** 3) @ W


class MGOracle:
    def __init__(self, n, p, beta, kappa, nu_dg=2, omega_dg=0.8, nu_v=1, omega_v=0.7, c=0.0, cvol=None):
        self.n, self.p = tuple(n), p
        self.nu_dg, self.omega_dg, self.nu_v, self.omega_v = nu_dg, omega_dg, nu_v, omega_v
        E = n[0] * n[1] * n[2]
        vol = 1.0 / E   # element volume on the unit cube
        mean = None if cvol is None else element_means(cvol, E, p + 1)
        self.dg = []
        q = p
        while True:
            op = ora.DGOperator(n, q, beta, (0, 0, 0), c, kappa=kappa)
            # the blocks of the operator without cvol, shifted by the
            # element's largest cvol (finest level) or its mean (coarse
            # levels, where cvol is that mean); the Legendre mass block is
            # vol * I
            B = op.block_diag()
            if cvol is not None:
                k3 = (q + 1) ** 3
                if not self.dg:
                    op.set_reaction_qp(cvol)
                    shift = np.asarray(cvol, dtype=np.float64).reshape(E, k3).max(axis=1)
                else:
                    op.set_reaction_qp(np.repeat(mean, k3))
                    shift = mean
                B = B + (shift * vol)[:, None, None] * np.eye(k3)[None]
            binv = np.linalg.inv(B)
            self.dg.append((op, q + 1, binv))
            if q == 1:
                break
            q = max(1, q // 2)
        self.E = E
        reaction = c != 0 or cvol is not None
        A1 = assemble_dg(self.dg[-1][0], n, 1)
        self.Pc = conforming_P(n)
        A = (self.Pc.T @ A1 @ self.Pc).tocsr()
        nv = [n[0] + 1, n[1] + 1, n[2] + 1]
        self.vx = []
        while True:
            d = A.diagonal()
            if reaction:
                d = np.maximum(d, omega_v / 1.9 * np.asarray(abs(A).sum(axis=1)).ravel())
            lev = {"A": A, "nv": list(nv), "dinv": 1.0 / d}
            self.vx.append(lev)
            if A.shape[0] <= DENSE_MAX:
                break
            Ps, ncs = zip(*(interp_1d(v) for v in nv))
            if list(ncs) == nv:
                break
            P = sp.kron(Ps[2], sp.kron(Ps[1], Ps[0])).tocsr()
            lev["P"] = P
            A = (P.T @ A @ P).tocsr()
            nv = list(ncs)
        self.coarse = np.linalg.inv(self.vx[-1]["A"].toarray())

    def _bj(self, l, v):
        _, k, binv = self.dg[l]
        nb = k ** 3
        return np.einsum("eij,ej->ei", binv, v.reshape(self.E, nb)).ravel()

    def _cycle_vertex(self, l, r):
        L = self.vx[l]
        if l + 1 == len(self.vx):
            return self.coarse @ r
        A, dinv, om = L["A"], L["dinv"], self.omega_v
        x = om * dinv * r
        for _ in range(1, self.nu_v):
            x = x + om * dinv * (r - A @ x)
        rc = L["P"].T @ (r - A @ x)
        x = x + L["P"] @ self._cycle_vertex(l + 1, rc)
        for _ in range(self.nu_v):
            x = x + om * dinv * (r - A @ x)
        return x

    def _cycle_dg(self, l, r):
        op, k, _ = self.dg[l]
        om = self.omega_dg
        x = om * self._bj(l, r)
        for _ in range(1, self.nu_dg):
            x = x + om * self._bj(l, r - op.apply(x, nthreads=8))
        t = r - op.apply(x, nthreads=8)
        if l + 1 < len(self.dg):
            kc = self.dg[l + 1][1]
            P = embed_modes(self.E, k, kc)
            x = x + P @ self._cycle_dg(l + 1, P.T @ t)
        else:
            x = x + self.Pc @ self._cycle_vertex(0, self.Pc.T @ t)
        for _ in range(self.nu_dg):
            x = x + om * self._bj(l, r - op.apply(x, nthreads=8))
        return x

    def apply(self, r):
        return self._cycle_dg(0, np.asarray(r, dtype=np.float64))
