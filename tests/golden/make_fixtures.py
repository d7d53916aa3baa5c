"""Writes the small golden fixtures under tests/golden/ from the CPU oracle.

Run from the repository root after `make -C oracle`:
    python tests/golden/make_fixtures.py

The reference holds no stored DoF vectors (its tests check invariants), so
these fixtures are the oracle's own outputs on tiny seeded inputs; they pin the
oracle against drift and give the GPU tests a fixed target.  The quadrature
known answers are the reference's (quadrature/test/quad_test.cc:246-278 with
the Func of quad_test.hh)."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as ora  # noqa: E402


def dg_faces_shape(n, m):
    return [((n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2))) * m * m for d in range(3)]


def main():
    rng = np.random.default_rng(20261202)
    json.dump({"source": "dune/fem/quadrature/test/quad_test.cc:246-278, quad_test.hh",
               "line": -0.25, "quad": -0.2708333333, "hex": -0.03854166666},
              open(os.path.join(HERE, "quad_test_kats.json"), "w"), indent=1)
    # DG SIPG Q2 on 2x2x2, cellwise kappa, advection and reaction
    n, p = (2, 2, 2), 2
    kc = rng.uniform(0.1, 10.0, size=8)
    op = ora.DGOperator(n, p, 10.0 * p * p, (1.0, 0.5, 0.25), 0.3, kappa=("cell", kc))
    u = rng.normal(size=op.N)
    fq = rng.normal(size=8 * 27)
    np.savez(os.path.join(HERE, "dg_sipg_q2_2x2x2_cell.npz"), n=n, p=p, beta=40.0, b=(1.0, 0.5, 0.25), c=0.3,
             kcell=kc, u=u, w=op.apply(u), fq=fq, F=op.rhs(fq),
             diag=np.einsum("eii->ei", op.block_diag()).ravel())
    # DG SIPG Q3 on 3x2x2, pointwise kappa
    n, p = (3, 2, 2), 3
    m = p + 1
    E = 12
    kv = rng.uniform(0.1, 10.0, size=E * m ** 3)
    kf = [rng.uniform(0.1, 10.0, size=s) for s in dg_faces_shape(n, m)]
    op = ora.DGOperator(n, p, 10.0 * p * p, (1.0, 0.5, 0.25), 0.0, kappa=("qp", kv, tuple(kf)))
    u = rng.normal(size=op.N)
    np.savez(os.path.join(HERE, "dg_sipg_q3_3x2x2_qp.npz"), n=n, p=p, beta=90.0, b=(1.0, 0.5, 0.25), c=0.0,
             kvol=kv, kx=kf[0], ky=kf[1], kz=kf[2], u=u, w=op.apply(u))
    # continuous Q2 on 4x4, pointwise kappa, reaction 1
    kv = rng.uniform(0.1, 10.0, size=16 * 9)
    cop = ora.CGOperator((4, 4), c=1.0, kappa=kv)
    u = rng.normal(size=cop.N)
    fq = rng.normal(size=16 * 9)
    np.savez(os.path.join(HERE, "cg_q2_4x4_qp.npz"), n=(4, 4), c=1.0, kvol=kv, u=u, w=cop.apply(u), fq=fq,
             F=cop.rhs(fq), diag=cop.diag())
    # Jacobi CG on continuous Q2 8x8 (kappa 1, c 1): the reference recurrence
    cop = ora.CGOperator((8, 8), c=1.0)
    fq = rng.normal(size=64 * 9)
    F = cop.rhs(fq)
    x, its = ora.cg(cop, F, pre=ora.jacobi(1.0 / cop.diag()), tol=1e-10, maxit=5000, criteria=1)
    np.savez(os.path.join(HERE, "cg_q2_8x8_jacobi_solve.npz"), n=(8, 8), c=1.0, fq=fq, F=F, x=x, iterations=its)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
