"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/dunefem_amd.h, and its host-only helpers agree with the oracle."""
import os
import re

import numpy as np

import dunefem_amd as fa
import pyoracle as ora

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dunefem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = fa.load_library()
    decl = declared_functions()
    assert len(decl) >= 40
    missing = [f for f in decl if not hasattr(L, f)]
    assert not missing, missing
    assert sorted(fa.EXPORTS) == decl


def test_host_helpers():
    L = fa.load_library()
    assert L.fem_version() >= 10000
    for m in range(1, 9):
        x, w = fa.gauss_points(m)
        xo, wo = ora.gauss(m)
        assert np.array_equal(x, xo) and np.array_equal(w, wo)
    p = fa.solver_parameters(**{"fem.solver.method": "cg", "linear.tolerance": "1e-10",
                                "fem.solver.errormeasure": "relative", "fem.solver.gmres.restart": "30"})
    assert p.method == fa.CG and p.tolerance == 1e-10 and p.errormeasure == fa.RELATIVE and p.restart == 30
    try:
        fa.solver_parameters(**{"fem.solver.bogus": "1"})
        raise AssertionError("unknown key accepted")
    except fa.FemError:
        pass


def test_workload_layouts_match_direct_evaluation():
    from dunefem_amd import workloads as wl
    rng = np.random.default_rng(0)
    modes = wl.fourier_modes(rng, 3, 3)
    n, m = (2, 3, 2), 2
    xg, _ = fa.gauss_points(m)

    def g(x, y, z):
        return sum(a * np.sin(k[0] * np.pi * x) * np.sin(k[1] * np.pi * y) * np.sin(k[2] * np.pi * z) for k, a in modes)

    G = wl.sine_sum_grid(modes, [wl.qp_coords(n[0], m), wl.qp_coords(n[1], m), wl.qp_coords(n[2], m)])
    V = wl.vol_layout3(G, n, m)
    for e in range(12):
        ix, iy, iz = e % 2, (e // 2) % 3, e // 6
        for q in range(m ** 3):
            qx, qy, qz = q % m, (q // m) % m, q // (m * m)
            assert abs(V[e, q] - g((ix + xg[qx]) / 2, (iy + xg[qy]) / 3, (iz + xg[qz]) / 2)) < 1e-14
    F = wl.face_layouts3(lambda cx, cy, cz: wl.sine_sum_grid(modes, [cx, cy, cz]), n, m)
    # y-normal face at (ix, fy, iz), point (qx, qz)
    for ix in range(2):
        for fy in range(4):
            for iz in range(2):
                f = ix + 2 * (fy + 4 * iz)
                for q in range(m * m):
                    qx, qz = q % m, q // m
                    assert abs(F[1][f, q] - g((ix + xg[qx]) / 2, fy / 3, (iz + xg[qz]) / 2)) < 1e-14


def test_every_reference_solver_key_is_accepted():
    # the linear-solver keys of the reference's parameter file
    # (dune/fem/solver/test/parameter) and Python scheme parameters
    # (pydemo/testthreading.py:70-71, pydemo/advectiondiffusion.py:69-79)
    keys = {
        "fem.solver.verbose": "false", "fem.solver.method": "gmres", "fem.solver.tolerance": "1e-15",
        "fem.solver.errormeasure": "absolute", "fem.solver.gmres.restart": "15",
        "fem.solver.maxiterations": "100", "fem.solver.linear.verbose": "false",
        "fem.solver.linear.method": "gmres", "fem.solver.linear.tolerance": "1e-12",
        "fem.solver.linear.gmres.restart": "15", "fem.solver.linear.maxiterations": "100",
        "newton.linear.preconditioning.method": "jacobi", "newton.linear.tolerance": "1e-13",
        "linear.verbose": "False".lower(), "linear.tolerance": "1e-5",
        "fem.solver.preconditioning.relaxation": "1.1", "fem.solver.preconditioning.iterations": "1",
        "fem.solver.preconditioning.level": "0", "fem.solver.threading": "true",
        "fem.solver.knolltrick": "false", "fem.solver.maxlineariterations": "1000",
        "fem.solver.absolutetol": "1e-11", "fem.solver.reductiontol": "1e-11",
    }
    p = fa.solver_parameters(**keys)
    assert p.method == fa.GMRES and p.restart == 15 and p.maxiterations == 1000
    assert p.preconditioning == fa.PRE_JACOBI and p.tolerance == 1e-11 and p.verbose == 0
    for v, kind in (("none", fa.PRE_NONE), ("jacobi", fa.PRE_JACOBI), ("blockjacobi", fa.PRE_BLOCK_JACOBI),
                    ("multigrid", fa.PRE_MULTIGRID)):
        assert fa.solver_parameters(**{"fem.solver.preconditioning.method": v}).preconditioning == kind
    # the reference's assembled-matrix preconditioners are refused by name,
    # Newton keys and malformed values are argument errors
    import pytest
    for v in ("ilu", "sor", "ssor", "gauss-seidel", "amg-jacobi", "ildl", "oas", "icc"):
        with pytest.raises(fa.FemError, match="assembled matrix"):
            fa.solver_parameters(**{"fem.solver.preconditioning.method": v})
    for k, v in (("fem.solver.nonlinear.tolerance", "1e-8"), ("nonlinear.verbose", "true"),
                 ("fem.solver.tolerance", "abc"), ("fem.solver.maxiterations", "1.5"),
                 ("fem.solver.verbose", "maybe"), ("fem.solver.preconditioning.method", "bogus")):
        with pytest.raises(fa.FemError):
            fa.solver_parameters(**{k: v})
    d = fa.solver_parameters()
    assert d.preconditioning == fa.PRE_NONE and d.relaxation == 1.1 and d.precond_iterations == 1


def test_every_newton_key_is_accepted():
    # NewtonParameter's keys (dune/fem/solver/newtoninverseoperator.hh:101-380)
    # with their defaults, the deprecated prefix newton. and lineSearch
    # spelling, and the split of scheme parameters between Newton and its
    # linear solver
    import pytest
    d = fa.newton_parameters()
    assert d.tolerance == 1e-6 and d.maxiterations == 2 ** 31 - 1 and d.maxlineariterations < 0
    assert d.linesearch == fa.LINESEARCH_NONE and d.forcing == fa.FORCING_NONE and d.simplified == 0
    assert d.eisenstatwalker_etamax == 0.99 and d.eisenstatwalker_gamma == 0.1
    p = fa.newton_parameters({
        "fem.solver.nonlinear.tolerance": "1e-9", "fem.solver.nonlinear.maxiterations": "25",
        "fem.solver.nonlinear.maxlineariterations": "4000", "fem.solver.nonlinear.linesearch": "simple",
        "fem.solver.nonlinear.maxlinesearchiterations": "7", "nonlinear.forcing": "eisenstatwalker",
        "nonlinear.eisenstatwalker.etamax": "0.5", "nonlinear.eisenstatwalker.gamma": "0.3",
        "nonlinear.simplified": "true", "newton.verbose": "false"})
    assert (p.tolerance, p.maxiterations, p.maxlineariterations, p.linesearch, p.maxlinesearchiterations) == \
        (1e-9, 25, 4000, fa.LINESEARCH_SIMPLE, 7)
    assert p.forcing == fa.FORCING_EISENSTATWALKER and p.eisenstatwalker_etamax == 0.5
    assert p.eisenstatwalker_gamma == 0.3 and p.simplified == 1 and p.verbose == 0
    assert fa.newton_parameters({"nonlinear.lineSearch": "simple"}).linesearch == fa.LINESEARCH_SIMPLE
    assert fa.newton_parameters({"nonlinear.linear.tolerance.strategy": "eisenstatwalker"}).forcing == \
        fa.FORCING_EISENSTATWALKER
    for k, v in (("nonlinear.tolerance", "0"), ("nonlinear.linesearch", "wolfe"), ("nonlinear.forcing", "x"),
                 ("nonlinear.linear.tolerance", "1e-8"), ("fem.solver.tolerance", "1e-8"),
                 ("nonlinear.bogus", "1"), ("nonlinear.maxiterations", "-1")):
        with pytest.raises(fa.FemError):
            fa.newton_parameters({k: v})
    # scheme parameters: nonlinear.* to Newton, the rest (and nonlinear.linear.*) to the linear solver
    assert fa._is_newton_key("fem.solver.nonlinear.tolerance") and fa._is_newton_key("newton.linesearch")
    assert fa._is_newton_key("fem.solver.nonlinear.linear.tolerance.strategy")
    assert not fa._is_newton_key("fem.solver.nonlinear.linear.method")
    assert not fa._is_newton_key("fem.solver.method") and not fa._is_newton_key("linear.tolerance")
