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
