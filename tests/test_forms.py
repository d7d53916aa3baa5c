"""Compiled pointwise integrands (dunefem_amd.forms): a reaction r(u, x) of
-div(kappa grad u) + r(u, x) = f written as HIP code with its symbolic
derivative, compiled into a library the operator loads.

CPU: the symbolic derivative against central differences of the numpy
evaluation, the quadrature-point coordinates against the workload layout,
and the compiled libraries of forms.PREBUILT (they load and name their
expression).  GPU: S[u] = A u + int r(u_h, x) v and the Jacobian against the
oracle (A u by the oracle's apply, the reaction by numpy at the oracle's
quadrature-point values and the oracle's load-vector assembly), the compiled
u^3 against the built-in law, and Newton / scheme solves of manufactured
problems.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as ora
from dunefem_amd import forms
from dunefem_amd import workloads as wl

TOL64 = 1e-12


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("r", forms.PREBUILT)
def test_symbolic_derivative_matches_central_difference(r):
    rng = np.random.default_rng(0)
    uq, px, py, pz = rng.uniform(-1.5, 1.5, 50), rng.uniform(0, 1, 50), rng.uniform(0, 1, 50), rng.uniform(0, 1, 50)
    h = 1e-5
    rp, _ = forms.evaluate(r, uq + h, px, py, pz)
    rm, _ = forms.evaluate(r, uq - h, px, py, pz)
    _, dr = forms.evaluate(r, uq, px, py, pz)
    assert np.abs((rp - rm) / (2 * h) - dr).max() < 1e-6 * max(1.0, np.abs(dr).max())


def test_unknown_symbols_are_rejected():
    with pytest.raises(ValueError):
        forms.as_expr("u**2 + t")


def test_qp_coordinates_match_the_workload_layout():
    n, m = (3, 4, 2), 3
    gx, _ = ora.gauss(m)
    X, Y, Z = forms.qp_coordinates(n, m, gx)
    cx, cy, cz = wl.qp_coords(n[0], m), wl.qp_coords(n[1], m), wl.qp_coords(n[2], m)
    zz, yy, xx = np.meshgrid(cz, cy, cx, indexing="ij")
    for got, grid in ((X, xx), (Y, yy), (Z, zz)):
        assert np.abs(got - wl.vol_layout3(grid, n, m)).max() < 1e-15


@pytest.mark.parametrize("r", forms.PREBUILT)
def test_prebuilt_libraries_load_and_name_their_expression(r):
    path = forms.compile_reaction(r, build=False)
    lib = C.CDLL(path)
    f = lib.fem_reaction_expression
    f.restype = C.c_char_p
    assert f().decode() == str(forms.as_expr(r))
    assert hasattr(lib, "fem_reaction_eval")
    with pytest.raises(FileNotFoundError):
        forms.compile_reaction("u**5 + 7*u", build=False)


# ---- GPU --------------------------------------------------------------------------------
def _oracle_S_J(n, p, kc, r, u, w):
    """S[u] and J(u) w on the oracle: A u + int r(u_h, x) v and A w + int r'(u_h) w v"""
    oop = ora.DGOperator(n, p, 10.0 * p * p, kappa=("cell", kc))
    m = p + 1
    X, Y, Z = forms.qp_coordinates(n, m, ora.gauss(m)[0])
    uq = oop.eval_qp(u)
    rq, drq = forms.evaluate(r, uq, X, Y, Z)
    S = oop.apply(u) + oop.rhs(rq)
    oop.set_reaction_qp(drq)
    J = oop.apply(w)
    return S, J


def _dg(ctx, fa, n, p, kc, precision=None):
    precision = fa.F64 if precision is None else precision
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, 10.0 * p * p, (0, 0, 0), 0.0, fa.KAPPA_CELL,
                             precision=precision)
    op.set_kappa_cell(kc)
    return op


@pytest.mark.gpu
@pytest.mark.parametrize("r", forms.PREBUILT)
@pytest.mark.parametrize("p", [1, 2, 3])
def test_gpu_compiled_reaction_and_jacobian_match_oracle(ctx, r, p):
    import dunefem_amd as fa
    n, rng = (4, 3, 5), np.random.default_rng(30 + p)
    kc = rng.uniform(0.5, 2.0, size=60)
    op = _dg(ctx, fa, n, p, kc)
    op.set_reaction(r, build=False)
    assert op.nonlinear()
    u, w = rng.normal(size=op.N), rng.normal(size=op.N)
    S_ref, J_ref = _oracle_S_J(n, p, kc, r, u, w)
    U, Wv, Y = op.vector(u), op.vector(w), op.vector()
    op(U, Y)
    errS = rel(Y.download(), S_ref)
    J = op.jacobian(U)
    J(Wv, Y)
    errJ = rel(Y.download(), J_ref)
    print("r", r, "p", p, "S rel err", errS, "J rel err", errJ, flush=True)
    assert errS < TOL64 and errJ < TOL64


@pytest.mark.gpu
def test_gpu_compiled_cubic_equals_builtin_law(ctx):
    # the generated u^3 and the law fused in the apply kernel give the same S
    # and J (different code paths: quadrature-point kernel + load-vector
    # assembly against the term inside the apply)
    import dunefem_amd as fa
    n, p, rng = (5, 4, 3), 2, np.random.default_rng(5)
    kc = rng.uniform(0.5, 2.0, size=60)
    gen, law = _dg(ctx, fa, n, p, kc), _dg(ctx, fa, n, p, kc)
    gen.set_reaction("u**3", build=False)
    law.set_reaction_law(1.0, 2)
    U, Wv = gen.vector(rng.normal(size=gen.N)), gen.vector(rng.normal(size=gen.N))
    a, b = gen.vector(), gen.vector()
    gen(U, a)
    law(U, b)
    assert rel(a.download(), b.download()) < 1e-13
    gen.jacobian(U)(Wv, a)
    law.jacobian(U)(Wv, b)
    assert rel(a.download(), b.download()) < 1e-13


@pytest.mark.gpu
def test_gpu_compiled_reaction_errors(ctx):
    import dunefem_amd as fa
    n, p = (3, 3, 3), 2
    kc = np.ones(27)
    with pytest.raises(fa.FemError):
        _dg(ctx, fa, n, p, kc, fa.F32).set_reaction("exp(u)", build=False)
    op = _dg(ctx, fa, n, p, kc)
    with pytest.raises(fa.FemError):   # a library without fem_reaction_eval
        fa._check(fa._lib.fem_op_set_reaction_plugin(op.h, fa.LIBPATH.encode()))
    op.set_reaction("exp(u)", build=False)
    op.set_reaction(None)
    assert not op.nonlinear()


@pytest.mark.gpu
def test_gpu_newton_with_compiled_reaction_returns_manufactured_solution(ctx):
    # -div(kappa grad u) + sinh(u) (2 + cos cos cos) = f: with F = S[u*] for
    # a given u*, Newton from 0 returns u*; then the scheme with the source
    # given as an expression (its load vector against the oracle's) ends
    # with a residual below its tolerance
    import dunefem_amd as fa
    r = forms.PREBUILT[3]
    n, p = (6, 5, 4), 2
    rng = np.random.default_rng(6)
    kc = rng.uniform(0.5, 2.0, size=120)
    op = _dg(ctx, fa, n, p, kc)
    op.set_reaction(r, build=False)
    ustar = op.vector(0.2 * rng.normal(size=op.N))
    F = op.vector()
    op(ustar, F)
    tol = 1e-10 * np.sqrt(F.normSquaredDofs())
    inv = fa.NewtonInverseOperator(ctx, {"nonlinear.tolerance": tol},
                                   linear=dict(method=fa.CG, tolerance=1e-12, errormeasure=fa.RELATIVE,
                                               preconditioning=fa.PRE_BLOCK_JACOBI))
    inv.bind(op)
    W = op.vector()
    inv(F, W)
    info = inv.info()
    print("newton with the compiled reaction", info, flush=True)
    assert info["converged"] and 2 <= info["iterations"] <= 12
    assert rel(W.download(), ustar.download()) < 1e-8
    # scheme: source f = 1 + x y z as an expression, S[u] = F
    Fx = op.rhs_expr("1 + x*y*z")
    oop = ora.DGOperator(n, p, 10.0 * p * p, kappa=("cell", kc))
    X, Y, Z = op.qp_coordinates()
    assert rel(Fx.download(), oop.rhs(1 + X * Y * Z)) < 1e-13
    scheme = fa.Scheme(op, Fx, parameters={"fem.solver.nonlinear.tolerance": 1e-10 * np.sqrt(Fx.normSquaredDofs())},
                       method=fa.CG, tolerance=1e-12, errormeasure=fa.RELATIVE,
                       preconditioning=fa.PRE_BLOCK_JACOBI)
    Us = op.vector()
    si = scheme.solve(Us)
    R = op.vector()
    scheme(Us, R)
    assert si["converged"] and np.sqrt(R.normSquaredDofs()) < 1e-10 * np.sqrt(Fx.normSquaredDofs())


@pytest.mark.gpu
@pytest.mark.parametrize("pre", ["blockjacobi", "multigrid"])
def test_gpu_galerkin_scheme_from_symbolic_integrands(ctx, pre):
    # dune.fem.scheme.galerkin's path: integrands (cellwise kappa, reaction
    # exp(u), a source expression) on a DG Q2 space, scheme.solve(target)
    # with the reference's parameter keys; the solution's residual, and the
    # load vector against the oracle's at the same points.  With multigrid
    # the V-cycle is built on each Jacobian (the reaction exp(u_h) in its
    # levels)
    import dunefem_amd as fa
    n, p = (5, 4, 6), 2
    kc = np.random.default_rng(13).uniform(0.5, 2.0, size=120)
    space = forms.Space(ctx, n, p)
    form = forms.Form(kappa=kc, reaction="exp(u)", source="10*sin(pi*x)*sin(pi*y)*sin(pi*z)")
    scheme = fa.galerkin(form, space, solver="cg",
                         parameters={"fem.solver.nonlinear.tolerance": 1e-9,
                                     "fem.solver.nonlinear.linear.tolerance": 1e-11,
                                     "fem.solver.nonlinear.linear.errormeasure": "relative",
                                     "fem.solver.nonlinear.linear.preconditioning.method": pre})
    U = scheme.op.vector()
    info = scheme.solve(U)
    print("galerkin scheme solve", pre, info, flush=True)
    assert set(info) == {"converged", "iterations", "linear_iterations", "timing"}
    assert info["converged"] and 2 <= info["iterations"] <= 10
    R = scheme.op.vector()
    scheme(U, R)
    assert np.sqrt(R.normSquaredDofs()) < 1e-9
    oop = ora.DGOperator(n, p, 10.0 * p * p, kappa=("cell", kc))
    X, Y, Z = scheme.op.qp_coordinates()
    fq = 10 * np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    assert rel(scheme.F.download(), oop.rhs(fq)) < 1e-13


def test_lagrange_space_arguments():
    with pytest.raises(ValueError):
        forms.Space(None, (4, 4), 1, kind="lagrange")
    with pytest.raises(ValueError):
        forms.Space(None, (4, 4, 4), 2, kind="lagrange")
    with pytest.raises(ValueError):
        forms.Space(None, (4, 4), 2, kind="raviartthomas")


@pytest.mark.gpu
def test_gpu_galerkin_scheme_on_a_lagrange_space(ctx):
    # the reference's most common path: a Lagrange space of order 2 with
    # u = 0 on the boundary, cellwise kappa, reaction exp(u) and a source
    # expression; scheme.solve with Jacobi-preconditioned CG inside Newton,
    # the solution's residual, and the load vector against the oracle's
    import dunefem_amd as fa
    n = (24, 20)
    kc = np.random.default_rng(31).uniform(0.5, 2.0, size=n[0] * n[1])
    space = forms.Space(ctx, n, 2, kind="lagrange")
    form = forms.Form(kappa=kc, reaction="exp(u)", source="10*sin(pi*x)*sin(pi*y)")
    scheme = fa.galerkin(form, space, solver="cg",
                         parameters={"fem.solver.nonlinear.tolerance": 1e-9,
                                     "fem.solver.nonlinear.linear.tolerance": 1e-12,
                                     "fem.solver.nonlinear.linear.errormeasure": "relative",
                                     "fem.solver.nonlinear.linear.maxiterations": 20000,
                                     "fem.solver.nonlinear.linear.preconditioning.method": "jacobi"})
    U = scheme.op.vector()
    info = scheme.solve(U)
    print("lagrange galerkin scheme solve", info, flush=True)
    assert info["converged"] and 2 <= info["iterations"] <= 10
    R = scheme.op.vector()
    scheme(U, R)
    assert np.sqrt(R.normSquaredDofs()) < 1e-9
    X, Y, _ = scheme.op.qp_coordinates()
    oop = ora.CGOperator(n, kappa=np.repeat(kc, 9))
    assert rel(scheme.F.download(), oop.rhs(10 * np.sin(np.pi * X) * np.sin(np.pi * Y))) < 1e-13


def test_kappa_points_match_the_quadrature_coordinates():
    # the volume values of an expression kappa at the points qp_coordinates
    # names (3D and 2D), and the face arrays in the face layout of the
    # workloads (one value per face point of each direction)
    n, m = (3, 4, 2), 3
    kv, kf = forms.kappa_points("exp(x)*(1 + y*z)", n, m)
    X, Y, Z = forms.qp_coordinates(n, m, ora.gauss(m)[0])
    assert np.abs(kv - np.exp(X) * (1 + Y * Z)).max() < 1e-14
    assert [a.shape for a in kf] == [((n[0] + 1) * n[1] * n[2], m * m), (n[0] * (n[1] + 1) * n[2], m * m),
                                     (n[0] * n[1] * (n[2] + 1), m * m)]
    kv2, none = forms.kappa_points("1 + x - y", (5, 3), 3)
    X2, Y2, _ = forms.qp_coordinates_2d((5, 3), 3, ora.gauss(3)[0])
    assert none is None and np.abs(kv2 - (1 + X2 - Y2)).max() < 1e-15
    with pytest.raises(ValueError):
        forms.kappa_points("1 + u", n, m)


@pytest.mark.gpu
def test_gpu_galerkin_with_a_kappa_expression(ctx):
    # kappa = 1 + x y z on a DG Q2 space (pointwise kappa at volume and face
    # points; a linear problem: one Newton step, no nonlinear iteration) and
    # kappa = 1 + x - y / 2 with the reaction exp(u) on the Lagrange space;
    # each operator against the oracle with the same point values
    import dunefem_amd as fa
    n, p = (5, 4, 6), 2
    space = forms.Space(ctx, n, p)
    form = forms.Form(kappa="1 + x*y*z", c=0.5, source="10*sin(pi*x)*sin(pi*y)*sin(pi*z)")
    scheme = fa.galerkin(form, space, solver="cg",
                         parameters={"fem.solver.nonlinear.tolerance": 1e-9,
                                     "fem.solver.nonlinear.linear.tolerance": 1e-12,
                                     "fem.solver.nonlinear.linear.errormeasure": "relative",
                                     "fem.solver.nonlinear.linear.preconditioning.method": "blockjacobi"})
    kv, kf = forms.kappa_points("1 + x*y*z", n, p + 1)
    oop = ora.DGOperator(n, p, 10.0 * p * p, c=0.5, kappa=("qp", kv, kf))
    u = np.random.default_rng(41).normal(size=scheme.op.N)
    Uv, Wv = scheme.op.vector(u), scheme.op.vector()
    scheme.op(Uv, Wv)
    assert rel(Wv.download(), oop.apply(u)) < 1e-12
    U = scheme.op.vector()
    info = scheme.solve(U)
    R = scheme.op.vector()
    scheme(U, R)
    print("dg kappa expression", info, np.sqrt(R.normSquaredDofs()), flush=True)
    assert info["converged"] and info["iterations"] == 0
    assert np.sqrt(R.normSquaredDofs()) < 1e-10 * np.sqrt(scheme.F.normSquaredDofs())
    # with a reaction on DG: the reaction variant of the pointwise-kappa
    # apply, S[u] against the oracle (u^3 as cvol = u^2 at the points)
    sr = fa.galerkin(forms.Form(kappa="1 + x", reaction="u**3", c=0.5), space, build=False)
    kv1, kf1 = forms.kappa_points("1 + x", n, p + 1)
    oop1 = ora.DGOperator(n, p, 10.0 * p * p, c=0.5, kappa=("qp", kv1, kf1))
    oop1.set_reaction_qp(oop1.eval_qp(u) ** 2)
    sr.op(Uv, Wv)
    assert rel(Wv.download(), oop1.apply(u)) < 1e-12
    # Lagrange: kappa expression and reaction
    n2 = (20, 16)
    space2 = forms.Space(ctx, n2, 2, kind="lagrange")
    form2 = forms.Form(kappa="1 + x - y/2", reaction="exp(u)", source="10*sin(pi*x)*sin(pi*y)")
    scheme2 = fa.galerkin(form2, space2, solver="cg",
                          parameters={"fem.solver.nonlinear.tolerance": 1e-9,
                                      "fem.solver.nonlinear.linear.tolerance": 1e-12,
                                      "fem.solver.nonlinear.linear.errormeasure": "relative",
                                      "fem.solver.nonlinear.linear.maxiterations": 20000,
                                      "fem.solver.nonlinear.linear.preconditioning.method": "jacobi"})
    oop2 = ora.CGOperator(n2, kappa=forms.kappa_points("1 + x - y/2", n2, 3)[0].ravel())
    u2 = np.random.default_rng(42).normal(size=scheme2.op.N)
    S2 = scheme2.op.vector()
    scheme2.op.set_reaction(None)
    scheme2.op(scheme2.op.vector(u2), S2)
    assert rel(S2.download(), oop2.apply(u2)) < 1e-12
    scheme2.op.set_reaction("exp(u)", build=False)
    U2 = scheme2.op.vector()
    info2 = scheme2.solve(U2)
    R2 = scheme2.op.vector()
    scheme2(U2, R2)
    print("lagrange kappa expression", info2, flush=True)
    assert info2["converged"] and np.sqrt(R2.normSquaredDofs()) < 1e-9


def test_boundary_values_layouts():
    # Q2: the nodal values of g on the node lattice; DG: the face arrays of
    # the pointwise-kappa layout
    space = forms.Space(None, (4, 3), 2, kind="lagrange")
    g = forms.boundary_values("1 + x + 2*y", space)
    xs, ys = np.linspace(0, 1, 9), np.linspace(0, 1, 7)
    assert g.shape == (63,) and abs(g[0] - 1) < 1e-15 and abs(g[8] - 2) < 1e-15 and abs(g[-1] - 4) < 1e-15
    assert np.abs(g.reshape(7, 9) - (1 + xs[None, :] + 2 * ys[:, None])).max() < 1e-15
    dg = forms.Space(None, (3, 2, 2), 2)
    gf = forms.boundary_values("x*y + z", dg)
    assert [a.shape for a in gf] == [(4 * 2 * 2, 9), (3 * 3 * 2, 9), (3 * 2 * 3, 9)]
    with pytest.raises(ValueError):
        forms.boundary_values("u + 1", dg)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["lagrange", "dglegendre"])
def test_gpu_galerkin_with_dirichlet_data_reproduces_a_polynomial_solution(ctx, kind):
    # u = 1 + x + 2y (+ 3z) with u = g on the boundary, -div grad u + u/2 +
    # u^3 = f: the quadrature rules integrate every term exactly for this u,
    # so the discrete nonlinear problem has u itself as its solution (Q2
    # nodal values; DG values at the quadrature points), reached by Newton
    # with the preconditioner built on each Jacobian
    import dunefem_amd as fa
    if kind == "lagrange":
        space, ue, pre = forms.Space(ctx, (16, 12), 2, kind="lagrange"), "1 + x + 2*y", "jacobi"
    else:
        space, ue, pre = forms.Space(ctx, (6, 5, 4), 2), "1 + x + 2*y + 3*z", "multigrid"
    form = forms.Form(kappa=1.0, c=0.5, reaction="u**3", source="0.5*(%s) + (%s)**3" % (ue, ue), dirichlet=ue)
    scheme = fa.galerkin(form, space, solver="cg", build=False,
                         parameters={"fem.solver.nonlinear.tolerance": 1e-11,
                                     "fem.solver.nonlinear.linear.tolerance": 1e-13,
                                     "fem.solver.nonlinear.linear.errormeasure": "relative",
                                     "fem.solver.nonlinear.linear.maxiterations": 20000,
                                     "fem.solver.nonlinear.linear.preconditioning.method": pre})
    U = scheme.op.vector()
    info = scheme.solve(U)
    print("dirichlet data", kind, info, flush=True)
    assert info["converged"] and info["iterations"] >= 2
    if kind == "lagrange":
        xs, ys = np.linspace(0, 1, 33), np.linspace(0, 1, 25)
        ref = (1 + xs[None, :] + 2 * ys[:, None]).ravel()
        got = U.download()
    else:
        X, Y, Z = scheme.op.qp_coordinates()
        ref = (1 + X + 2 * Y + 3 * Z).ravel()
        got = scheme.op.eval_qp(U).download()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("  max rel error against u", err, flush=True)
    assert err < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["gmres", "bicgstab"])
def test_gpu_galerkin_advection_reaction_with_a_nonsymmetric_solver(ctx, method):
    # -div grad u + b . grad u + u^3 = f with upwind advection on DG Q2: a
    # nonsymmetric Jacobian, Newton with GMRES or BiCGStab and Jacobi inside;
    # the solution's residual, and S[u] against the oracle (advection with
    # the reaction as cvol = u^2 at the points)
    import dunefem_amd as fa
    n, p, b = (6, 5, 4), 2, (1.0, 0.5, -0.25)
    space = forms.Space(ctx, n, p)
    form = forms.Form(kappa=1.0, b=b, reaction="u**3", source="10*sin(pi*x)*sin(pi*y)*sin(pi*z)")
    scheme = fa.galerkin(form, space, solver=method, build=False,
                         parameters={"fem.solver.nonlinear.tolerance": 1e-9,
                                     "fem.solver.nonlinear.linear.tolerance": 1e-12,
                                     "fem.solver.nonlinear.linear.errormeasure": "relative",
                                     "fem.solver.nonlinear.linear.maxiterations": 5000,
                                     "fem.solver.nonlinear.linear.preconditioning.method": "jacobi"})
    u = np.random.default_rng(51).normal(size=scheme.op.N)
    oop = ora.DGOperator(n, p, 10.0 * p * p, b, 0.0, kappa=1.0)
    oop.set_reaction_qp(oop.eval_qp(u) ** 2)
    S = scheme.op.vector()
    scheme.op(scheme.op.vector(u), S)
    assert rel(S.download(), oop.apply(u)) < 1e-12
    U = scheme.op.vector()
    info = scheme.solve(U)
    R = scheme.op.vector()
    scheme(U, R)
    print("advection-reaction", method, info, np.sqrt(R.normSquaredDofs()), flush=True)
    assert info["converged"] and 2 <= info["iterations"] <= 10
    assert np.sqrt(R.normSquaredDofs()) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["dglegendre", "lagrange"])
def test_gpu_newton_solutions_converge_at_the_expected_order(ctx, kind):
    # what the mathematics fixes: for -lap u + u^3 = f with the smooth
    # manufactured u = prod sin(pi x_d) (zero on the boundary), the L2 error
    # of the Newton solution on Q2 spaces falls as h^3 (SIPG DG Q2 on n^3
    # hexes, Lagrange Q2 on n^2 quads): halving h divides it by about 8
    import dunefem_amd as fa
    if kind == "dglegendre":
        ue, d, sizes, pre = "sin(pi*x)*sin(pi*y)*sin(pi*z)", 3, (4, 8), "multigrid"
    else:
        ue, d, sizes, pre = "sin(pi*x)*sin(pi*y)", 2, (8, 16), "jacobi"
    source = "%d*pi**2*(%s) + (%s)**3" % (d, ue, ue)
    errs = []
    for n in sizes:
        space = forms.Space(ctx, (n,) * d, 2, kind=kind)
        scheme = fa.galerkin(forms.Form(kappa=1.0, reaction="u**3", source=source), space, solver="cg",
                             build=False,
                             parameters={"fem.solver.nonlinear.tolerance": 1e-10,
                                         "fem.solver.nonlinear.linear.tolerance": 1e-12,
                                         "fem.solver.nonlinear.linear.errormeasure": "relative",
                                         "fem.solver.nonlinear.linear.maxiterations": 20000,
                                         "fem.solver.nonlinear.linear.preconditioning.method": pre})
        U = scheme.op.vector()
        info = scheme.solve(U)
        assert info["converged"]
        errs.append(scheme.op.l2_error(U, ue))
    rate = np.log2(errs[0] / errs[1])
    print(kind, "L2 errors", errs, "rate", rate, flush=True)
    assert 2.7 < rate < 3.5 and errs[1] < 1e-3


@pytest.mark.gpu
def test_gpu_sipg_q3_with_pointwise_kappa_and_advection_converges_at_order_four(ctx):
    # the discretisation of the headline apply (Q3, kappa at the volume and
    # face points, upwind advection): -div(kappa grad u) + b . grad u = f
    # with kappa = 1 + x y z and the manufactured u = prod sin(pi x_d); GMRES
    # + Jacobi; the L2 error falls as h^4 (p + 1)
    import sympy as sp
    import dunefem_amd as fa
    b = (1.0, 0.5, 0.25)
    X, Y, Z = forms.x, forms.y, forms.z
    ue = sp.sin(sp.pi * X) * sp.sin(sp.pi * Y) * sp.sin(sp.pi * Z)
    kap = 1 + X * Y * Z
    f = -sum(sp.diff(kap * sp.diff(ue, v), v) for v in (X, Y, Z)) + sum(bb * sp.diff(ue, v) for bb, v in zip(b, (X, Y, Z)))
    errs = []
    for n in (3, 6):
        space = forms.Space(ctx, (n, n, n), 3)
        scheme = fa.galerkin(forms.Form(kappa=kap, b=b, source=f), space, solver="gmres",
                             parameters={"fem.solver.nonlinear.tolerance": 1e-10,
                                         "fem.solver.nonlinear.linear.tolerance": 1e-13,
                                         "fem.solver.nonlinear.linear.errormeasure": "relative",
                                         "fem.solver.nonlinear.linear.maxiterations": 20000,
                                         "fem.solver.nonlinear.linear.preconditioning.method": "jacobi"})
        U = scheme.op.vector()
        info = scheme.solve(U)
        assert info["converged"]
        errs.append(scheme.op.l2_error(U, ue))
    rate = np.log2(errs[0] / errs[1])
    print("SIPG Q3 pointwise kappa + advection: L2 errors", errs, "rate", rate, flush=True)
    assert 3.5 < rate < 4.8


@pytest.mark.gpu
@pytest.mark.parametrize("p,sizes,lo,hi", [(1, (4, 8), 1.8, 2.4), (4, (3, 6), 4.3, 5.8)])
def test_gpu_sipg_converges_at_order_p_plus_one(ctx, p, sizes, lo, hi):
    # the other orders' kernels (Q1, and Q4 of the C4 workload) against the
    # h^(p+1) the theory gives: -2 lap u = f, u = prod sin(pi x_d), constant
    # kappa 2, CG + block Jacobi
    import dunefem_amd as fa
    ue = "sin(pi*x)*sin(pi*y)*sin(pi*z)"
    errs = []
    for n in sizes:
        space = forms.Space(ctx, (n, n, n), p)
        scheme = fa.galerkin(forms.Form(kappa=2.0, source="6*pi**2*" + ue), space, solver="cg",
                             parameters={"fem.solver.nonlinear.linear.tolerance": 1e-13,
                                         "fem.solver.nonlinear.linear.errormeasure": "relative",
                                         "fem.solver.nonlinear.linear.maxiterations": 20000,
                                         "fem.solver.nonlinear.linear.preconditioning.method": "blockjacobi"})
        U = scheme.op.vector()
        assert scheme.solve(U)["converged"]
        errs.append(scheme.op.l2_error(U, ue))
    rate = np.log2(errs[0] / errs[1])
    print("SIPG Q%d" % p, "L2 errors", errs, "rate", rate, flush=True)
    assert lo < rate < hi
