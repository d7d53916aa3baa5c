"""GPU CNN kernels (BF16 MFMA, FP32 accumulation) against the numpy oracle,
and the learned preconditioner inside FGMRES."""
import numpy as np
import pytest

import cnn_oracle as cnn
import dunefem_amd as fa
from dunefem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def relf(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30)


def grid_input(rng, n):
    nx, ny, nz = n
    x = rng.normal(size=(nz, ny, nx, 2)).astype(np.float32)
    return cnn.bf16(x).astype(np.float32)


@pytest.mark.parametrize("n,dil", [((10, 9, 8), (1, 2, 4, 1)), ((16, 16, 16), (1, 2, 4, 1)), ((7, 5, 3), (1, 1, 2, 1))])
def test_cnn_forward_and_gradients_match_oracle(ctx, n, dil):
    rng = np.random.default_rng(0)
    net = fa.CNN(ctx, n, dil)
    w = (rng.normal(size=net.nparams) * 0.2).astype(np.float32)
    net.set_weights(w)
    Ws = [x.astype(np.float64) for x in net.split(w)]
    x = grid_input(rng, n)
    t = rng.normal(size=(n[2], n[1], n[0], 1))
    y = net.forward(x.reshape(-1, 2))
    loss_o, grads_o, yo = cnn.loss_and_grads(x.astype(np.float64), t, Ws, list(dil), bf=True)
    ef = relf(y, yo.ravel())
    print("forward rel err", n, ef, flush=True)
    # the oracle rounds to BF16 where the kernels do; what remains is FP32
    # accumulation order, which can move a BF16 rounding of a backpropagated
    # value by one unit in the last place (2^-8 = 3.9e-3). Measured: forward
    # below 3e-5, gradients up to 3.3e-3 (layer 0 on the 7x5x3 grid)
    assert ef < 1e-3
    loss, g = net.loss_grad(x.reshape(-1, 2), t.ravel())
    assert abs(loss - loss_o) < 1e-3 * abs(loss_o)
    gs = net.split(g)
    for l in range(4):
        eg = relf(gs[l], grads_o[l])
        print("layer", l, "grad rel err", eg, flush=True)
        assert eg < 2.0 ** -8


def test_cnn_forward_matches_oracle_on_the_48cubed_grid(ctx):
    # the element grid of the learned-preconditioner workload (48^3 Q2)
    n, dil = (48, 48, 48), (1, 2, 4, 1)
    rng = np.random.default_rng(2)
    net = fa.CNN(ctx, n, dil)
    w = (rng.normal(size=net.nparams) * 0.2).astype(np.float32)
    net.set_weights(w)
    Ws = [x.astype(np.float64) for x in net.split(w)]
    x = grid_input(rng, n)
    y = net.forward(x.reshape(-1, 2))
    yo = cnn.forward(x.astype(np.float64), Ws, list(dil), bf=True)[-1]
    ef = relf(y, yo.ravel())
    print("48^3 forward rel err", ef, flush=True)
    assert ef < 1e-3


def test_cnn_training_reduces_loss_on_a_linear_target(ctx):
    # learn t = smoothing of the first input channel: loss must drop
    n = (12, 12, 12)
    rng = np.random.default_rng(1)
    net = fa.CNN(ctx, n)
    x = grid_input(rng, n)
    t = cnn.conv(x[..., :1].astype(np.float64), np.full((27, 1, 1), 1.0 / 27), 1)
    l0, _ = net.loss_grad(x.reshape(-1, 2), t.ravel())
    w = net.get_weights().astype(np.float64)
    m = np.zeros_like(w); v = np.zeros_like(w)
    for it in range(1, 201):   # Adam on the GPU gradients (host update keeps the test independent of the trainer)
        loss, g = net.loss_grad(x.reshape(-1, 2), t.ravel())
        m = 0.9 * m + 0.1 * g; v = 0.999 * v + 0.001 * g * g
        w -= 3e-3 * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-8)
        net.set_weights(w.astype(np.float32))
    l1, _ = net.loss_grad(x.reshape(-1, 2), t.ravel())
    print("toy training loss", l0, "->", l1, flush=True)
    assert l1 < 0.2 * l0


def learned_setup(ctx, n=12, solves=8, pairs_per_solve=60, epochs=20):
    P = wl.problem3(n=n)
    op64 = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, fa.KAPPA_CELL)
    op64.set_kappa_cell(P["kcell"])
    bj64 = fa.Preconditioner(op64, fa.PRE_BLOCK_JACOBI)
    net = fa.CNN(ctx, P["n"])
    net.set_coefficient(P["kcell"], P["beta"])
    rng = np.random.default_rng(7)
    m = P["p"] + 1
    for s in range(solves):
        modes = wl.fourier_modes(rng, 16, 3)
        c = wl.qp_coords(n, m)
        fq = wl.vol_layout3(wl.sine_sum_grid(modes, [c, c, c]), P["n"], m)
        b = op64.rhs(fq)
        net.collect_pairs(op64, bj64, b, pairs_per_solve, every=4)
    tl, vr = net.train(epochs=epochs, batch=8, lr=1e-3)
    return P, op64, bj64, net, tl, vr


def test_coarse_operator_is_galerkin_projection(ctx):
    # A_c = P^T A P of element means, checked against the oracle's DG operator
    import pyoracle as ora
    P = wl.problem3(n=5)
    oop = ora.DGOperator(P["n"], P["p"], P["beta"], kappa=("cell", P["kcell"]))
    net = fa.CNN(ctx, P["n"])
    net.set_coefficient(P["kcell"], P["beta"])
    rng = np.random.default_rng(3)
    c = rng.normal(size=125)
    u = np.zeros(oop.N)
    u[::27] = c
    ref = oop.apply(u)[::27]
    got = net.coarse_apply(c)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_coarse_cg_matches_textbook_pcg(ctx):
    # the device coarse solve (Chronopoulos-Gear recurrences, FP32 vectors,
    # FP64 inner products) against k steps of the textbook two-reduction
    # Jacobi-PCG in float64 on A_c = P^T A P taken from the oracle's DG
    # operator; many steps must solve A_c x = b
    import pyoracle as ora
    P = wl.problem3(n=8)
    oop = ora.DGOperator(P["n"], P["p"], P["beta"], kappa=("cell", P["kcell"]))
    nb = (P["p"] + 1) ** 3
    V = int(np.prod(P["n"]))
    Ac = np.zeros((V, V))
    for j in range(V):
        u = np.zeros(oop.N)
        u[j * nb] = 1.0
        Ac[:, j] = oop.apply(u)[::nb]
    d = np.diag(Ac).copy()
    net = fa.CNN(ctx, P["n"])
    net.set_coefficient(P["kcell"], P["beta"])
    b = np.random.default_rng(5).normal(size=V).astype(np.float32).astype(np.float64)

    def pcg(k):
        x = np.zeros(V); r = b.copy(); z = r / d; p = z.copy(); rz = r @ z
        for _ in range(k):
            q = Ac @ p
            a = rz / (p @ q)
            x += a * p; r -= a * q; z = r / d
            rz, rz_old = r @ z, rz
            p = z + (rz / rz_old) * p
        return x, rz

    for k in (1, 2, 5, 10, 20, 50):
        x, g, g0 = net.coarse_solve(b, k)
        xr, gr = pcg(k)
        ex = relf(x, xr)
        print("coarse CG steps", k, "x rel err", ex, "gamma", g, gr, flush=True)
        assert ex < 1e-5
        assert abs(g0 - b @ (b / d)) <= 1e-6 * (b @ (b / d))
        if k <= 20:
            assert abs(g - gr) <= 1e-4 * gr
    x, g, g0 = net.coarse_solve(b, 200)
    res = np.linalg.norm(b - Ac @ x) / np.linalg.norm(b)
    print("coarse CG 200 steps: true rel residual", res, flush=True)
    assert res < 1e-5 and g <= 1e-12 * g0
    x, g, g0 = net.coarse_solve(np.zeros(V), 3)
    assert not np.any(x) and g == 0 and g0 == 0


def fgmres_run(ctx, op, pre, F):
    inv = fa.InverseOperator(ctx, method=fa.FGMRES, tolerance=1e-8, errormeasure=fa.RELATIVE,
                             maxiterations=20000, restart=30)
    inv.bind(op, pre)
    X = op.vector()
    inv(F, X)
    return inv.info()


def test_learned_preconditioner_fgmres_ablation_and_fallback(ctx):
    # trained on pairs of the plain CG + block-Jacobi solves of 8 other
    # right-hand sides; measured by FGMRES(30) iterations to a 1e-8 FP64
    # residual on the workload's own right-hand side
    P, op64, bj64, net, tl, vr = learned_setup(ctx)
    print("train energy loss", tl, "val rel coarse energy error", vr, flush=True)
    assert 0 <= vr < 0.6   # a zero correction scores 1
    F = op64.rhs(P["fq"])
    res = {"bj": fgmres_run(ctx, op64, bj64, F)}
    for name, mode in (("coarse_only", fa.LEARNED_COARSE_ONLY), ("net", fa.LEARNED_NET)):
        res[name] = fgmres_run(ctx, op64, fa.LearnedPreconditioner(op64, net, coarse_steps=5, mode=mode), F)
    for name, info in res.items():
        print(name, info, flush=True)
        assert info["converged"] and info["true_rel_residual"] <= 1e-8 and not info["fallback"]
    # the network's gain over the same two-level method without it, and over
    # block Jacobi
    assert res["net"]["iterations"] < 0.75 * res["coarse_only"]["iterations"]
    assert res["coarse_only"]["iterations"] < res["bj"]["iterations"]
    # a network with destroyed weights must not break the solve
    bad = fa.CNN(ctx, P["n"])
    bad.set_coefficient(P["kcell"], P["beta"])
    bad.set_weights(np.full(bad.nparams, 3.0, dtype=np.float32))
    info = fgmres_run(ctx, op64, fa.LearnedPreconditioner(op64, bad, coarse_steps=5), F)
    print("bad network", info, flush=True)
    assert info["converged"] and info["true_rel_residual"] <= 1e-8


def test_learned_correction_after_multigrid(ctx):
    # the multigrid form: a V-cycle, the network's coarse correction of the
    # residual it leaves, one block-Jacobi sweep; trained on post-V-cycle
    # residuals of plain CG + multigrid solves of other right-hand sides.
    # Both forms (with and without the network; the sweep in both) reach the
    # FP64 1e-8 residual; the iterations with and without the network are
    # the measured gain (the V-cycle already holds the element-mean space: on
    # the CPU at 12^3 FGMRES takes 27 iterations with the V-cycle alone, 25
    # with the sweep, 26 with an exact element-mean correction and the sweep)
    n = 12
    P = wl.problem3(n=n)
    op64 = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, fa.KAPPA_CELL)
    op64.set_kappa_cell(P["kcell"])
    mg64 = fa.Preconditioner(op64, fa.PRE_MULTIGRID)
    net = fa.CNN(ctx, P["n"])
    net.set_coefficient(P["kcell"], P["beta"])
    rng = np.random.default_rng(7)
    m = P["p"] + 1
    c = wl.qp_coords(n, m)
    total = 0
    for s in range(12):
        modes = wl.fourier_modes(rng, 16, 3)
        b = op64.rhs(wl.vol_layout3(wl.sine_sum_grid(modes, [c, c, c]), P["n"], m))
        total += net.collect_pairs(op64, mg64, b, 40, every=1)
    tl, vr = net.train(epochs=20, batch=8, lr=1e-3)
    print("post-multigrid pairs", total, "train energy loss", tl, "val rel coarse energy error", vr, flush=True)
    assert total > 100 and 0 <= vr < 1.0
    F = op64.rhs(P["fq"])
    its = {}
    for name, mode in (("mg", fa.LEARNED_OFF), ("mg+net", fa.LEARNED_NET)):
        pre = fa.LearnedPreconditioner(op64, net, coarse_steps=0, mode=mode, smoother=fa.PRE_MULTIGRID)
        info = fgmres_run(ctx, op64, pre, F)
        print(name, info, flush=True)
        assert info["converged"] and info["true_rel_residual"] < 1.01e-8
        its[name] = info["iterations"]
    print("FGMRES(30) iterations multigrid alone / with the network:", its, flush=True)
