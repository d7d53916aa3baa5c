"""Benchmark: matrix-free SIPG DG apply (and solves) on one MI355X.

Default workload (the apply the headline metric is quoted on): 3D SIPG DG Q3
advection-diffusion on 64^3 hexes (16,777,216 DoFs), pointwise kappa at the
volume and face quadrature points, b = (1, 0.5, 0.25), beta = 10 p^2, input
DoFs ~ N(0,1).  A step is one FP64 operator apply w = A u with all inputs
resident in HBM.  ``value`` is FP64 GDoF/s = N / (time per apply).

Also reported on the same line:
  roofline      algorithmic bytes per apply / measured kernel time vs 8 TB/s
  cpu_baseline  the oracle (the reference algorithm restated: dense
                per-element tables, element/face loop, OpenMP element chunks)
                on the host cores, same 64^3 workload
  mixed         FP32 and BF16-MFMA applies: GDoF/s and relative error vs FP64
  solve         time-to-solve comparisons (48^3 Q2 channels, 96^3 Q4 log-normal)
  nonlinear     Newton solves of -div(kappa grad u) + c3 u^3 = f (48^3 Q2)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--solve]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dune-fem_amd"))

import numpy as np

HBM_PEAK_GBS = 8000.0
# CG steps captured in a HIP graph (off under --no-graph, for profiler runs:
# rocprofv3's kernel trace does not follow graph launches reliably)
GRAPH = 1


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="solve section without the learned preconditioner")
    ap.add_argument("--no-combined", action="store_true", help="skip the 96^3 Q4 solves")
    ap.add_argument("--no-nonlinear", action="store_true", help="skip the Newton solves")
    ap.add_argument("--nonlinear-only", action="store_true", help="only the Newton section (profiling)")
    ap.add_argument("--no-graph", action="store_true", help="no HIP graph capture of CG steps (profiler runs)")
    ap.add_argument("--apply-only", choices=["c1", "c3", "c4"], default=None,
                    help="time only the operator apply of one workload (for profiling): c1 = 64^3 Q3 pointwise "
                         "kappa with advection, c3 = 48^3 Q2 channels, c4 = 96^3 Q4 log-normal")
    ap.add_argument("--prec", choices=["f64", "f32", "bf16"], default="f64", help="precision of --apply-only")
    ap.add_argument("--calib", action="store_true",
                    help="also run a plain FP64 axpy of known size (HBM counter calibration)")
    ap.add_argument("--solve-only", choices=["c3", "c4"], default=None,
                    help="only FP64 CG on one workload, --reps times (for profiling); --pre mg (multigrid, --prec "
                         "f32 runs the V-cycle in FP32) or bj (exact element block Jacobi)")
    ap.add_argument("--pre", choices=["mg", "bj"], default="mg", help="preconditioner of --solve-only")
    ap.add_argument("--reps", type=int, default=2, help="solves of --solve-only")
    return ap.parse_args()


def time_applies(fa, ctx, op, U, W, steps, warmup):
    for _ in range(warmup):
        op(U, W)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.mark(0)
    for _ in range(steps):
        op(U, W)
    ctx.mark(1)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ctx.elapsed_ms()
    return wall, ev_ms / steps


def measured_traffic(p, n):
    """HBM bytes per launch of the FP64 apply from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against an axpy of known
    size); only for the workload it was measured on, else None."""
    f = os.path.join(ROOT, "profiles", "k_dg_apply_q3_64cubed_traffic.json")
    if p != 3 or tuple(n) != (64, 64, 64) or not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh)["traffic_bytes_per_launch"]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _info(v):
    return {"converged": v["converged"], "iterations": v["iterations"], "outer_iterations": v["outer_iterations"],
            "time_s": round(v["timing"][0], 4), "true_rel_residual": v["true_rel_residual"],
            "fallback": v.get("fallback", False)}


def plain_section(fa, wl, ctx, cpu=True):
    """The reference's CPU-runnable case: 2D continuous Q2, 64x64 quads,
    -lap u + u = f, strong u = 0, CG + Jacobi to 1e-10 relative.  GPU solve
    (FP64) against the oracle's CG on one host core."""
    P = wl.problem0(n=64)
    op = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, P["n"], 2, c=1.0)
    pre = fa.Preconditioner(op, fa.PRE_JACOBI)
    F = op.rhs(P["fq"])
    res = {}
    for rep in range(2):   # the first call includes one-time set-up
        X = op.vector()
        inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-10, errormeasure=fa.RELATIVE, maxiterations=5000)
        inv.bind(op, pre)
        inv(F, X)
        res = _info(inv.info())
    out = {"workload": "2D continuous Q2, 64x64 quads (16,641 DoFs), -lap u + u = f (8 Fourier modes), strong "
                       "u = 0, CG + Jacobi, rel tol 1e-10", "gpu_fp64_cg_jacobi": res}
    if cpu:
        import pyoracle as ora
        oop = ora.CGOperator(P["n"], c=1.0)
        Fh = oop.rhs(P["fq"])
        t0 = time.perf_counter()
        xo, ito = ora.cg(oop, Fh, pre=ora.jacobi(1.0 / oop.diag()), tol=1e-10, maxit=5000, criteria=1)
        cpu_s = time.perf_counter() - t0
        out["cpu_oracle_cg_jacobi"] = {"iterations": int(ito), "time_s": round(cpu_s, 4), "cores": 1}
        out["iterations_match"] = int(ito) == res["iterations"]
    log("plain", out)
    return out


def _dg_cell(fa, ctx, P, prec):
    op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, fa.KAPPA_CELL,
                             precision=prec)
    op.set_kappa_cell(P["kcell"])
    return op


def _runner(fa, ctx, op64, F, runs, tag):
    def run(name, method, pre, inner=None, maxit=20000, **kw):
        X = op64.vector()
        inv = fa.InverseOperator(ctx, method=method, tolerance=1e-8, errormeasure=fa.RELATIVE,
                                 maxiterations=maxit, restart=30, residual_check=1, **kw)
        if inner is None:
            inv.bind(op64, pre)
        else:
            inv.bind(op64, None, inner, pre)
        inv(F, X)
        runs[name] = _info(inv.info())
        log(tag, name, runs[name])
    return run


def _train_learned(fa, wl, ctx, P, op64, mg64, pairs, epochs, lr, out, key):
    """Pairs (post-V-cycle residual, error) from plain FP64 CG + multigrid
    solves of other seeded right-hand sides, every iterate; training on the GPU."""
    n = P["n"][0]
    net = fa.CNN(ctx, P["n"])
    net.set_coefficient(P["kcell"], P["beta"])
    rng = np.random.default_rng(7)
    m = P["p"] + 1
    c = wl.qp_coords(n, m)
    t0 = time.perf_counter()
    npairs, solves = 0, 0
    while npairs < pairs and solves < 400:
        modes = wl.fourier_modes(rng, 16, 3)
        b = op64.rhs(wl.vol_layout3(wl.sine_sum_grid(modes, [c, c, c]), P["n"], m))
        npairs += net.collect_pairs(op64, mg64, b, pairs - npairs, every=1)
        solves += 1
    ctx.synchronize()
    collect_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tl, vr = net.train(epochs=epochs, batch=8, lr=lr)
    train_s = time.perf_counter() - t0
    log(key, "trained on", npairs, "pairs in", round(train_s, 1), "s, val", vr)
    out[key] = {"pairs": npairs, "solves": solves, "collect_s": round(collect_s, 2), "epochs": epochs, "lr": lr,
                "train_s": round(train_s, 2), "train_energy_loss": round(tl, 4),
                "val_rel_coarse_energy_error": round(vr, 4),
                "note": "pairs: element means of the residual one V-cycle leaves and of the error, from plain "
                        "FP64 CG + multigrid solves; loss 1 = zero correction, 0 = exact coarse solve"}
    return net


def solve_section(fa, wl, ctx, n=48, pairs=2000, epochs=3, lr=1e-4, learned=True):
    """High-contrast channel problem: time to a 1e-8 FP64 relative residual.
    The plain FP64 GPU path (CG + exact block Jacobi, the reference's Krylov
    recurrence with element-block preconditioning), block-Jacobi FGMRES(30)
    (the configuration's comparison), CG with the multigrid preconditioner
    in FP64, FP64 CG with an FP32 V-cycle, FP64 refinement around FP32 inner
    solves with an FP32 multigrid and with a multigrid on the BF16-MFMA
    operator, and the learned correction after the V-cycle in FGMRES(30)
    (V-cycle, network correction, one fine block-Jacobi sweep): with and
    without the network (FP64), with an FP32 V-cycle in FP64 FGMRES, and
    inside FP64 refinement."""
    P = wl.problem3(n=n)
    E = n ** 3
    N = E * (P["p"] + 1) ** 3
    out = {"workload": "3D SIPG DG Q2 diffusion, %d^3 hexes (%d DoFs), kappa=10^(4H), H binary channels 20%%, "
                       "f = 16 Fourier modes; rel tol 1e-8 on the FP64 true residual" % (n, N), "dofs": N}
    op64 = _dg_cell(fa, ctx, P, fa.F64)
    op32 = _dg_cell(fa, ctx, P, fa.F32)
    opbf = _dg_cell(fa, ctx, P, fa.BF16MFMA)
    F = op64.rhs(P["fq"])
    runs = {}
    run = _runner(fa, ctx, op64, F, runs, "solve")
    bj64 = fa.Preconditioner(op64, fa.PRE_BLOCK_JACOBI)
    t0 = time.perf_counter()
    mg64 = fa.Preconditioner(op64, fa.PRE_MULTIGRID)
    mg32 = fa.Preconditioner(op32, fa.PRE_MULTIGRID)
    mgbf = fa.Preconditioner(opbf, fa.PRE_MULTIGRID)
    ctx.synchronize()
    out["multigrid_setup_s_three_precisions"] = round(time.perf_counter() - t0, 3)
    run("fp64_cg_bj", fa.CG, bj64)
    run("fp64_fgmres_bj", fa.FGMRES, bj64, maxit=3000)
    run("fp64_cg_mg", fa.CG, mg64, graph=GRAPH)
    # FP64 CG whose V-cycle runs in FP32 (rounded in, widened out)
    run("mixed_fp64cg_fp32mg", fa.CG, mg32, graph=GRAPH)
    run("mixed_fp32_cg_mg_ir", fa.CG, mg32, op32, refinement=1, inner_tolerance=1e-3, inner_maxiterations=2000,
        graph=GRAPH)
    # the V-cycle on the BF16-MFMA operator: the inner Krylov solve keeps the
    # FP32 operator; FGMRES, since BF16 rounding makes the V-cycle slightly
    # nonlinear
    run("mixed_bf16mg_fgmres_ir", fa.FGMRES, mgbf, op32, refinement=1, inner_tolerance=1e-3,
        inner_maxiterations=2000)
    if learned:
        net = _train_learned(fa, wl, ctx, P, op64, mg64, pairs, epochs, lr, out, "learned_training")
        run("fp64_fgmres_mg", fa.FGMRES,
            fa.LearnedPreconditioner(op64, net, coarse_steps=0, mode=fa.LEARNED_OFF, smoother=fa.PRE_MULTIGRID))
        run("fp64_fgmres_mg_learned", fa.FGMRES,
            fa.LearnedPreconditioner(op64, net, coarse_steps=0, mode=fa.LEARNED_NET, smoother=fa.PRE_MULTIGRID))
        lp32 = fa.LearnedPreconditioner(op32, net, coarse_steps=0, mode=fa.LEARNED_NET, smoother=fa.PRE_MULTIGRID)
        # FP64 FGMRES with the network and the V-cycle in FP32
        run("mixed_fp64fgmres_fp32mg_learned", fa.FGMRES, lp32)
        run("mixed_fp32_fgmres_mg_learned_ir", fa.FGMRES, lp32, op32, refinement=1, inner_tolerance=1e-3,
            inner_maxiterations=2000)
    out["runs"] = runs
    base = runs["fp64_cg_bj"]["time_s"]
    out["speedup_vs_fp64_cg_bj"] = {k: round(base / v["time_s"], 2) for k, v in runs.items()}
    mgb = runs["fp64_cg_mg"]["time_s"]
    out["speedup_vs_fp64_cg_mg"] = {k: round(mgb / v["time_s"], 2) for k, v in runs.items()}
    return out


def combined_section(fa, wl, ctx, n=96, learned=True, pairs=400, epochs=2, lr=1e-4):
    """Q4 log-normal problem (110.6 M DoFs): the plain FP64 GPU path (CG +
    block Jacobi) against CG + multigrid in FP64, FP64 CG with an FP32
    V-cycle, FP64 refinement around FP32 inner solves, and the learned
    correction after an FP32 V-cycle in FP64 FGMRES and inside FP64
    refinement (trained on a reduced budget of pairs at this size)."""
    P = wl.problem4(n=n)
    E = n ** 3
    N = E * (P["p"] + 1) ** 3
    out = {"workload": "3D SIPG DG Q4 diffusion, %d^3 hexes (%d DoFs), log-normal kappa (sigma 1.5, corr. "
                       "length 0.1, per element), f = 32 Fourier modes; rel tol 1e-8 on the FP64 true residual"
                       % (n, N), "dofs": N}
    op64 = _dg_cell(fa, ctx, P, fa.F64)
    F = op64.rhs(P["fq"])
    del P["fq"]
    runs = {}
    run = _runner(fa, ctx, op64, F, runs, "combined")
    run("fp64_cg_bj", fa.CG, fa.Preconditioner(op64, fa.PRE_BLOCK_JACOBI))
    t0 = time.perf_counter()
    mg64 = fa.Preconditioner(op64, fa.PRE_MULTIGRID)
    ctx.synchronize()
    out["multigrid_setup_s_fp64"] = round(time.perf_counter() - t0, 3)
    run("fp64_cg_mg", fa.CG, mg64, graph=GRAPH)
    op32 = _dg_cell(fa, ctx, P, fa.F32)
    mg32 = fa.Preconditioner(op32, fa.PRE_MULTIGRID)
    run("mixed_fp64cg_fp32mg", fa.CG, mg32, graph=GRAPH)
    run("mixed_fp32_cg_mg_ir", fa.CG, mg32, op32, refinement=1, inner_tolerance=1e-3, inner_maxiterations=2000,
        graph=GRAPH)
    # (a V-cycle on the BF16-MFMA operator does not converge at this size:
    # BF16 rounding noise swamps the smooth residual; not run here)
    if learned:
        net = _train_learned(fa, wl, ctx, P, op64, mg64, pairs, epochs, lr, out, "learned_training")
        lp32 = fa.LearnedPreconditioner(op32, net, coarse_steps=0, mode=fa.LEARNED_NET, smoother=fa.PRE_MULTIGRID)
        run("mixed_fp64fgmres_fp32mg_learned", fa.FGMRES, lp32)
        run("mixed_fp32_fgmres_mg_learned_ir", fa.FGMRES, lp32, op32, refinement=1, inner_tolerance=1e-3,
            inner_maxiterations=2000)
    out["runs"] = runs
    base = runs["fp64_cg_bj"]["time_s"]
    out["speedup_vs_fp64_cg_bj"] = {k: round(base / v["time_s"], 2) for k, v in runs.items()}
    mgb = runs["fp64_cg_mg"]["time_s"]
    out["speedup_vs_fp64_cg_mg"] = {k: round(mgb / v["time_s"], 2) for k, v in runs.items()}
    return out


def nonlinear_section(fa, wl, ctx, n=48, p=2, c3=1.0e4, amp=50.0, cpu=True):
    """Newton on -div(kappa grad u) + c3 u^3 = f: SIPG DG Q2 on n^3 hexes,
    log-normal cellwise kappa (sigma 1), f = amp x (32 Fourier modes), the
    cubic reaction at the volume quadrature points.  Newton (the reference's
    NewtonInverseOperator loop) to |S[w] - F| < 1e-8 |F| from w = 0 with
    matrix-free Jacobians: CG + block Jacobi rebuilt on each Jacobian, the
    same with Eisenstat-Walker forcing, CG with a multigrid built once on
    the diffusion part (FP64, and FP32 widened into the FP64 CG), and FP64
    refinement around FP32 inner CG on an FP32 twin of each Jacobian.  Also the
    apply rates of S[u], of the Jacobian and of the linear part, with their
    algorithmic bytes against 8 TB/s, and the oracle's S[u] on the host."""
    P = wl.problem4(n=n, p=p, sigma=1.0)
    E = n ** 3
    N = E * (p + 1) ** 3
    out = {"workload": "3D SIPG DG Q%d, %d^3 hexes (%d DoFs), -div(kappa grad u) + %g u^3 = f, log-normal kappa "
                       "(sigma 1, per element), f = %g x 32 Fourier modes; Newton to |S[w] - F| < 1e-8 |F| from "
                       "w = 0" % (p, n, N, c3, amp), "dofs": N}
    lin64 = _dg_cell(fa, ctx, P, fa.F64)       # the diffusion part alone
    S = _dg_cell(fa, ctx, P, fa.F64)
    S.set_reaction_law(c3, 2)
    F = lin64.rhs(amp * P["fq"])
    fnorm = float(np.sqrt(F.normSquaredDofs()))
    tol = 1e-8 * fnorm
    runs = {}

    def run(name, nl, linear, pre=None):
        inv = fa.NewtonInverseOperator(ctx, dict(nl, **{"nonlinear.tolerance": tol}), linear=linear)
        inv.bind(S, pre)
        W = S.vector()
        ctx.synchronize()
        t0 = time.perf_counter()
        inv(F, W)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        inv.unbind()
        i = inv.info()
        R = S.vector()
        S(W, R)
        R.axpy(-1.0, F)
        runs[name] = {"converged": i["converged"], "failure": i["failure"], "newton_iterations": i["iterations"],
                      "linear_iterations": i["linear_iterations"], "time_s": round(wall, 4),
                      "timing_total_jacobian_solve": [round(t, 4) for t in i["timing"]],
                      "rel_residual": float(np.sqrt(R.normSquaredDofs()) / fnorm)}
        log("nonlinear", name, runs[name])
        return W

    exact = dict(method=fa.CG, tolerance=1e-10, errormeasure=fa.RELATIVE, maxiterations=20000)
    W = run("newton_cg_bj", {}, dict(exact, preconditioning=fa.PRE_BLOCK_JACOBI))
    run("newton_ew_cg_bj", {"nonlinear.forcing": "eisenstatwalker"},
        dict(method=fa.CG, maxiterations=20000, preconditioning=fa.PRE_BLOCK_JACOBI))
    t0 = time.perf_counter()
    mg64 = fa.Preconditioner(lin64, fa.PRE_MULTIGRID)
    lin32 = _dg_cell(fa, ctx, P, fa.F32)
    mg32 = fa.Preconditioner(lin32, fa.PRE_MULTIGRID)
    ctx.synchronize()
    out["multigrid_setup_s"] = round(time.perf_counter() - t0, 3)
    run("newton_cg_mg_diffusion", {}, dict(exact, graph=GRAPH), mg64)
    run("newton_ew_cg_mg32_diffusion", {"nonlinear.forcing": "eisenstatwalker"},
        dict(method=fa.CG, maxiterations=20000, graph=GRAPH), mg32)
    # the multigrid built on each Jacobian (reaction taken into its levels)
    run("newton_cg_mg_jacobian", {}, dict(exact, preconditioning=fa.PRE_MULTIGRID, graph=GRAPH))
    run("newton_ew_cg_mg_jacobian", {"nonlinear.forcing": "eisenstatwalker"},
        dict(method=fa.CG, maxiterations=20000, preconditioning=fa.PRE_MULTIGRID, graph=GRAPH))
    # each Newton system by FP64 refinement around FP32 inner CG on an FP32
    # twin of the Jacobian, the FP32 multigrid of the diffusion part inside
    run("newton_ew_ir_fp32cg_mg32_diffusion", {"nonlinear.forcing": "eisenstatwalker"},
        dict(method=fa.CG, maxiterations=20000, refinement=1, inner_tolerance=1e-2, inner_maxiterations=2000,
             graph=GRAPH), mg32)
    out["runs"] = runs
    base = runs["newton_cg_bj"]["time_s"]
    out["speedup_vs_newton_cg_bj"] = {k: round(base / v["time_s"], 2) for k, v in runs.items()}

    # apply rates at the solution: S[u] (law in the kernel), J(u) w (cvol
    # read), the linear part; cellwise kappa: E values
    J = S.jacobian(W)
    rates = {}
    for name, op, nbytes in (("S_apply", S, 8 * (2 * N + E)), ("jacobian_apply", J, 8 * (3 * N + E)),
                             ("linear_apply", lin64, 8 * (2 * N + E))):
        Y = op.vector()
        wall, kms = time_applies(fa, ctx, op, W, Y, 50, 5)
        rates[name] = {"gdofs": round(N / (wall / 50) / 1e9, 3), "kernel_ms": round(kms, 5),
                       "alg_bytes": nbytes, "hbm_frac": round(nbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["applies"] = rates
    log("nonlinear applies", rates)

    if cpu:
        # the oracle's S[u] at the solution (the reference algorithm restated:
        # cvol = c3 u^2 at the quadrature points, then the dense-table apply),
        # one apply on the host cores
        import pyoracle as ora
        threads = cpu_threads()
        oop = ora.DGOperator(P["n"], p, P["beta"], kappa=("cell", P["kcell"]))
        w = W.download()
        t0 = time.perf_counter()
        uq = oop.eval_qp(w)
        oop.set_reaction_qp(c3 * uq * uq)
        so = oop.apply(w, nthreads=threads)
        tc = time.perf_counter() - t0
        Y = S.vector()
        S(W, Y)
        out["cpu_oracle_S_apply"] = {"seconds": round(tc, 3), "gdofs": round(N / tc / 1e9, 5), "cores": threads,
                                     "gpu_vs_oracle_rel_err": float(np.abs(Y.download() - so).max() /
                                                                    np.abs(so).max())}
    return out


def batch_section(fa, wl, ctx, count=256):
    """256 uneven 2D Q2 problems: one batched FP32-CG + FP64-refinement solve
    vs the same problems solved one by one with the FP64 GPU CG."""
    t0 = time.perf_counter()
    inst = wl.problem2_instances(count=count)
    gen_s = time.perf_counter() - t0
    sides = [d["n"] for d in inst]
    B = fa.Batch(ctx, sides, c=0.0)
    B.set_kappa_qp(np.concatenate([d["kvol"] for d in inst]))
    (tx, ty), tune = B.autotune()
    log("batch: autotuned tile", tx, ty)
    F = B.rhs(np.concatenate([d["fq"] for d in inst]))
    X = B.vector()
    B.solve(F, X, tol=1e-8, use_graph=bool(GRAPH))          # warm-up (captures the graph)
    X.clear()
    info = B.solve(F, X, tol=1e-8, use_graph=bool(GRAPH))
    # per-instance FP64 loop: one operator + CG(Jacobi) solve per problem
    Fh = F.download()
    t1 = time.perf_counter()
    loop_its, loop_res, loop_conv = 0, 0.0, 0
    for i, d in enumerate(inst):
        op = fa.GalerkinOperator(ctx, fa.OP_CG_Q2_2D, d["n"], 2, c=0.0, kappa_mode=fa.KAPPA_QP)
        op.set_kappa_qp(d["kvol"])
        pre = fa.Preconditioner(op, fa.PRE_JACOBI)
        inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, errormeasure=fa.RELATIVE,
                                 maxiterations=100000, residual_check=1)
        inv.bind(op, pre)
        b = op.vector(Fh[B.offsets[i]:B.offsets[i + 1]])
        x = op.vector()
        inv(b, x)
        r = inv.info()
        loop_its += abs(r["iterations"])
        loop_res = max(loop_res, r["true_rel_residual"])
        loop_conv += int(r["converged"])
    ctx.synchronize()
    loop_s = time.perf_counter() - t1
    log("batch: batched", round(info["time_s"], 3), "s, per-instance loop", round(loop_s, 3), "s")
    return {"workload": "%d x 2D continuous Q2, -div(kappa grad u) = f, sides U{32..256} (%d DoFs total), "
                        "log-normal kappa (sigma 1, corr. length 0.1) at quadrature points, 8-mode forcing; "
                        "tol 1e-8 FP64 true residual per instance" % (count, B.N),
            "dofs": int(B.N), "input_generation_s": round(gen_s, 2),
            "tile_shape": [tx, ty], "autotune_ms_per_apply": [round(m, 4) for m in tune],
            "batched_mixed": {"time_s": round(info["time_s"], 4), "converged": info["converged"],
                              "outer_iterations": info["outer_iterations"],
                              "max_inner_iterations": info["max_inner_iterations"],
                              "max_true_rel_residual": info["max_true_rel_residual"]},
            "per_instance_fp64_loop": {"time_s": round(loop_s, 4), "converged": loop_conv,
                                       "total_iterations": loop_its, "max_true_rel_residual": loop_res},
            "speedup": round(loop_s / info["time_s"], 2),
            "problems_per_s_batched": round(count / info["time_s"], 1)}


def cpu_threads():
    """Threads of the CPU baseline: OMP_NUM_THREADS when set (the host-core
    share a one-GPU job is given), else every core the process may run on."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        return env
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_info(ora):
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": affinity,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "oracle_library": ora._LIBNAME,
            "oracle_flags": "-O3 -fopenmp " + ("-march=native" if "native" in ora._LIBNAME else "-march=x86-64-v3")}


def apply_only(a, fa, wl):
    """One workload's operator apply, timed alone (the profiling driver)."""
    ctx = fa.Context(0)
    prec = {"f64": fa.F64, "f32": fa.F32, "bf16": fa.BF16MFMA}[a.prec]
    if a.apply_only == "c1":
        P = wl.problem1(n=a.n)
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], P["b"], 0.0, fa.KAPPA_QP,
                                 precision=prec)
        op.set_kappa_qp(P["kvol"], P["kfaces"])
        n, k = P["n"], P["p"] + 1
        qc = n[0] * n[1] * n[2] * k ** 3 + sum((n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2))
                                               for d in range(3)) * k * k
        u = P["u"]
    else:
        P = wl.problem3() if a.apply_only == "c3" else wl.problem4()
        del P["fq"]
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, P["n"], P["p"], P["beta"], (0, 0, 0), 0.0, fa.KAPPA_CELL,
                                 precision=prec)
        op.set_kappa_cell(P["kcell"])
        qc = int(np.prod(P["n"]))
        u = np.random.default_rng(1).normal(size=op.N)
    U, W = op.vector(u), op.vector()
    del u
    wall, kern_ms = time_applies(fa, ctx, op, U, W, a.steps, a.warmup)
    if a.calib and prec == fa.F64:
        # y += a x on N doubles (reads 16 N, writes 8 N bytes): the HBM counter calibration
        for _ in range(5):
            W.axpy(0.5, U)
        ctx.synchronize()
    sv = 8 if prec == fa.F64 else 4
    b_alg = sv * (2 * op.N + qc)
    ach = b_alg / (kern_ms * 1e-3) / 1e9
    print(json.dumps({"apply_only": a.apply_only, "precision": a.prec, "dofs": int(op.N), "kernel_ms": round(kern_ms, 5),
                      "gdofs": round(op.N / (kern_ms * 1e-3) / 1e9, 3), "alg_bytes_per_launch": int(b_alg),
                      "achieved_gbs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}), flush=True)


def solve_only(a, fa, wl):
    """FP64 CG with the multigrid or the block-Jacobi preconditioner on the
    C3 or C4 workload, solved --reps times without graph capture (the
    profiling driver of the solve kernels)."""
    ctx = fa.Context(0)
    P = wl.problem3() if a.solve_only == "c3" else wl.problem4()
    op64 = _dg_cell(fa, ctx, P, fa.F64)
    F = op64.rhs(P["fq"])
    del P["fq"]
    opm = op64 if a.prec == "f64" else _dg_cell(fa, ctx, P, fa.F32)
    pre = fa.Preconditioner(opm, fa.PRE_MULTIGRID if a.pre == "mg" else fa.PRE_BLOCK_JACOBI)
    for rep in range(a.reps):
        X = op64.vector()
        inv = fa.InverseOperator(ctx, method=fa.CG, tolerance=1e-8, errormeasure=fa.RELATIVE, maxiterations=20000,
                                 residual_check=1)
        inv.bind(op64, pre)
        inv(F, X)
        info = inv.info()
        print(json.dumps({"solve_only": a.solve_only, "preconditioner": a.pre, "precision": a.prec, "rep": rep,
                          "iterations": info["iterations"], "time_s": round(info["timing"][0], 4),
                          "true_rel_residual": info["true_rel_residual"]}), flush=True)


def main():
    global GRAPH
    a = parse()
    if a.no_graph:
        GRAPH = 0
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (--gpus %d is not supported)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    import dunefem_amd as fa
    from dunefem_amd import workloads as wl
    if a.apply_only:
        apply_only(a, fa, wl)
        return
    if a.solve_only:
        solve_only(a, fa, wl)
        return
    if a.nonlinear_only:
        if not a.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as ora
            ora.select_native()
        print(json.dumps({"nonlinear": nonlinear_section(fa, wl, fa.Context(0), cpu=not a.no_cpu)}), flush=True)
        return
    if not a.no_cpu:
        # the CPU baseline runs the oracle built for this host (-march=native)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as ora
        ora.select_native()

    P = wl.problem1(n=a.n)
    n, p = P["n"], P["p"]
    k = p + 1
    E = n[0] * n[1] * n[2]
    N = E * k ** 3
    ctx = fa.Context(0)

    def make(prec):
        op = fa.GalerkinOperator(ctx, fa.OP_DG_SIPG_3D, n, p, P["beta"], P["b"], 0.0, fa.KAPPA_QP, precision=prec)
        op.set_kappa_qp(P["kvol"], P["kfaces"])
        return op

    op64 = make(fa.F64)
    U, W = op64.vector(P["u"]), op64.vector()
    wall, kern_ms = time_applies(fa, ctx, op64, U, W, a.steps, a.warmup)
    w64 = W.download()
    ms_per_step = wall / a.steps * 1e3
    gdofs = N / (wall / a.steps) / 1e9

    # algorithmic bytes: 8 (N in + N out) + 8 Q_c, Q_c = volume + face quadrature values
    qc = E * k ** 3 + sum((n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2)) for d in range(3)) * k * k
    b_alg = 8 * (2 * N + qc)
    achieved = b_alg / (kern_ms * 1e-3) / 1e9

    result = {
        "metric": "DG apply GDoF/s + time-to-solve at 1e-8 rel residual: mixed/learned vs FP64",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D SIPG DG Q%d advection-diffusion apply, %dx%dx%d hexes, %d DoFs, pointwise "
                               "kappa=clip(exp(16 Fourier modes),0.1,10) at volume+face Gauss points, "
                               "b=(1,0.5,0.25), upwind, beta=10p^2, u~N(0,1)" % (p, n[0], n[1], n[2], N),
                   "dofs": N, "order": p, "elements": E},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(p, n),
                     "kernel": "k_dg_apply<%d,double,false,true>" % k, "kernel_ms": round(kern_ms, 5),
                     "alg_bytes_per_launch": b_alg},
    }

    if a.calib:
        # y += a x on N doubles: reads 16 N bytes, writes 8 N bytes
        X, Y = op64.vector(P["u"]), op64.vector(P["u"])
        for _ in range(5):
            Y.axpy(0.5, X)
        ctx.synchronize()
        result["calibration"] = {"kernel": "k_axpy<double,double>", "read_bytes": 16 * N, "write_bytes": 8 * N}

    # ---- mixed precision applies --------------------------------------------------
    mixed = {}
    nrm = np.linalg.norm(w64)
    for name, prec in (() if a.no_mixed else (("fp32", fa.F32), ("bf16_mfma", fa.BF16MFMA))):
        try:
            op = make(prec)
        except fa.FemError as e:
            mixed[name] = {"error": str(e)}
            continue
        U2, W2 = op.vector(P["u"]), op.vector()
        try:
            wall2, kms2 = time_applies(fa, ctx, op, U2, W2, a.steps, a.warmup)
        except fa.FemError as e:
            mixed[name] = {"error": str(e)}
            continue
        w2 = W2.download()
        b2 = 4 * (2 * N + qc)
        mixed[name] = {"gdofs": round(N / (wall2 / a.steps) / 1e9, 3), "kernel_ms": round(kms2, 5),
                       "rel_l2_err_vs_fp64": float(np.linalg.norm(w2 - w64) / nrm),
                       "hbm_frac": round(b2 / (kms2 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        del op, U2, W2
    result["mixed"] = mixed

    if not a.no_solve:
        result["plain"] = plain_section(fa, wl, ctx, cpu=not a.no_cpu)

    # ---- time to solve (problem 3: 48^3 SIPG Q2, kappa = 10^(4H) channels) --------
    if not a.no_solve:
        result["solve"] = solve_section(fa, wl, ctx, learned=not a.no_learned)

    if not a.no_batch:
        result["batch"] = batch_section(fa, wl, ctx)

    if not a.no_combined and not a.no_solve:
        result["combined"] = combined_section(fa, wl, ctx, learned=not a.no_learned)

    if not a.no_nonlinear and not a.no_solve:
        result["nonlinear"] = nonlinear_section(fa, wl, ctx, cpu=not a.no_cpu)

    # ---- CPU baseline: the oracle (reference algorithm) on the host cores ----------
    if not a.no_cpu:
        import pyoracle as ora
        threads = cpu_threads()
        oop = ora.DGOperator(n, p, P["beta"], P["b"], 0.0, kappa=("qp", P["kvol"], tuple(P["kfaces"])))
        wc = oop.apply(P["u"], nthreads=threads)   # warm-up and parity spot check
        times = []
        for _ in range(a.cpu_reps):
            t0 = time.perf_counter()
            oop.apply(P["u"], nthreads=threads)
            times.append(time.perf_counter() - t0)
        tc = float(np.median(times))
        # the same workload by the sum-factorised CPU apply (labelled apart:
        # it is this library's algorithm on the host, not the reference's)
        ws = oop.apply_sf(P["u"], nthreads=threads)
        sf_times = []
        for _ in range(a.cpu_reps):
            t0 = time.perf_counter()
            oop.apply_sf(P["u"], nthreads=threads)
            sf_times.append(time.perf_counter() - t0)
        tsf = float(np.median(sf_times))
        result["cpu_baseline"] = {"value": round(N / tc / 1e9, 5), "unit": "GDoF/s", "cores": threads,
                                  "kind": "port",
                                  "sample": "full %d^3 Q%d apply (the same workload), median of %d after 1 warm-up; "
                                            "oracle = reference algorithm restated (dense per-element tables, "
                                            "skeleton once per face pair, OpenMP element chunks)" % (n[0], p, a.cpu_reps),
                                  "seconds_per_apply": round(tc, 4),
                                  "gpu_fp64_vs_oracle_rel_err": float(np.abs(w64 - wc).max() / np.abs(wc).max()),
                                  "host": host_info(ora),
                                  "sum_factorised_port": {
                                      "value": round(N / tsf / 1e9, 5), "unit": "GDoF/s", "cores": threads,
                                      "seconds_per_apply": round(tsf, 4),
                                      "note": "the same apply by 1D contractions per direction (owner-computes "
                                              "faces), on the host cores; a CPU port of the GPU algorithm, not "
                                              "the reference's dense-table evaluation",
                                      "rel_err_vs_dense": float(np.abs(ws - wc).max() / np.abs(wc).max())}}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
