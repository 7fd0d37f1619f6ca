"""Benchmark of the diffusion_2D hot path on one MI355X.

Workload (the GPU configuration the headline metric is quoted on): the
manufactured 2-D heat equation u_t = kx u_xx + ky u_yy + b(x,y,t) on [0,1]^2,
4096 x 4096 interior nodes, kx = 1, ky = 0.5, FP64, LSRKStep RKL2 with the
spectral radius from power iteration every 25 steps, rtol 1e-6, atol 1e-10;
fused stencil/stage kernels replayed as one HIP graph per step.  A "step" is
one adaptive integrator step (s RKL2 stages + the embedded error estimate,
plus the power iteration when it is due, plus any rejected attempts).  The
solve is first advanced, untimed, to t = 0.01 — past the start-up transient of
tiny steps into the stage-limited regime that makes up almost all of the solve
to t = 1 — then W warm-up steps and K timed steps follow.  Inputs are resident
in HBM before the timed region; the reported value counts every right-hand-side
evaluation (stages, error estimate, power-iteration J v products) times the
cells.

Extra objects on the JSON line:
  roofline      the dominant kernel (>97% of the step's HBM traffic): the
                multi-stage pass (K stages per launch, 48 algorithmic bytes per
                cell: Y_{j-1}, Y_{j-2}, Y_n, F_n read, the last two stages
                written) when the autotuner enables it, else the single fused
                stage (40 bytes per cell); mean launch time from HIP events on
                the integrator's stream
  cpu_baseline  the CPU oracle (the C restatement, -O3 with OpenMP loops)
                timed on the first adaptive steps of the same workload, on
                all the process's cores and (reported beside it) on one
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "cell-RHS evals/s; time-to-solution at fixed L2 err, batched/learned vs plain"
HBM_PEAK_GBS = 8000.0
N_SIDE = 4096
CPU_OMP_STEPS = 10  # adaptive steps of the CPU baseline's OpenMP leg
T_SETUP = 0.01


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def hip_sync():
    C.CDLL("libamdhip64.so").hipDeviceSynchronize()


def read_traffic(name, shape):
    """HBM bytes per launch of a kernel at one launch shape (K, rows, nt) from
    its committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated), or None when that shape was not measured."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    key = "traffic K=%d rows=%d nt=%d " % shape
    for line in open(path):
        if line.startswith(key):
            return float(line.split("=")[-1])
    return None


def make_integrator(pk, ctx, fast=2):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=N_SIDE, ny=N_SIDE, kx=1.0, ky=0.5)
    I = pk.Integrator(G, "rkl", G.initial())
    I.ARKodeSStolerances(1e-6, 1e-10)
    I.internal_eig(100, 0.01)
    I.LSRKStepSetDomEigFrequency(25)
    I.ARKodeSetMaxNumSteps(1000000)
    I.ARKodeSetStopTime(1.0)
    if fast:
        assert I.LSRKStepSetDiffusionFastPath(G.h, 1 if fast > 1 else 0) == 0
    return G, I


def roofline(pk, ctx, G):
    L = pk.lib()
    vs = [G.vector() for _ in range(5)]
    for k, v in enumerate(vs):
        L.N_VConst(1.0 + 0.1 * k, v.h)
    avg = C.c_double()
    rc = L.DiffusionProblem_TimeKernel(G.h, 1, vs[0].h, vs[1].h, vs[2].h, vs[3].h, vs[4].h, 50,
                                       C.byref(avg))
    assert rc == 0
    stage_ms = avg.value
    stage_gbs = 40.0 * G.n / (stage_ms * 1e-3) / 1e9
    rc = L.DiffusionProblem_TimeKernel(G.h, 0, vs[0].h, None, None, None, vs[4].h, 50, C.byref(avg))
    stencil_gbs = 16.0 * G.n / (avg.value * 1e-3) / 1e9
    K, rows, nt = G.get_config(2)
    if K >= 2:
        rc = L.DiffusionProblem_TimeKernel(G.h, 2, vs[0].h, vs[1].h, vs[2].h, vs[3].h, vs[4].h, 50,
                                           C.byref(avg))
        assert rc == 0
        launch_ms = avg.value
        bytes_alg = 48.0 * G.n
        kname = (f"k_pass<2,false,true,{K},{'true' if nt & 1 else 'false'},{'true' if nt & 2 else 'false'}> "
                 f"rows={rows} ({K} fused STS stages per launch)")
        traffic = read_traffic("stage_pass_4096_pmc.txt", (K, rows, nt))
    else:
        launch_ms = stage_ms
        bytes_alg = 40.0 * G.n
        v, r, n1 = G.get_config(1)
        kname = (f"k_stencil<2,false,true,true> rows={r}" if v == 0 else
                 f"k_march<2,false,true,true,{'true' if n1 else 'false'}> rows={r}") + " (fused STS stage)"
        traffic = read_traffic("fused_stage_4096_pmc.txt", (1, r, n1)) if v == 1 else None
    for v in vs:
        v.destroy()
    achieved = bytes_alg / (launch_ms * 1e-3) / 1e9
    return {"kernel": kname, "bound": "hbm",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch_algorithmic": bytes_alg, "launch_ms": round(launch_ms, 4),
            "stages_per_launch": max(K, 1),
            # what K single-stage launches would have to stream (40 B/cell each)
            # in the pass's time: the gain of temporal blocking, not HBM traffic
            "single_stage_equivalent_GBs": round(40.0 * G.n * max(K, 1) / (launch_ms * 1e-3) / 1e9, 1),
            "single_stage_fused_ms": round(stage_ms, 4), "single_stage_fused_GBs": round(stage_gbs, 1),
            "rhs_stencil_GBs": round(stencil_gbs, 1)}


def kernels_only(pk, reps, sweep, pin_pass=None):
    """Launch the hot-path kernels alone at the workload's size (the workload
    of the rocprofv3 counter runs): the autotuned multi-stage pass, fused
    STS stage and RHS stencil, and a plain copy (N_VScale by 1: 8 B read +
    8 B written per element) whose counters calibrate FETCH_SIZE/WRITE_SIZE.
    With sweep, first time every candidate launch shape of the kernels; with
    pin_pass (K, rows, nt), launch the pass at that shape instead of the
    autotuned one (counter runs at a fixed shape)."""
    ctx = pk.Context()
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=N_SIDE, ny=N_SIDE, kx=1.0, ky=0.5)
    L = pk.lib()
    vs = [G.vector() for _ in range(5)]
    for k, v in enumerate(vs):
        L.N_VConst(1.0 + 0.1 * k, v.h)
    res = {}
    avg = C.c_double()
    if sweep:
        cands = [(0, r, 0) for r in (1, 2, 4, 8)] + [(1, r, nt) for r in (4, 8, 16, 32, 64) for nt in (0, 1)]
        for (v, r, nt) in cands:
            G.set_config(0, v, r, nt)
            G.set_config(1, v, r, nt)
            L.DiffusionProblem_TimeKernel(G.h, 1, vs[0].h, vs[1].h, vs[2].h, vs[3].h, vs[4].h, reps,
                                          C.byref(avg))
            res[f"fused_v{v}_r{r}_nt{nt}_GBs"] = round(40.0 * G.n / (avg.value * 1e-3) / 1e9, 1)
            L.DiffusionProblem_TimeKernel(G.h, 0, vs[0].h, None, None, None, vs[4].h, reps, C.byref(avg))
            res[f"rhs_v{v}_r{r}_nt{nt}_GBs"] = round(16.0 * G.n / (avg.value * 1e-3) / 1e9, 1)
        for K in (2, 3, 4):
            for r in (8, 16, 32, 64, 128):
                for nt in (0, 1):
                    G.set_config(2, K, r, nt)
                    L.DiffusionProblem_TimeKernel(G.h, 2, vs[0].h, vs[1].h, vs[2].h, vs[3].h, vs[4].h,
                                                  reps, C.byref(avg))
                    res[f"pass_K{K}_r{r}_nt{nt}_us_per_stage"] = round(avg.value * 1e3 / K, 2)
    G.autotune(3)
    res["autotuned_rhs"] = G.get_config(0)
    res["autotuned_fused"] = G.get_config(1)
    L.DiffusionProblem_TimeKernel(G.h, 1, vs[0].h, vs[1].h, vs[2].h, vs[3].h, vs[4].h, reps, C.byref(avg))
    res["tuned_fused_GBs"] = round(40.0 * G.n / (avg.value * 1e-3) / 1e9, 1)
    L.DiffusionProblem_TimeKernel(G.h, 0, vs[0].h, None, None, None, vs[4].h, reps, C.byref(avg))
    res["tuned_rhs_GBs"] = round(16.0 * G.n / (avg.value * 1e-3) / 1e9, 1)
    res["autotuned_pass"] = G.get_config(2)
    if pin_pass:
        assert G.set_config(2, *pin_pass) == 0
        res["pinned_pass"] = G.get_config(2)
    if G.get_config(2)[0] >= 2:
        L.DiffusionProblem_TimeKernel(G.h, 2, vs[0].h, vs[1].h, vs[2].h, vs[3].h, vs[4].h, reps,
                                      C.byref(avg))
        res["tuned_pass_ms"] = round(avg.value, 4)
        res["tuned_pass_GBs"] = round(48.0 * G.n / (avg.value * 1e-3) / 1e9, 1)
        res["tuned_pass_us_per_stage"] = round(avg.value * 1e3 / G.get_config(2)[0], 2)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.N_VScale(1.0, vs[0].h, vs[4].h)
    hip_sync()
    res["copy_GBs"] = round(16.0 * G.n * reps / (time.perf_counter() - t0) / 1e9, 1)
    print(json.dumps(res))


def c4_problems(n=256, seed=20260101):
    """The batched configuration: square grids with side uniform over
    {64,128,256,512}, kx and ky log-uniform in [0.1, 10]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(rng.choice([64, 128, 256, 512]))
        kx, ky = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
        out.append(dict(kind=1, nx=side, ny=side, kx=float(kx), ky=float(ky)))
    return out


def batch_workload(pk, nprob, tout, batch_only=False, kpass=0, tile_rows=0, thin=None):
    """Batched RKC (one step group per step: every problem plans its attempt,
    then one launch per phase over all problems still in their stage loop)
    against the same problems solved one after another with the
    single-problem fused+graph path; absolute L2 errors against the exact
    solution (u_exact(0.5) = 0, so a relative error is undefined there).
    batch_only: the batched solve alone (a profiling workload)."""
    import numpy as np

    ctx = pk.Context()
    plist = c4_problems(nprob)
    B = pk.Batch(ctx, plist, "rkc")
    B.SetTolerances(1e-5, 1e-10)
    B.SetMaxNumSteps(1000000)
    if kpass:
        assert B.SetStagesPerPass(kpass) == 0
    if tile_rows:
        assert pk.lib().DiffusionBatch_SetTileRows(B.h, tile_rows) == 0
    if thin is not None:
        assert pk.lib().DiffusionBatch_SetThinPhases(B.h, thin[0], thin[1]) == 0
    hip_sync()
    log(f"batched: {nprob} problems, {sum(p['nx'] * p['ny'] for p in plist)} cells, "
        f"{kpass or 'default'} stages per pass")
    t0 = time.perf_counter()
    assert B.evolve(tout) == 0
    hip_sync()
    t_batch = time.perf_counter() - t0
    groups, launches = B.groups()
    log(f"batched done in {t_batch:.2f} s ({groups} step groups, {launches} phase launches)")
    if batch_only:
        nfe = sum(B.stats(k)["nfe"] for k in range(nprob))
        cells_rhs = sum(B.stats(k)["nfe"] * p["nx"] * p["ny"] for k, p in enumerate(plist))
        alg = B.algorithmic_bytes()
        print(json.dumps({"metric": "batched STS solve (256 uneven problems)", "n_problems": nprob, "tout": tout,
                          "batched_s": t_batch, "step_groups": groups, "phase_launches": launches,
                          "rhs_evals": nfe, "batched_cell_rhs_per_s": cells_rhs / t_batch,
                          "phase_kernels_algorithmic_bytes": alg,
                          "algorithmic_GBs_over_the_solve": alg / t_batch / 1e9}))
        B.release()
        return
    # one after another, each with the fused+graph single-problem path
    probs, integs = [], []
    for prm in plist:
        G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=prm["nx"], ny=prm["ny"], kx=prm["kx"], ky=prm["ky"])
        I = pk.Integrator(G, "rkc", G.initial())
        I.ARKodeSStolerances(1e-5, 1e-10)
        I.ARKodeSetMaxNumSteps(1000000)
        assert I.LSRKStepSetDiffusionFastPath(G.h, 1) == 0
        probs.append(G)
        integs.append(I)
    hip_sync()
    log("one after another")
    t0 = time.perf_counter()
    for I in integs:
        flag, t = I.evolve(tout)
        assert flag == 0
    hip_sync()
    t_seq = time.perf_counter() - t0
    errs, diffs, nfe_b, nfe_s, cells_rhs = [], [], 0, 0, 0
    for k, (G, I) in enumerate(zip(probs, integs)):
        yb = B.solution(k)
        ys = I.y.get()
        errs.append(float(np.sqrt(np.mean(yb ** 2))))
        diffs.append(float(np.abs(yb - ys).max() / max(np.abs(ys).max(), 1e-300)))
        nb, ns = B.stats(k)["nfe"], I.stats()["nfe"]
        nfe_b += nb
        nfe_s += ns
        cells_rhs += nb * G.n
    out = {"metric": "batched STS solves/s vs one after another (256 uneven problems)",
           "n_problems": nprob, "tout": tout, "method": "RKC", "rtol": 1e-5,
           "batched_s": t_batch, "sequential_s": t_seq,
           "batched_solves_per_s": nprob / t_batch, "sequential_solves_per_s": nprob / t_seq,
           "speedup": t_seq / t_batch, "step_groups": groups, "phase_launches": launches,
           "rhs_evals_batched": nfe_b, "rhs_evals_sequential": nfe_s,
           "batched_cell_rhs_per_s": cells_rhs / t_batch,
           "abs_l2_error_vs_exact_max": max(errs), "abs_l2_error_vs_exact_median": float(np.median(errs)),
           "max_rel_state_diff_batched_vs_sequential": max(diffs)}
    for I in integs:
        I.free()
    B.release()
    print(json.dumps(out))


def implicit_workload(pk, args):
    """The implicit path: mms-dirichlet 2048^2, kx = ky = 1, ARKStep DIRK3
    (ARK3(2)4L[2]SA), fixed dt = 1e-3, linear solves to relative residual
    1e-8: Jacobi-PCG against FGMRES with the learned preconditioner (Jacobi-PCG
    fallback at the iteration cap).  The networks (4 layers of 3x3
    convolutions, 16 channels, BF16: the symmetric K^T K with dilations
    1,3,9,27 and the ReLU network) are trained on the GPU first, on pairs from
    Jacobi-PCG solves with the same gamma*dt/dx^2, outside the timed solves.  A prefix of `--isteps` steps is timed for each method;
    errors are L2 against the exact solution and between the two runs."""
    import numpy as np

    ctx = pk.Context()
    side, dt, n_steps = args.iside, 1e-3, args.isteps
    dx = 1.0 / (side + 1)
    gamma_a = 1767732205903.0 / 4055673282236.0
    ratio = gamma_a * dt / dx ** 2
    log(f"implicit: {side}^2, dt {dt}, gamma*dt/dx^2 = {ratio:.1f}, {n_steps} steps per method")
    methods = args.imethods.split(",")
    if not args.no_untrained:
        methods.append("identity")
    dil = tuple(int(x) for x in args.dil.split(","))
    sdil = tuple(int(x) for x in args.sym_dil.split(","))
    training = {}

    def train(symmetric):
        """Pairs from batched Jacobi-PCG solves at relative residual 1e-12,
        then Adam on the GPU; the network starts at the identity (Jacobi)."""
        tn = args.sym_train_n if symmetric else args.train_n
        steps = args.sym_train_steps if symmetric else args.train_steps
        lr = args.sym_train_lr if symmetric else args.train_lr
        nsamp = args.sym_nsamp if symmetric else args.nsamp
        t0 = time.perf_counter()
        R, E, iters, rc = pk.training_set(ctx, nsamp, tn, ratio, reltol=1e-12, seed=1234)
        hip_sync()
        t_data = time.perf_counter() - t0
        assert rc == 0
        net = pk.LearnedPrec(ctx, sdil if symmetric else dil, symmetric=symmetric)
        net.init_identity(8e-4, 0)
        net.set_output_scale(1.0 / (1.0 + 4.0 * ratio))
        t0 = time.perf_counter()
        hist = net.train(nsamp, tn, ratio, R, E, steps, args.train_batch, lr, 0)
        hip_sync()
        t_train = time.perf_counter() - t0
        tag = "symmetric" if symmetric else "relu"
        log(f"{tag} network: training set {nsamp} pairs of {tn}^2 (Jacobi-PCG <= {iters} its) "
            f"{t_data:.1f} s; {steps} Adam steps {t_train:.1f} s; loss {hist[:20].mean():.3f} -> "
            f"{hist[-50:].mean():.3f}")
        training[tag] = {"pairs": nsamp, "grid": tn, "data_s": t_data, "train_s": t_train,
                         "steps": steps, "batch": args.train_batch, "lr": lr,
                         "dilations": list(sdil if symmetric else dil),
                         "loss_first20": float(hist[:20].mean()), "loss_last50": float(hist[-50:].mean())}
        R.destroy()
        E.destroy()
        return net

    net = train(False) if any(m in ("learned", "learned-fcg") for m in methods) else None
    snet = train(True) if any(m.startswith("learned-sym") for m in methods) else None
    ident = pk.LearnedPrec(ctx, dil)
    ident.init_identity(0.0, 0)

    def run(kind):
        """kind: 'jacobi' (Jacobi-PCG), 'mg' (PCG + one multigrid V-cycle),
        'learned' (FGMRES + network), 'learned-fcg' (flexible PCG + network),
        'learned-sym' (FGMRES + symmetric network), 'learned-sym-pcg' (PCG +
        symmetric network), 'identity' (FGMRES + untrained network)."""
        G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=side, ny=side, kx=1.0, ky=1.0)
        if kind not in ("jacobi", "mg"):
            {"identity": ident, "learned-sym": snet, "learned-sym-pcg": snet}.get(kind, net).attach(G)
        ls = {"jacobi": "cg", "mg": "cg", "learned-fcg": "fcg", "learned-sym-pcg": "cg"}.get(kind, "fgmres")
        prec = {"jacobi": True, "mg": "mg"}.get(kind, "learned")
        I = pk.Integrator(G, "dirk", G.initial(), ls=ls, maxl=20000 if ls == "cg" else args.imaxl,
                          precond=prec)
        if ls == "fgmres":
            pk.lib().SUNLinSol_SPFGMRSetMaxRestartDim(I.LS, args.sym_mrestart if kind == "learned-sym" else args.mrestart)
        if args.jtimes:
            I.analytic_jtimes()
            if args.fused_pcg:
                I.fused_pcg()
        I.ARKodeSetOrder(3)
        I.ARKodeSetFixedStep(dt)
        I.ARKodeSetLSRelativeTolerance(1e-8)
        I.ARKodeSetMaxNumSteps(1000000)
        flag, t = I.evolve(dt)  # first step untimed (setup, allocation)
        hip_sync()
        s0 = I.stats()
        t0 = time.perf_counter()
        for k in range(n_steps):
            flag, t = I.evolve(dt * (k + 2))
            assert flag == 0, flag
            st = I.stats()
            log(f"  {kind} step {k + 1}: {time.perf_counter() - t0:.2f} s, Krylov its {st['nli'] - s0['nli']}, "
                f"fallbacks {I.fallbacks()}")
        hip_sync()
        el = time.perf_counter() - t0
        s1 = I.stats()
        y = I.y.get()
        ex = G.exact(t).get()
        out = dict(s_per_step=el / n_steps, lin_iters=(s1["nli"] - s0["nli"]) / n_steps,
                   newton_iters=(s1["nni"] - s0["nni"]) / n_steps, fallbacks=I.fallbacks(),
                   l2_err_vs_exact=float(np.sqrt(np.mean((y - ex) ** 2))), t=t)
        I.free()
        G.destroy()
        log(f"{kind}: {out}")
        return out, y

    res, ys = {}, {}
    for kind in methods:
        res[kind], ys[kind] = run(kind)
    names = {"jacobi": "jacobi_pcg", "mg": "multigrid_pcg", "learned": "learned_fgmres",
             "learned-fcg": "learned_flexible_pcg", "identity": "untrained_fgmres",
             "learned-sym": "learned_symmetric_fgmres", "learned-sym-pcg": "learned_symmetric_pcg"}
    out = {"metric": "time-to-solution at fixed L2 err, learned-preconditioned FGMRES vs Jacobi-PCG",
           "config": f"mms-dirichlet {side}x{side} kx=ky=1 FP64, DIRK3 fixed dt=1e-3, LS reltol 1e-8",
           "steps_timed": n_steps, "gamma_dt_over_dx2": ratio,
           "jtimes": "analytic (stencil)" if args.jtimes else "difference quotient (as the reference)"}
    for kind in methods:
        out[names[kind]] = res[kind]
    if "jacobi" in res:
        yj = ys["jacobi"]
        for kind in methods:
            if kind == "jacobi":
                continue
            tag = names[kind].replace("_pcg", "").replace("_fgmres", "")
            out[f"speedup_{names[kind]}_vs_jacobi_pcg"] = res["jacobi"]["s_per_step"] / res[kind]["s_per_step"]
            out[f"rel_l2_diff_{names[kind]}_vs_jacobi_pcg"] = float(
                np.sqrt(np.mean((ys[kind] - yj) ** 2)) / np.sqrt(np.mean(yj ** 2)))
    for kind in ("learned", "learned-sym"):
        if kind in res and "identity" in res:
            out[f"krylov_iters_ratio_{names[kind]}_vs_untrained"] = res[kind]["lin_iters"] / res["identity"]["lin_iters"]
    if "jacobi" in res:
        # time to solution over the timed steps with the network's training
        # (pairs + Adam) charged to the learned method, and the step count
        # from which training pays off
        tj = res["jacobi"]["s_per_step"]
        tts = {"steps": n_steps, "t_span": n_steps * dt, "jacobi_pcg_s": tj * n_steps,
               "note": "the first step (setup) is untimed for every method"}
        for kind, tag in (("learned-sym", "symmetric"), ("learned-sym-pcg", "symmetric"), ("learned", "relu"),
                          ("learned-fcg", "relu")):
            if kind not in res or tag not in training:
                continue
            tr = training[tag]["data_s"] + training[tag]["train_s"]
            tl = res[kind]["s_per_step"]
            tts[f"{names[kind]}_s_incl_training"] = tl * n_steps + tr
            tts[f"speedup_{names[kind]}_incl_training"] = tj * n_steps / (tl * n_steps + tr)
            tts[f"break_even_steps_{names[kind]}"] = tr / (tj - tl) if tj > tl else None
            # the whole solve to t = 1 (1000 steps) from the measured per-step times
            tts[f"projected_speedup_{names[kind]}_incl_training_to_t1"] = 1000 * tj / (1000 * tl + tr)
        out["time_to_solution"] = tts
    out["training"] = training
    out["fgmres_restart"] = args.mrestart
    out["fgmres_restart_symmetric"] = args.sym_mrestart
    for nt in (net, snet, ident):
        if nt is not None:
            nt.release()
    print(json.dumps(out))


def learned_apply_workload(pk, args):
    """The symmetric learned network (dilations --sym-dil, random weights:
    the cost does not depend on them) applied --apply-reps times to one
    2048^2 FP64 image: the network's kernels alone, for timing and counters."""
    import numpy as np

    ctx = pk.Context()
    n = args.iside
    rng = np.random.default_rng(0)
    R, Z = pk.Vector(ctx, n=n * n), pk.Vector(ctx, n=n * n)
    R.set(rng.standard_normal(n * n))
    dil = tuple(int(x) for x in args.sym_dil.split(","))
    net = pk.LearnedPrec(ctx, dil, symmetric=True)
    net.set_params((0.1 * rng.standard_normal(net.nparams)).astype(np.float32))
    net.apply(1, n, n, R, Z)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.apply_reps):
        net.apply(1, n, n, R, Z)
    hip_sync()
    ms = (time.perf_counter() - t0) / args.apply_reps * 1e3
    out = {"metric": "symmetric learned network application time", "value": ms, "unit": "ms",
           "config": {"image": f"{n}x{n} FP64", "dilations": list(dil), "reps": args.apply_reps,
                      "conv_kernel": {0: "direct", 1: "LDS row chains",
                                      2: "LDS row chains, two layers per kernel"}[args.conv_kernel],
                      "pair_rows": args.pair_rows, "pair_ahead": args.pair_ahead}}
    ctx.close()
    print(json.dumps(out))


def c5_problems(n=64, seed=20260102):
    """The batched implicit configuration: square grids with side uniform in
    [256, 1024], kx and ky log-uniform in [0.1, 10]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(rng.integers(256, 1025))
        kx, ky = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
        out.append(dict(nx=side, ny=side, kx=float(kx), ky=float(ky)))
    return out


def c5_workload(pk, args):
    """Many implicit problems at once: 64 manufactured problems of different
    sizes and coefficients, each ARKStep DIRK3 with fixed dt = 1e-3 and linear
    solves to relative residual 1e-8 (equal accuracy for every method),
    integrated to t = steps*dt.  The problems run as independent integrators
    driven concurrently by `--c5-workers` host threads, each with its own
    SUNContext (HIP stream); a worker integrates its problems one after
    another.  Methods: Jacobi-PCG (the reference's configuration), PCG and
    FGMRES with learned symmetric networks (one per class of problem
    anisotropy and magnitude, --c5-classes x --c5-mag-classes, or with
    --c5-classes 0 one isotropic network at the batch's median
    gamma*dt*k/dx^2; trained on the GPU outside the timed region; the output
    scaled per problem by its inverse Jacobi diagonal) and PCG with the
    multigrid V-cycle.  Time-to-solution is the wall time of the whole batch,
    problem setup included (the L2 check against the exact solution and the
    teardown follow the timed region; the workers' contexts and the trained
    networks precede it); Jacobi-PCG is also run by a single worker."""
    import threading
    import numpy as np

    probs = c5_problems(args.c5_n, 20260102)
    dt, steps = 1e-3, args.c5_steps
    gamma_a = 1767732205903.0 / 4055673282236.0
    ratios = [gamma_a * dt * 0.5 * (p["kx"] + p["ky"]) * (p["nx"] + 1) ** 2 for p in probs]
    cells = sum(p["nx"] * p["ny"] for p in probs)
    log(f"c5: {len(probs)} problems, {cells} cells, sides {min(p['nx'] for p in probs)}..."
        f"{max(p['nx'] for p in probs)}, gamma*dt*k/dx^2 {min(ratios):.0f}..{max(ratios):.0f}, {steps} steps")
    methods = args.c5_methods.split(",")
    t_end = dt * steps  # every method integrates each problem to the same time
    sdil = tuple(int(x) for x in args.sym_dil.split(","))
    # network shape per learned method: (dilations, symmetric); learned-sym-pcg
    # is PCG with the symmetric network (a fixed SPD preconditioner)
    shapes = {"learned": ((1, 1, 1, 1), False), "learned-sym": (sdil, True), "learned-sym-pcg": (sdil, True)}
    # training operators: with --c5-classes K the problems are split by their
    # anisotropy r_x/r_y (r = gamma*dt*k/dx^2 along each axis) into K classes of
    # equal count (each split again by r_x r_y into --c5-mag-classes classes)
    # and one network is trained per class, on the class's geometric-mean
    # (r_x, r_y); otherwise one isotropic network for all
    rxy = [(gamma_a * dt * p["kx"] * (p["nx"] + 1) ** 2, gamma_a * dt * p["ky"] * (p["ny"] + 1) ** 2)
           for p in probs]
    if args.c5_classes > 0:
        order = sorted(range(len(probs)), key=lambda k: rxy[k][0] / rxy[k][1])
        K = min(args.c5_classes, len(probs))
        groups = [order[c * len(order) // K:(c + 1) * len(order) // K] for c in range(K)]
        if args.c5_mag_classes > 1:  # each anisotropy class split by r_x r_y into equal-count classes
            sub = []
            for g in groups:
                g = sorted(g, key=lambda k: rxy[k][0] * rxy[k][1])
                m = min(args.c5_mag_classes, len(g))
                sub += [g[c * len(g) // m:(c + 1) * len(g) // m] for c in range(m)]
            groups = sub
        cls_of = {k: c for c, g in enumerate(groups) for k in g}
        targets = [tuple(float(np.exp(np.mean([np.log(rxy[k][a]) for k in g]))) for a in (0, 1)) for g in groups]
    else:
        ratio = args.c5_train_ratio if args.c5_train_ratio > 0 else float(np.median(ratios))
        cls_of = {k: 0 for k in range(len(probs))}
        targets = [(ratio, ratio)]
    shared = {}
    train = {}
    saved = {}
    if args.c5_nets and os.path.exists(args.c5_nets):
        with np.load(args.c5_nets) as z:
            saved = {k: z[k] for k in z.files}
    for kind in methods:
        if kind not in shapes or (kind == "learned-sym-pcg" and "learned-sym" in shared):
            continue
        key = "sym" if shapes[kind][1] else "relu"
        if key + "_targets" in saved and np.allclose(saved[key + "_targets"], targets, rtol=1e-12, atol=0):
            shared[kind] = list(saved[key + "_params"])
            train[kind] = [{"loaded_from": args.c5_nets}]
            log(f"{kind}: {len(targets)} trained network(s) loaded from {args.c5_nets}")
            if kind in ("learned-sym", "learned-sym-pcg"):
                shared["learned-sym"] = shared["learned-sym-pcg"] = shared[kind]
            continue
        dil, sym = shapes[kind]
        tn = args.sym_train_n if sym else args.train_n
        train_steps = args.sym_train_steps if sym else args.train_steps
        lr = args.sym_train_lr if sym else args.train_lr
        nsamp = args.sym_nsamp if sym else args.nsamp
        shared[kind], train[kind] = [], []
        for (rx, ry) in targets:
            ctx = pk.Context()
            t0 = time.perf_counter()
            R, E, iters, rc = pk.training_set(ctx, nsamp, tn, rx, reltol=1e-12, seed=1234, ratio_y=ry)
            assert rc == 0
            net = pk.LearnedPrec(ctx, dil, symmetric=sym)
            net.init_identity(8e-4, 0)
            net.set_output_scale(1.0 / (1.0 + 2.0 * rx + 2.0 * ry))
            hist = net.train(nsamp, tn, rx, R, E, train_steps, args.train_batch, lr, 0, ratio_y=ry)
            hip_sync()
            shared[kind].append(net.params())
            rec = dict(ratio_x=rx, ratio_y=ry, pairs=nsamp, grid=tn, steps=train_steps, lr=lr,
                       dilations=list(dil), symmetric=sym, seconds=time.perf_counter() - t0,
                       loss_first20=float(hist[:20].mean()), loss_last50=float(hist[-50:].mean()))
            if len(targets) > 1:
                rec["problems"] = sum(1 for c in cls_of.values() if c == len(train[kind]))
            train[kind].append(rec)
            log(f"{kind} network trained at ratios ({rx:.0f}, {ry:.0f}): {rec}")
            ctx.close()
        if kind in ("learned-sym", "learned-sym-pcg"):  # one symmetric network for both
            shared["learned-sym"] = shared["learned-sym-pcg"] = shared[kind]
        if args.c5_nets:
            saved[key + "_targets"] = np.array(targets)
            saved[key + "_params"] = np.array(shared[kind])
            np.savez(args.c5_nets, **saved)

    def solve_one(ctx, net, k, kind, t_end):
        p = probs[k]
        t0 = time.perf_counter()
        G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=p["nx"], ny=p["ny"], kx=p["kx"], ky=p["ky"])
        ls, prec, maxl = "cg", True, 20000
        if kind == "mg":
            prec = "mg"
        elif kind in shapes:
            dx2 = 1.0 / (p["nx"] + 1) ** 2
            net.set_output_scale(1.0 / (1.0 + 2.0 * gamma_a * dt * (p["kx"] + p["ky"]) / dx2))
            net.attach(G)
            ls, prec, maxl = ("cg" if kind == "learned-sym-pcg" else "fgmres"), "learned", args.imaxl
        I = pk.Integrator(G, "dirk", G.initial(), ls=ls, maxl=maxl, precond=prec)
        if ls == "fgmres":
            pk.lib().SUNLinSol_SPFGMRSetMaxRestartDim(I.LS, args.sym_mrestart if kind == "learned-sym" else args.mrestart)
        I.analytic_jtimes()
        if args.fused_pcg:
            I.fused_pcg()
            if args.pcg_bulk != 0.75:
                assert pk.lib().ARKodeSetFusedPCGBulk(I.h, args.pcg_bulk) == 0
        I.ARKodeSetOrder(3)
        I.ARKodeSetFixedStep(dt)
        I.ARKodeSetLSRelativeTolerance(1e-8)
        I.ARKodeSetMaxNumSteps(1000000)
        t1 = time.perf_counter()
        flag, t = I.evolve(t_end)
        assert flag == 0 and t == t_end, (kind, k, flag, t, t_end)
        # host seconds of this worker in the problem's setup and integration
        return G, I, t, (t1 - t0, time.perf_counter() - t1)

    def check_one(G, I, t, phase):
        """After the timed region: statistics, L2 error against the exact
        solution, teardown."""
        t0 = time.perf_counter()
        st = I.stats()
        y = I.y.get()
        ex = G.exact(t).get()
        rec = dict(lin_iters=int(st["nli"]), fallbacks=I.fallbacks(),
                   l2_err=float(np.sqrt(np.mean((y - ex) ** 2))))
        I.free()
        G.destroy()
        rec["phase_s"] = phase + (time.perf_counter() - t0,)
        return rec

    def run(kind, workers):
        todo = list(range(len(probs)))
        lock = threading.Lock()
        recs = [None] * len(probs)
        errors = []
        solved_count = [0]

        # the timed region runs from every worker's start to the last solve's
        # end: each worker's context and trained networks exist before it
        # (the networks are trained before any batch), the L2 check and the
        # teardown follow it
        start = threading.Barrier(workers + 1)
        solved = threading.Barrier(workers + 1)

        def worker():
            ctx, nets, live = None, [], []
            try:
                ctx = pk.Context()
                if kind in shapes:
                    for params in shared[kind]:
                        nets.append(pk.LearnedPrec(ctx, shapes[kind][0], symmetric=shapes[kind][1]))
                        nets[-1].set_params(params)
            except Exception as e:  # reported after the join
                with lock:
                    errors.append(repr(e))
            start.wait()
            try:
                while True:
                    with lock:
                        if not todo or errors:
                            break
                        k = todo.pop(0)
                    live.append((k, solve_one(ctx, nets[cls_of[k]] if nets else None, k, kind, t_end)))
                    with lock:
                        solved_count[0] += 1
                        done = solved_count[0]
                    if done % 8 == 0 or workers == 1:
                        log(f"  {kind} x{workers}: {done}/{len(probs)} problems solved")
            except Exception as e:
                with lock:
                    errors.append(repr(e))
            solved.wait()
            try:
                for k, res_k in live:
                    recs[k] = check_one(*res_k)
            except Exception as e:
                with lock:
                    errors.append(repr(e))
            finally:
                if ctx is not None:
                    ctx.close()

        th = [threading.Thread(target=worker) for _ in range(workers)]
        for x in th:
            x.start()
        hip_sync()
        start.wait()
        t0 = time.perf_counter()
        solved.wait()
        hip_sync()
        el = time.perf_counter() - t0
        for x in th:
            x.join()
        assert not errors, errors
        errs = np.array([r["l2_err"] for r in recs])
        its = np.array([r["lin_iters"] for r in recs])
        out = dict(seconds=el, workers=workers, solves_per_s=len(probs) / el,
                   krylov_iters_total=int(its.sum()), krylov_iters_max=int(its.max()),
                   fallbacks=int(sum(r["fallbacks"] for r in recs)),
                   l2_err_max=float(errs.max()), l2_err_median=float(np.median(errs)),
                   krylov_iters_per_problem=[int(x) for x in its],
                   worker_s_setup_evolve_check=[round(float(sum(r["phase_s"][i] for r in recs)), 3)
                                                for i in range(3)],
                   timed="problem setup and integration of every problem; the L2 check and teardown after")
        log(f"{kind} x{workers}: { {k: v for k, v in out.items() if k != 'krylov_iters_per_problem'} }")
        return out, errs

    res = {}
    errs = {}
    for kind in methods:
        res[kind], errs[kind] = run(kind, args.c5_workers)
    if "jacobi" in res and args.c5_workers != 1:
        res["jacobi_one_worker"], _ = run("jacobi", 1)
    out = {"metric": "time-to-solution at fixed L2 err, batched implicit: learned preconditioners (PCG, "
                     "FGMRES) and multigrid (PCG) vs Jacobi-PCG",
           "config": {"workload": f"{len(probs)} mms-dirichlet problems, side U[256,1024], kx,ky logU[0.1,10] "
                                  f"(seed 20260102), {cells} cells, DIRK3 fixed dt=1e-3, {steps} steps, "
                                  "LS reltol 1e-8, analytic J v",
                      "learned_networks": (f"{len(targets)}, one per class of anisotropy"
                                           + (" and magnitude" if args.c5_mag_classes > 1 else "")
                                           if args.c5_classes > 0
                                           else "1, isotropic, shared by all problems"),
                      "concurrency": f"{args.c5_workers} host threads, one SUNContext/stream each"},
           "results": res, "training": train,
           "problems": [[p["nx"], round(p["kx"], 4), round(p["ky"], 4)] for p in probs]}
    if "jacobi" in res:
        for kind in res:
            if kind != "jacobi":
                out[f"speedup_{kind}_vs_jacobi"] = res["jacobi"]["seconds"] / res[kind]["seconds"]
        for kind in errs:
            if kind != "jacobi":
                out[f"max_rel_l2_err_change_{kind}_vs_jacobi"] = float(
                    np.max(np.abs(errs[kind] - errs["jacobi"]) / errs["jacobi"]))
    print(json.dumps(out))


def cpu_threads():
    """Host threads for the OpenMP leg: the process's CPU share (OMP_NUM_THREADS
    when set, as on the GPU hosts), else the cores it may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline():
    """The CPU oracle (this repository's C restatement of the reference path,
    built -O3 -march=x86-64-v3 with OpenMP loops) timed on the same workload
    on this host: serially on one core, the first adaptive step (initial-step
    estimate, one power iteration, the step's stages and error estimate); and
    with OpenMP on every core of the process's share, the first ten steps
    (their stage counts still ramp up: about 34 RHS evaluations per step
    against 195 in the timed GPU steps, so the per-step work outside the
    stages weighs more on the CPU rate).  `value` is the OpenMP rate; the
    serial leg is reported beside it."""
    import oracle_lib as O

    def leg(threads, steps):
        L = O.lib(O.TIMING_LIB_PATH)
        got = L.orc_set_threads(threads)
        P = O.OracleProblem(O.MMS, lib_path=O.TIMING_LIB_PATH, nx=N_SIDE, ny=N_SIDE, kx=1.0, ky=0.5)
        R = O.OracleIntegrator(P, O.RKL, P.initial())
        R.set_tolerances(1e-6, 1e-10)
        R.set_internal_eig(1, 100, 0.01)
        R.set_dom_eig_frequency(25)
        R.set_max_num_steps(steps)
        t0 = time.perf_counter()
        R.evolve(1.0)
        dt = time.perf_counter() - t0
        st = R.stats()
        return got, st, dt, st["nfe"] * P.n / dt

    nthr = cpu_threads()
    _, st1, dt1, v1 = leg(1, 1)
    got, stn, dtn, vn = leg(nthr, CPU_OMP_STEPS)
    log(f"cpu baseline: serial {v1:.3e}, {got} threads {vn:.3e} cell-RHS evals/s")
    return {"value": vn, "unit": "cell-RHS evals/s", "cores": got, "kind": "port",
            "sample": f"first {CPU_OMP_STEPS} adaptive steps of the 4096^2 RKL2 workload (initial-step estimate, power "
                      f"iteration, {stn['nst']} steps): {stn['nfe']} RHS evals in {dtn:.1f} s on {got} threads",
            "build": "oracle/ (C restatement), gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off",
            "host_cpus": os.cpu_count(), "process_cpus": len(os.sched_getaffinity(0)),
            "serial": {"value": v1, "cores": 1,
                       "sample": f"first adaptive step: {st1['nfe']} RHS evals in {dt1:.1f} s"}}


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--fastpath", type=int, default=2, help="0 API path, 1 fused, 2 fused+graphs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--apply-reps", type=int, default=10, help="learned-apply: network applications")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain implicit comparison after the headline")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-vs-sequential comparison after the headline")
    ap.add_argument("--kernels-only", type=int, default=0, metavar="REPS",
                    help="only launch the hot-path kernels REPS times each (profiling workload)")
    ap.add_argument("--sweep", action="store_true", help="with --kernels-only: time every launch shape")
    ap.add_argument("--pin-pass", default=None, metavar="K,ROWS,NT",
                    help="with --kernels-only: launch the multi-stage pass at this shape")
    ap.add_argument("--workload", default="sts", choices=["sts", "batch", "implicit", "c5", "learned-apply"],
                    help="sts: the headline 4096^2 RKL2 line; batch: the 256-problem batched "
                         "comparison; implicit: learned-preconditioned FGMRES vs Jacobi-PCG")
    ap.add_argument("--iside", type=int, default=2048)
    ap.add_argument("--isteps", type=int, default=3)
    ap.add_argument("--nsamp", type=int, default=20000)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-n", type=int, default=64, help="side of the training grids")
    ap.add_argument("--train-batch", type=int, default=16)
    ap.add_argument("--train-lr", type=float, default=3e-4)
    ap.add_argument("--no-untrained", action="store_true", help="skip the untrained-network FGMRES run")
    ap.add_argument("--jtimes", action="store_true", help="analytic J v (stencil) instead of the difference quotient")
    ap.add_argument("--mrestart", type=int, default=30)
    ap.add_argument("--dil", default="1,1,1,1", help="dilations of the network's 3x3 layers")
    ap.add_argument("--fused-pcg", action="store_true",
                    help="implicit/c5: PCG in two fused kernels per iteration with Jacobi (with the analytic J v)")
    ap.add_argument("--sym-dil", default="1,3,9,27", help="dilations of the symmetric network's layers")
    ap.add_argument("--sym-train-n", type=int, default=192, help="training grid side, symmetric network")
    ap.add_argument("--sym-nsamp", type=int, default=20000,
                    help="training pairs of the symmetric network (20000, the implicit configuration's count; "
                         "4096 gave the same 63 Krylov iterations per 2048^2 DIRK3 step)")
    ap.add_argument("--sym-train-steps", type=int, default=1000)
    ap.add_argument("--sym-train-lr", type=float, default=3e-3)
    ap.add_argument("--sym-mrestart", type=int, default=20, help="FGMRES restart with the symmetric network")
    ap.add_argument("--imethods", default="jacobi,mg,learned-sym,learned-sym-pcg,learned",
                    help="implicit: methods to run (jacobi first)")
    ap.add_argument("--imaxl", type=int, default=3000, help="FGMRES iteration cap before the PCG fallback")
    ap.add_argument("--nprob", type=int, default=256)
    ap.add_argument("--c5-n", type=int, default=64, help="c5: number of problems")
    ap.add_argument("--c5-steps", type=int, default=3, help="c5: DIRK3 steps per problem")
    ap.add_argument("--c5-workers", type=int, default=8, help="c5: concurrent host threads")
    ap.add_argument("--c5-methods", default="jacobi,mg,learned-sym-pcg,learned-sym",
                    help="c5: methods, from jacobi, mg, learned-sym-pcg (PCG + symmetric network), learned-sym "
                         "(FGMRES + symmetric network), learned (FGMRES + ReLU network)")
    ap.add_argument("--c5-train-ratio", type=float, default=0.0,
                    help="c5: gamma*dt*k/dx^2 the shared network is trained at (0: the batch median)")
    ap.add_argument("--pcg-bulk", type=float, default=0.75,
                    help="c5: fraction of the previous solve's iterations the fused PCG runs before testing "
                         "the residual before each preconditioner call")
    ap.add_argument("--c5-nets", default="",
                    help="c5: .npz file of trained networks: loaded when it holds networks for the same training "
                         "operators, else written after training")
    ap.add_argument("--c5-mag-classes", type=int, default=3,
                    help="c5: with --c5-classes, split each anisotropy class into this many classes of r_x r_y")
    ap.add_argument("--c5-classes", type=int, default=4,
                    help="c5: learned methods train one network per class of problem anisotropy (this many "
                         "classes of equal count); 0: one isotropic network shared by every problem")
    ap.add_argument("--tout", type=float, default=0.5)
    ap.add_argument("--conv-kernel", type=int, default=1, choices=[0, 1, 2],
                    help="learned network's convolutions: 1 LDS row chains, 2 LDS row chains and two layers per "
                         "kernel in the symmetric network's application, 0 the direct kernel")
    ap.add_argument("--pair-rows", type=int, default=16, help="chain rows per workgroup of the two-layer kernel")
    ap.add_argument("--chain-tasks", type=int, default=1024,
                    help="row-chain conv kernels: shorter row segments per workgroup on images that would launch "
                         "fewer workgroups than this (1: 16 rows everywhere)")
    ap.add_argument("--pair-ahead", type=int, default=2, help="input rows in flight per workgroup, two-layer kernel")
    ap.add_argument("--kpass", type=int, default=0, help="batch: stages per pass (0: the library default)")
    ap.add_argument("--tile-rows", type=int, default=0, help="batch: rows per wavefront tile (0: autotuned)")
    ap.add_argument("--thin-phases", default="",
                    help="batch: ROWS,MIN_TILES of the fine tiles of thin phases (ROWS 0: never; default 12,2048)")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch: the batched solve alone, without the one-after-another comparison")
    return ap.parse_args(argv)


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (every box here has one MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    from hiplib import package

    pk = package()
    if args.conv_kernel != 1:
        assert pk.lib().LearnedPrec_SetConvKernel(args.conv_kernel) == 0
    if (args.pair_rows, args.pair_ahead) != (16, 2):
        assert pk.lib().LearnedPrec_SetPairTuning(args.pair_rows, args.pair_ahead) == 0
    if args.chain_tasks != 1024:
        assert pk.lib().LearnedPrec_SetChainTasks(args.chain_tasks) == 0
    if args.kernels_only:
        pin = tuple(int(x) for x in args.pin_pass.split(",")) if args.pin_pass else None
        kernels_only(pk, args.kernels_only, args.sweep, pin)
        return
    if args.workload == "batch":
        thin = tuple(int(v) for v in args.thin_phases.split(",")) if args.thin_phases else None
        batch_workload(pk, args.nprob, args.tout, args.batch_only, args.kpass, args.tile_rows, thin)
        return
    if args.workload == "implicit":
        implicit_workload(pk, args)
        return
    if args.workload == "c5":
        c5_workload(pk, args)
        return
    if args.workload == "learned-apply":
        learned_apply_workload(pk, args)
        return
    ctx = pk.Context()
    G, I = make_integrator(pk, ctx, args.fastpath)
    log(f"workload ready: {G.nx}x{G.ny} cells; advancing to t = {T_SETUP} (untimed)")
    flag, t = I.evolve(T_SETUP)
    assert flag >= 0, flag
    log(f"setup done at t = {t}, {I.stats()['nst']} steps; warmup {args.warmup} steps")
    for _ in range(args.warmup):
        flag, t = I.evolve(1.0, pk.ARK_ONE_STEP)
        assert flag >= 0, flag
    hip_sync()
    s0 = I.stats()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        flag, t = I.evolve(1.0, pk.ARK_ONE_STEP)
        assert flag >= 0, flag
    hip_sync()
    dt = time.perf_counter() - t0
    s1 = I.stats()
    nfe = s1["nfe"] - s0["nfe"]
    value = nfe * G.n / dt
    log(f"{nfe} RHS evals in {dt:.3f} s; stages/step {nfe / args.steps:.1f}; t = {t:.4e}")
    rl = roofline(pk, ctx, G)
    log(f"dominant kernel {rl['kernel']}: {rl['launch_ms']} ms -> {rl['achieved']} GB/s")
    I.free()
    G.destroy()
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        cpu = cpu_baseline()
    out = {
        "metric": METRIC, "value": value, "unit": "cell-RHS evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (manufactured-solution initial state, no dataset)",
        "config": {"workload": "mms-dirichlet 4096x4096 interior FP64, kx=1 ky=0.5, LSRKStep RKL2, "
                               "power-iteration spectral radius every 25 steps, rtol=1e-6 atol=1e-10",
                   "path": ["api", "fused", "fused+hipgraph"][args.fastpath],
                   "cells": G.n, "rhs_evals_timed": nfe, "stages_per_step": nfe / args.steps,
                   "t_start_timed": T_SETUP, "attempts_per_step": (s1["attempts"] - s0["attempts"]) / args.steps,
                   "error_test_fails_timed": s1["netf"] - s0["netf"]},
        "roofline": rl, "cpu_baseline": cpu,
    }
    if not args.no_learned:
        out["learned_vs_plain"] = learned_vs_plain()
    if not args.no_batch:
        out["batched_vs_sequential"] = batched_vs_sequential()
    print(json.dumps(out))


def child_json(extra, keep, limit=300):
    """One JSON-printing run of this script as a child process with its own
    time limit (after the headline measurement, so nothing there can hold
    up the headline line); returns the kept keys, or the error."""
    import subprocess

    cmd = [sys.executable, os.path.abspath(__file__)] + extra
    log("child run: " + " ".join(cmd[1:]))
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=limit)
    except subprocess.TimeoutExpired:
        return {"error": f"time limit ({limit} s) reached"}
    if r.returncode != 0:
        return {"error": f"exit status {r.returncode}", "stderr_tail": r.stderr[-400:]}
    try:
        d = json.loads(r.stdout.strip().splitlines()[-1])
    except (IndexError, ValueError) as e:
        return {"error": f"no JSON line from the child ({e!r})", "stdout_tail": r.stdout[-400:]}
    return {k: d[k] for k in keep if k in d}


def learned_vs_plain():
    """The implicit configuration's time-to-solution comparison (mms-dirichlet
    2048^2, DIRK3, dt = 1e-3, 50 timed steps): fused Jacobi-PCG against FGMRES
    and against PCG with the symmetric learned network (one network, trained
    on the GPU first); the time to solution is reported per step and with the
    training included."""
    res = child_json(["--workload", "implicit", "--isteps", "50", "--jtimes", "--fused-pcg", "--imethods",
                      "jacobi,learned-sym,learned-sym-pcg", "--no-untrained"],
                     ("config", "steps_timed", "gamma_dt_over_dx2", "jacobi_pcg", "learned_symmetric_fgmres",
                      "learned_symmetric_pcg", "speedup_learned_symmetric_fgmres_vs_jacobi_pcg",
                      "speedup_learned_symmetric_pcg_vs_jacobi_pcg",
                      "rel_l2_diff_learned_symmetric_fgmres_vs_jacobi_pcg",
                      "rel_l2_diff_learned_symmetric_pcg_vs_jacobi_pcg", "time_to_solution", "training",
                      "fgmres_restart_symmetric"))
    log(f"learned vs plain: speedup {res.get('speedup_learned_symmetric_fgmres_vs_jacobi_pcg')}")
    return res


def batched_vs_sequential():
    """The 256-problem batched STS configuration: the batch engine against the
    same problems solved one after another on the single-problem fast path."""
    res = child_json(["--workload", "batch"],
                     ("metric", "n_problems", "tout", "method", "rtol", "batched_s", "sequential_s", "speedup",
                      "batched_solves_per_s", "sequential_solves_per_s", "step_groups", "phase_launches",
                      "batched_cell_rhs_per_s", "rhs_evals_batched",
                      "rhs_evals_sequential", "abs_l2_error_vs_exact_max",
                      "max_rel_state_diff_batched_vs_sequential"))
    log(f"batched vs sequential: speedup {res.get('speedup')}")
    return res


if __name__ == "__main__":
    main()
