"""Batched super-time-stepping on the GPU: every problem of an uneven batch
(sizes, coefficients, kinds) must take the same steps as the CPU oracle
integrating it alone, and reach the same state (1e-10 relative, FP64)."""
import numpy as np
import pytest

import oracle_lib as O
from hiplib import package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pk():
    return package()


@pytest.fixture(scope="module")
def ctx(pk):
    c = pk.Context()
    yield c
    c.close()


def problems(seed, n, sides, kind=1):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(rng.choice(sides))
        kx, ky = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
        out.append(dict(kind=kind, nx=side, ny=side + int(rng.integers(0, 3)), kx=float(kx), ky=float(ky)))
    return out


def oracle_run(prm, method, tout, rtol, atol, fixed=0.0, controller=0, touts=None):
    kw = {k: v for k, v in prm.items() if k != "kind"}
    P = O.OracleProblem(prm["kind"], **kw)
    R = O.OracleIntegrator(P, O.RKC if method == "rkc" else O.RKL, P.initial())
    R.set_tolerances(rtol, atol)
    R.set_max_num_steps(100000)
    if fixed > 0:
        R.set_fixed_step(fixed)
    else:
        R.set_controller(controller)
    for tt in (touts or [tout]):
        flag, t, y = R.evolve(tt)
    return flag, t, y, R.stats()


@pytest.mark.parametrize("method", ["rkc", "rkl"])
def test_batch_matches_oracle_per_problem(pk, ctx, method):
    plist = problems(11, 10, [16, 24, 33, 48, 64, 96])
    B = pk.Batch(ctx, plist, method)
    B.SetTolerances(1e-5, 1e-10)
    B.SetMaxNumSteps(100000)
    for tout in (0.05, 0.1):
        assert B.evolve(tout) == 0
    g, launches = B.groups()
    assert g > 0 and launches > g
    for k, prm in enumerate(plist):
        flag, t, yr, sr = oracle_run(prm, method, 0.1, 1e-5, 1e-10, touts=[0.05, 0.1])
        assert flag == 0 and B.time(k) == t
        sg = B.stats(k)
        for f in ("nst", "attempts", "nfe", "netf", "nstlim", "stage_max"):
            assert sg[f] == sr[f], (k, f, sg[f], sr[f])
        yg = B.solution(k)
        assert np.abs(yg - yr).max() <= 1e-10 * np.abs(yr).max(), k
    B.release()


def test_batch_periodic_and_controllers(pk, ctx):
    plist = problems(5, 6, [20, 31, 40], kind=0)
    for controller in (1, 2):
        B = pk.Batch(ctx, plist, "rkl")
        B.SetTolerances(1e-4, 1e-11)
        B.SetController(controller)
        B.SetMaxNumSteps(100000)
        assert B.evolve(0.2) == 0
        for k, prm in enumerate(plist):
            flag, t, yr, sr = oracle_run(prm, "rkl", 0.2, 1e-4, 1e-11, controller=controller)
            assert B.stats(k)["nst"] == sr["nst"]
            yg = B.solution(k)
            assert np.abs(yg - yr).max() <= 1e-10 * np.abs(yr).max()
        B.release()


@pytest.mark.parametrize("rows", [1, 4, 8])
def test_tile_heights_agree(pk, ctx, rows):
    plist = problems(3, 5, [32, 64, 128])
    ref = pk.Batch(ctx, plist, "rkc")
    ref.SetTolerances(1e-5, 1e-10)
    ref.SetTileRows(2)
    assert ref.evolve(0.05) == 0
    B = pk.Batch(ctx, plist, "rkc")
    B.SetTolerances(1e-5, 1e-10)
    B.SetTileRows(rows)
    assert B.evolve(0.05) == 0
    for k in range(len(plist)):
        a, b = ref.solution(k), B.solution(k)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    ref.release()
    B.release()


def test_batch_fixed_step(pk, ctx):
    plist = problems(8, 4, [40, 64])
    B = pk.Batch(ctx, plist, "rkl")
    B.SetFixedStep(2e-4)
    B.SetMaxNumSteps(100000)
    assert B.evolve(0.01) == 0
    for k, prm in enumerate(plist):
        flag, t, yr, sr = oracle_run(prm, "rkl", 0.01, 1e-5, 1e-10, fixed=2e-4)
        assert B.stats(k)["nfe"] == sr["nfe"]
        yg = B.solution(k)
        assert np.abs(yg - yr).max() <= 1e-10 * np.abs(yr).max()
    B.release()


def test_batch_step_cap_reports_too_much_work(pk, ctx):
    B = pk.Batch(ctx, problems(1, 3, [64]), "rkc")
    B.SetTolerances(1e-8, 1e-12)
    B.SetMaxNumSteps(3)
    assert B.evolve(0.5) == -1
    B.release()


def test_stages_per_pass_agree_bitwise(pk, ctx):
    """The batch result does not depend on how many stages a phase launch
    advances (1..5 per pass, the tile halo covering five): the passes are
    bitwise equal to single stages."""
    plist = problems(13, 6, [32, 64, 96])
    ref = None
    for kpass in (1, 2, 3, 4, 5):
        B = pk.Batch(ctx, plist, "rkc")
        B.SetTolerances(1e-5, 1e-10)
        B.SetMaxNumSteps(100000)
        assert B.SetStagesPerPass(kpass) == 0
        assert B.evolve(0.05) == 0
        sols = [B.solution(k) for k in range(len(plist))]
        stats = [B.stats(k)["nfe"] for k in range(len(plist))]
        B.release()
        if ref is None:
            ref = (sols, stats)
            continue
        assert stats == ref[1], kpass
        for a, b in zip(sols, ref[0]):
            assert np.array_equal(a, b), kpass
    B = pk.Batch(ctx, plist, "rkc")
    assert B.SetStagesPerPass(6) != 0 and B.SetStagesPerPass(0) != 0
    B.release()


@pytest.mark.parametrize("method", ["rkc", "rkl"])
def test_mixed_batch_equals_single_problems(pk, ctx, method):
    """40 mixed problems (both kinds, odd and even sides: the per-tile layout
    dispatch) end bitwise where the single-problem fused path ends, with the
    same counts, after two evolve calls."""
    plist = problems(21, 20, [16, 24, 33, 48]) + problems(22, 20, [16, 31, 40], kind=0)
    B = pk.Batch(ctx, plist, method)
    B.SetTolerances(1e-5, 1e-10)
    B.SetMaxNumSteps(100000)
    assert B.evolve(0.03) == 0
    assert B.evolve(0.06) == 0
    g, launches = B.groups()
    assert g > 0 and launches > g
    for k, prm in enumerate(plist):
        kw = {kk: v for kk, v in prm.items() if kk != "kind"}
        G = pk.Problem(ctx, prm["kind"], **kw)
        I = pk.Integrator(G, method, G.initial())
        I.ARKodeSStolerances(1e-5, 1e-10)
        I.ARKodeSetMaxNumSteps(100000)
        assert I.LSRKStepSetDiffusionFastPath(G.h, 1) == 0
        for tout in (0.03, 0.06):
            flag, t = I.evolve(tout)
            assert flag == 0
        assert B.time(k) == t
        si, sb = I.stats(), B.stats(k)
        for f in ("nst", "attempts", "nfe", "netf", "nstlim", "stage_max"):
            assert si[f] == sb[f], (k, f, si[f], sb[f])
        assert np.array_equal(I.y.get(), B.solution(k)), k
        I.free()
        G.destroy()
    B.release()


@pytest.mark.parametrize("thin", [(0, 1), (4, 1 << 30), (12, 2048), (1, 64)])
def test_thin_phase_tiles_do_not_change_the_result(pk, ctx, thin):
    """Thin phases may run on finer tiles (shorter column marches): never
    (rows 0), always with 4 rows, the default, or 1 row below 64 tiles.  The
    tiling does not enter the cells' values: every setting gives the
    coarse-only batch's states bit for bit, with the same counts."""
    plist = problems(23, 16, [16, 33, 48, 64]) + problems(24, 8, [24, 31], kind=0)
    out = []
    for rows, tiles in ((0, 1), thin):
        B = pk.Batch(ctx, plist, "rkc")
        B.SetTolerances(1e-5, 1e-10)
        B.SetMaxNumSteps(100000)
        assert pk.lib().DiffusionBatch_SetThinPhases(B.h, rows, tiles) == 0
        assert B.evolve(0.04) == 0
        out.append(([B.solution(k) for k in range(len(plist))], [B.stats(k)["nfe"] for k in range(len(plist))]))
        B.release()
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert out[0][1] == out[1][1]


def test_thin_phase_setting_rejects_bad_values(pk, ctx):
    B = pk.Batch(ctx, problems(25, 2, [16]), "rkc")
    assert pk.lib().DiffusionBatch_SetThinPhases(B.h, -1, 10) != 0
    assert pk.lib().DiffusionBatch_SetThinPhases(B.h, 257, 10) != 0
    assert pk.lib().DiffusionBatch_SetThinPhases(B.h, 8, 0) != 0
    B.release()
