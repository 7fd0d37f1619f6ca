"""Multi-stage STS passes on the GPU: K consecutive fused stages in one sweep
(intermediate stages kept in registers) must equal, bit for bit, K stages of
the oracle's stencil followed by the 5-term stage combination
(diffusion.cpp:48-53 order, N_VLinearCombination order), for every launch
shape, odd and tiny grids, both boundary kinds; and the integrator's fused
path must give bitwise the same trajectory with passes as without."""
import ctypes as C

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


def mms_factors(t):
    f = O.lib().orc_mms_time
    f.restype = None
    f.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    T1, T2 = C.c_double(), C.c_double()
    f(t, C.byref(T1), C.byref(T2))
    return T1.value, T2.value


def reference_pass(P, kind, K, coef, times, y1, y2, yn, fn):
    """K stages with the oracle's RHS; returns (stage K-1, stage K)."""
    prev2, cur = y2, y1
    out = []
    for m in range(K):
        c = coef[m]
        F = P.rhs(times[m], cur)
        new = c[0] * F
        new = new + c[1] * prev2
        new = new + c[2] * yn
        new = new + c[3] * cur
        new = new + c[4] * fn
        out.append(new)
        prev2, cur = cur, new
    return (out[-2] if K >= 2 else None), out[-1]


SHAPES = [(0, 300, 77), (0, 129, 40), (1, 256, 130), (1, 77, 65), (1, 2, 2), (0, 2, 3),
          (1, 5, 3), (0, 130, 7), (1, 500, 9), (1, 1024, 300)]


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("kind,nx,ny", SHAPES)
def test_stage_pass_bitwise(pk, ctx, K, kind, nx, ny):
    kw = dict(nx=nx, ny=ny, kx=1.3, ky=0.6, inhomogeneous=1 if kind == 0 else 0)
    G = pk.Problem(ctx, kind, **kw)
    P = O.OracleProblem(kind, **kw)
    rng = np.random.default_rng(nx * 31 + ny * 7 + K)
    y1, y2, yn, fn = (rng.standard_normal(P.n) for _ in range(4))
    times = rng.uniform(0.0, 1.0, K)
    coef = np.zeros((K, 8))
    for m in range(K):
        coef[m, :5] = rng.uniform(-1.0, 1.0, 5) * np.array([1e-4, 1.0, 0.5, 1.0, 1e-4])
        coef[m, 5:7] = mms_factors(times[m]) if kind == 1 else (0.0, 0.0)
    ra, rb = reference_pass(P, kind, K, coef, times, y1, y2, yn, fn)
    V = [G.vector(v) for v in (y1, y2, yn, fn)]
    for rows in (1, 3, 8, 32, 200):
        for nt in (0, 1, 2, 3):  # bit 0 non-temporal, bit 1 stage coefficients in LDS
            assert G.set_config(2, K, rows, nt) == 0
            outa, outb = G.vector(), G.vector()
            assert G.stage_pass(K, coef, *V, outa, outb) == 0
            np.testing.assert_array_equal(outb.get(), rb, err_msg=f"rows={rows} nt={nt}")
            if K >= 2:
                np.testing.assert_array_equal(outa.get(), ra, err_msg=f"rows={rows} nt={nt}")


def test_stage_pass_rejects_aliasing(pk, ctx):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=32, ny=32)
    V = [G.vector() for _ in range(5)]
    coef = np.zeros((3, 8))
    assert G.stage_pass(3, coef, V[0], V[1], V[2], V[3], V[0], V[4]) == -1
    assert G.stage_pass(3, coef, V[0], V[1], V[2], V[3], V[4], V[4]) == -1
    assert G.stage_pass(7, coef, V[0], V[1], V[2], V[3], V[4], V[1]) == -1


def fused_run(pk, ctx, kind, method, kw, cfg, graphs, tout, internal):
    G = pk.Problem(ctx, kind, **kw)
    G.autotune(1)
    assert G.set_config(2, *cfg) == 0
    I = pk.Integrator(G, method, G.initial())
    I.ARKodeSStolerances(1e-6, 1e-10)
    I.ARKodeSetMaxNumSteps(100000)
    if internal:
        I.internal_eig(100, 0.01)
    assert I.LSRKStepSetDiffusionFastPath(G.h, graphs) == 0
    assert G.get_config(2) == cfg
    flag, t = I.evolve(tout)
    y = I.y.get()
    st = I.stats()
    I.free()
    return flag, t, y, st


@pytest.mark.parametrize("method", ["rkc", "rkl"])
@pytest.mark.parametrize("kind,nx,ny,internal", [(1, 96, 80, True), (0, 70, 51, False)])
def test_fused_path_with_passes_is_bitwise(pk, ctx, method, kind, nx, ny, internal):
    kw = dict(nx=nx, ny=ny, kx=1.0, ky=0.5)
    base = fused_run(pk, ctx, kind, method, kw, (0, 32, 0), 1, 0.05, internal)
    assert base[0] == 0
    for cfg, graphs in [((2, 16, 0), 1), ((3, 8, 1), 1), ((4, 32, 0), 0), ((3, 200, 0), 1), ((5, 16, 1), 1),
                        ((6, 24, 0), 1), ((5, 48, 2), 1), ((4, 16, 3), 0)]:
        got = fused_run(pk, ctx, kind, method, kw, cfg, graphs, 0.05, internal)
        assert got[0] == 0 and got[1] == base[1]
        for f in ("nst", "attempts", "nfe", "netf", "stage_max"):
            assert got[3][f] == base[3][f], (cfg, f)
        np.testing.assert_array_equal(got[2], base[2], err_msg=str(cfg))


def test_fused_path_with_passes_matches_oracle(pk, ctx):
    kw = dict(nx=128, ny=128, kx=1.0, ky=1.0)
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, **kw)
    P = O.OracleProblem(O.MMS, **kw)
    G.autotune(1)
    assert G.set_config(2, 3, 16, 0) == 0
    I = pk.Integrator(G, "rkc", G.initial())
    R = O.OracleIntegrator(P, O.RKC, P.initial())
    for X in (I,):
        X.ARKodeSStolerances(1e-5, 1e-10)
    R.set_tolerances(1e-5, 1e-10)
    assert I.LSRKStepSetDiffusionFastPath(G.h, 1) == 0
    fg, tg = I.evolve(0.2)
    fr, tr, yr = R.evolve(0.2)
    assert fg == fr == 0 and tg == tr
    sg, sr = I.stats(), R.stats()
    for f in ("nst", "attempts", "nfe", "netf", "stage_max"):
        assert sg[f] == sr[f], f
    yg = I.y.get()
    assert np.abs(yg - yr).max() <= 1e-10 * np.abs(yr).max()
    I.free()
