"""The fused PCG of the implicit path (two kernels per iteration with
Jacobi, ARKodeSetDiffusionFastPCG) solves the same systems as the N_Vector PCG
(SUNLinSol_PCG with the analytic J v): the same steps and Newton iterations,
Krylov iterations within 1 %, final states within 1e-10, for the Jacobi and
multigrid preconditioners and none, both problems and both residual norms
(relative 2-norm, and ARKODE's weighted norm when no LS tolerance is set); it
matches the CPU oracle the same way, and stays off where it does not apply
(difference-quotient J v)."""
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


def dirk(pk, ctx, kind, side, dt, tout, precond, fused, reltol=1e-8, jtimes=True, ny=None, bulk=None):
    G = pk.Problem(ctx, kind, nx=side, ny=ny or side, kx=1.0, ky=0.7)
    I = pk.Integrator(G, "dirk", G.initial(), ls="cg", maxl=20000, precond=precond)
    I.ARKodeSetOrder(3)
    if dt:
        I.ARKodeSetFixedStep(dt)
    else:
        I.ARKodeSStolerances(1e-6, 1e-10)
    if reltol:
        I.ARKodeSetLSRelativeTolerance(reltol)
    if jtimes:
        I.analytic_jtimes()
    if fused:
        assert I.fused_pcg() == 0
    if bulk is not None:
        assert pk.lib().ARKodeSetFusedPCGBulk(I.h, bulk) == 0
    flag, t = I.evolve(tout)
    out = dict(flag=flag, t=t, y=I.y.get(), stats=I.stats())
    I.free()
    return out


def same_solve(a, b, exact_iters=False):
    assert a["flag"] == b["flag"] == 0 and a["t"] == b["t"]
    for f in ("nst", "nni", "npsetup"):
        assert a["stats"][f] == b["stats"][f], f
    na, nb = a["stats"]["nli"], b["stats"]["nli"]
    assert abs(na - nb) <= max(1, 0.01 * nb), (na, nb)
    assert np.abs(a["y"] - b["y"]).max() <= 1e-10 * np.abs(b["y"]).max()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("precond", [True, "mg", False])
@pytest.mark.parametrize("reltol", [1e-8, 0.0])
def test_fused_pcg_matches_nvector_pcg(pk, ctx, kind, precond, reltol):
    side, dt = (64, 4e-3) if kind == 0 else (96, 2e-3)
    ref = dirk(pk, ctx, kind, side, dt, 4 * dt, precond, False, reltol)
    got = dirk(pk, ctx, kind, side, dt, 4 * dt, precond, True, reltol)
    same_solve(got, ref)
    if precond is True:
        assert got["stats"]["npsolve"] == ref["stats"]["npsolve"] or \
            abs(got["stats"]["npsolve"] - ref["stats"]["npsolve"]) <= max(2, 0.02 * ref["stats"]["npsolve"])


@pytest.mark.parametrize("kind", [0, 1])
def test_fused_pcg_with_adaptive_steps(pk, ctx, kind):
    """Adaptive steps change gamma between preconditioner setups; the fused
    Jacobi PCG must apply the inverse diagonal of the last setup (what the
    N_Vector PCG's PSolve uses), not one of the current gamma."""
    ref = dirk(pk, ctx, kind, 64, None, 0.02, True, False)
    got = dirk(pk, ctx, kind, 64, None, 0.02, True, True)
    assert ref["stats"]["nst"] > 3
    same_solve(got, ref)


@pytest.mark.parametrize("side,ny", [(63, 40), (50, 37)])
def test_fused_pcg_odd_and_even_rows(pk, ctx, side, ny):
    """Rows of odd length take the cell-wise kernels, even ones the pair
    kernels."""
    ref = dirk(pk, ctx, pk.MMS_DIRICHLET, side, 4e-3, 1.2e-2, True, False, ny=ny)
    got = dirk(pk, ctx, pk.MMS_DIRICHLET, side, 4e-3, 1.2e-2, True, True, ny=ny)
    same_solve(got, ref)


def test_fused_pcg_matches_oracle(pk, ctx):
    side, dt, tout = 96, 2e-3, 0.01
    g = dirk(pk, ctx, pk.MMS_DIRICHLET, side, dt, tout, True, True)
    P = O.OracleProblem(O.MMS, nx=side, ny=side, kx=1.0, ky=0.7)
    R = O.OracleIntegrator(P, O.DIRK, P.initial())
    R.set_order(3)
    R.set_fixed_step(dt)
    R.set_lin_solver(20000, 0.0, 0, 1)
    R.set_ls_reltol(1e-8)
    R.set_jtimes_analytic(1)
    fr, tr, yr = R.evolve(tout)
    sr = R.stats()
    assert g["flag"] == fr == 0 and abs(g["t"] - tr) <= 1e-14
    for f in ("nst", "nni", "npsetup"):
        assert g["stats"][f] == sr[f], f
    assert abs(g["stats"]["nli"] - sr["nli"]) <= max(1, 0.01 * sr["nli"])
    assert np.abs(g["y"] - yr).max() <= 1e-10 * np.abs(yr).max()


def test_fused_pcg_off_with_difference_quotient(pk, ctx):
    """Without the analytic J v the integrator keeps the N_Vector PCG:
    bitwise the run without the option."""
    a = dirk(pk, ctx, pk.MMS_DIRICHLET, 48, 2e-3, 6e-3, True, False, jtimes=False)
    b = dirk(pk, ctx, pk.MMS_DIRICHLET, 48, 2e-3, 6e-3, True, True, jtimes=False)
    assert a["stats"] == b["stats"]
    np.testing.assert_array_equal(a["y"], b["y"])


@pytest.mark.parametrize("bulk", [0.0, 1.0])
def test_fused_pcg_bulk_fraction_changes_only_the_waits(pk, ctx, bulk):
    """With a preconditioner callback (here the multigrid V-cycle) the host
    enqueues a fraction of the previous solve's iterations before it tests
    the residual; the device stops each solve at the same iterate whatever
    the fraction, so the states, the counts and every bit agree."""
    ref = dirk(pk, ctx, pk.MMS_DIRICHLET, 96, 2e-3, 8e-3, "mg", True)
    got = dirk(pk, ctx, pk.MMS_DIRICHLET, 96, 2e-3, 8e-3, "mg", True, bulk=bulk)
    assert got["flag"] == ref["flag"] == 0
    for f in ("nst", "nni", "nli"):
        assert got["stats"][f] == ref["stats"][f], f
    assert np.array_equal(got["y"], ref["y"])


def test_fused_pcg_bulk_fraction_rejects_other_values(pk, ctx):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=16, ny=16, kx=1.0, ky=1.0)
    I = pk.Integrator(G, "dirk", G.initial(), ls="cg", maxl=100, precond=True)
    assert pk.lib().ARKodeSetFusedPCGBulk(I.h, -0.1) != 0
    assert pk.lib().ARKodeSetFusedPCGBulk(I.h, 1.5) != 0
    assert pk.lib().ARKodeSetFusedPCGBulk(I.h, float("nan")) != 0
    I.free()
