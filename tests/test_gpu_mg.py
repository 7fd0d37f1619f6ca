"""The multigrid preconditioner on the GPU: the V-cycle equals the CPU
oracle's bit for bit (odd and even sizes, anisotropic coefficients, the
manufactured Dirichlet and the reference's periodic problem), PCG
with it takes the oracle's steps on the implicit path, and it cuts the
Krylov iterations of Jacobi-PCG by an order of magnitude."""
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


@pytest.mark.parametrize("nx,ny,kx,ky", [(23, 19, 1.0, 1.0), (30, 18, 1.0, 1.0), (64, 64, 1.0, 0.5),
                                         (127, 65, 2.0, 0.3), (256, 200, 1.0, 1.0), (7, 7, 1.0, 1.0)])
@pytest.mark.parametrize("gamma", [1e-3, 5e-2])
def test_vcycle_bitwise(pk, ctx, nx, ny, kx, ky, gamma):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=nx, ny=ny, kx=kx, ky=ky)
    P = O.OracleProblem(O.MMS, nx=nx, ny=ny, kx=kx, ky=ky)
    r = np.random.default_rng(nx * ny).standard_normal(P.n)
    assert G.mg_psetup(gamma) == 0
    R, Z = G.vector(r), G.vector()
    assert G.mg_psolve(R, Z) == 0
    zr, nl = P.mg_apply(gamma, r)
    assert pk.lib().DiffusionProblem_MGLevels(G.h) == nl
    np.testing.assert_array_equal(Z.get(), zr)
    # a second setup with another gamma reuses the captured cycle
    assert G.mg_psetup(2 * gamma) == 0
    assert G.mg_psolve(R, Z) == 0
    np.testing.assert_array_equal(Z.get(), P.mg_apply(2 * gamma, r)[0])


@pytest.mark.parametrize("relax,nu", [(2, 1), (2, 3), (3, 2), (3, 1), (2, 0)])
def test_vcycle_options_bitwise(pk, ctx, relax, nu):
    """--pfmg_relax / --pfmg_nrelax: every smoother option equals the oracle's,
    switching options recaptures the cycle, and the cycle counter counts."""
    nx, ny, gamma = 97, 80, 2e-2
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=nx, ny=ny, kx=1.0, ky=0.5)
    P = O.OracleProblem(O.MMS, nx=nx, ny=ny, kx=1.0, ky=0.5)
    r = np.random.default_rng(relax * 10 + nu).standard_normal(P.n)
    R, Z = G.vector(r), G.vector()
    assert G.mg_psetup(gamma) == 0 and G.mg_psolve(R, Z) == 0
    np.testing.assert_array_equal(Z.get(), P.mg_apply(gamma, r)[0])
    assert G.mg_options(relax, nu) == 0
    assert G.mg_psetup(gamma) == 0 and G.mg_psolve(R, Z) == 0
    np.testing.assert_array_equal(Z.get(), P.mg_apply(gamma, r, relax, nu)[0])
    assert G.mg_cycles() == 2
    assert G.mg_options(1, 2) == -1 and G.mg_options(2, -1) == -1 and G.mg_options(3, 17) == -1


@pytest.mark.parametrize("nx,ny,kx,ky,inh", [(24, 20, 1.0, 1.0, 0), (64, 32, 1.0, 0.5, 0), (128, 96, 2.0, 0.3, 1),
                                             (6, 10, 1.0, 0.7, 0), (256, 256, 1.0, 1.0, 1)])
@pytest.mark.parametrize("relax,nu", [(2, 2), (3, 1)])
def test_periodic_vcycle_bitwise(pk, ctx, nx, ny, kx, ky, inh, relax, nu):
    """The reference's periodic problem (hypre PFMG's use in the reference):
    wrap-around coarsening, transfers and smoothing equal the oracle's."""
    G = pk.Problem(ctx, pk.CEDA_PERIODIC, nx=nx, ny=ny, kx=kx, ky=ky, inhomogeneous=inh)
    P = O.OracleProblem(O.CEDA, nx=nx, ny=ny, kx=kx, ky=ky, inhomogeneous=inh)
    r = np.random.default_rng(nx + 7 * ny).standard_normal(P.n)
    assert G.mg_options(relax, nu) == 0
    assert G.mg_psetup(3e-3) == 0
    R, Z = G.vector(r), G.vector()
    assert G.mg_psolve(R, Z) == 0
    zr, nl = P.mg_apply(3e-3, r, relax, nu)
    assert pk.lib().DiffusionProblem_MGLevels(G.h) == nl
    np.testing.assert_array_equal(Z.get(), zr)


def test_periodic_odd_side_is_rejected(pk, ctx):
    G = pk.Problem(ctx, pk.CEDA_PERIODIC, nx=15, ny=16)
    assert G.mg_psetup(1e-3) == -1


def dirk(pk, ctx, side, dt, tout, precond, ls="cg", kind=None):
    kind = pk.MMS_DIRICHLET if kind is None else kind
    G = pk.Problem(ctx, kind, nx=side, ny=side, kx=1.0, ky=1.0)
    I = pk.Integrator(G, "dirk", G.initial(), ls=ls, maxl=20000, precond=precond)
    I.ARKodeSetOrder(3)
    I.ARKodeSetFixedStep(dt)
    I.ARKodeSetLSRelativeTolerance(1e-8)
    I.analytic_jtimes()
    flag, t = I.evolve(tout)
    out = dict(flag=flag, t=t, y=I.y.get(), stats=I.stats())
    I.free()
    return out


def test_pcg_mg_matches_oracle(pk, ctx):
    side, dt, tout = 96, 2e-3, 0.01
    g = dirk(pk, ctx, side, dt, tout, "mg")
    P = O.OracleProblem(O.MMS, nx=side, ny=side, kx=1.0, ky=1.0)
    R = O.OracleIntegrator(P, O.DIRK, P.initial())
    R.set_order(3)
    R.set_fixed_step(dt)
    R.set_lin_solver(20000, 0.0, 0, 2)
    R.set_ls_reltol(1e-8)
    R.set_jtimes_analytic(1)
    fr, tr, yr = R.evolve(tout)
    sr = R.stats()
    assert g["flag"] == fr == 0 and abs(g["t"] - tr) <= 1e-14
    for f in ("nst", "nni", "npsetup"):
        assert g["stats"][f] == sr[f], f
    assert abs(g["stats"]["nli"] - sr["nli"]) <= max(1, 0.01 * sr["nli"])
    assert np.abs(g["y"] - yr).max() <= 1e-10 * np.abs(yr).max()


def test_periodic_pcg_mg_matches_oracle(pk, ctx):
    side, dt, tout = 64, 5e-3, 0.02
    g = dirk(pk, ctx, side, dt, tout, "mg", kind=pk.CEDA_PERIODIC)
    P = O.OracleProblem(O.CEDA, nx=side, ny=side, kx=1.0, ky=1.0)
    R = O.OracleIntegrator(P, O.DIRK, P.initial())
    R.set_order(3)
    R.set_fixed_step(dt)
    R.set_lin_solver(20000, 0.0, 0, 2)
    R.set_ls_reltol(1e-8)
    R.set_jtimes_analytic(1)
    fr, tr, yr = R.evolve(tout)
    sr = R.stats()
    assert g["flag"] == fr == 0 and abs(g["t"] - tr) <= 1e-14
    for f in ("nst", "nni", "npsetup"):
        assert g["stats"][f] == sr[f], f
    assert abs(g["stats"]["nli"] - sr["nli"]) <= max(1, 0.01 * sr["nli"])
    assert np.abs(g["y"] - yr).max() <= 1e-10 * np.abs(yr).max()


def test_mg_cuts_krylov_iterations(pk, ctx):
    side = 128
    gamma_a = 1767732205903.0 / 4055673282236.0
    dt = 1830.0 / (gamma_a * (side + 1) ** 2)   # gamma*dt/dx^2 of the 2048^2, dt = 1e-3 case
    jac = dirk(pk, ctx, side, dt, dt, True)
    mg = dirk(pk, ctx, side, dt, dt, "mg")
    assert jac["flag"] == mg["flag"] == 0
    assert mg["stats"]["nli"] * 10 <= jac["stats"]["nli"], (mg["stats"]["nli"], jac["stats"]["nli"])
    scale = np.abs(jac["y"]).max()
    assert np.abs(mg["y"] - jac["y"]).max() <= 1e-7 * scale
