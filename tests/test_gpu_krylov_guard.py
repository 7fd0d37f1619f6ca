"""A NaN can never pass a Krylov residual test.  Every linear solver of the
implicit path (PCG, flexible PCG, FGMRES with its Jacobi-PCG fallback, and
the fused PCG with Jacobi, multigrid or a callback preconditioner) must stop
with ARK_LSOLVE_FAIL and a message instead of running to its iteration cap
and having the integrator retry (which looks like a hang).  The NaN is
planted from the host: in the initial state (the right-hand side is NaN
before the first iteration), and by a preconditioner callback that returns a
NaN correction on its third call (mid-solve).  Every kernel exits normally;
the guard is on the scalars the solvers already compute.  The oracle's PCG
has the same guard (tests/test_oracle.py)."""
import ctypes as C

import numpy as np
import pytest

from hiplib import package

pytestmark = pytest.mark.gpu

ARK_LSOLVE_FAIL = -7
ARK_ILL_INPUT = -22
PSOLVEFN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                       C.c_double, C.c_int, C.c_void_p)


@pytest.fixture(scope="module")
def pk():
    return package()


@pytest.fixture(scope="module")
def ctx(pk):
    c = pk.Context()
    yield c
    c.close()


def make(pk, G, y0, ls, precond, fused):
    I = pk.Integrator(G, "dirk", y0, ls=ls, maxl=5000, precond=precond)
    I.ARKodeSetOrder(3)
    I.ARKodeSetFixedStep(1e-3)
    assert I.ARKodeSetLSRelativeTolerance(1e-8) == 0
    I.analytic_jtimes()
    if fused:
        assert I.fused_pcg() == 0
    return I


CASES = [("cg", True, False), ("fcg", True, False), ("fgmres", True, False), ("cg", True, True),
         ("cg", "mg", True), ("cg", False, True)]


@pytest.mark.parametrize("ls,precond,fused", CASES)
def test_nan_state_fails_the_linear_solve(pk, ctx, ls, precond, fused):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=64, ny=64, kx=1.0, ky=1.0)
    y = G.initial().get()
    y[1000] = np.nan
    I = make(pk, G, G.vector(y), ls, precond, fused)
    flag, t = I.evolve(1e-3)
    st = I.stats()
    I.free()
    assert flag == ARK_LSOLVE_FAIL, flag
    assert st["nli"] == 0 and st["nst"] == 0


@pytest.mark.parametrize("ls,fused", [("cg", False), ("fcg", False), ("fgmres", False), ("cg", True)])
def test_nan_from_the_preconditioner_mid_solve(pk, ctx, ls, fused):
    """The callback runs the Jacobi PSolve, and on its third call overwrites
    z with NaN: PCG's r.z (fused: computed on the device, read on the host),
    FGMRES's |w| turn NaN in that iteration."""
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=64, ny=64, kx=1.0, ky=1.0)
    L = pk.lib()
    calls = [0]

    @PSOLVEFN
    def ps(t, y, fy, r, z, gamma, delta, lr, ud):
        calls[0] += 1
        rc = L.PSolve(t, y, fy, r, z, gamma, delta, lr, ud)
        if calls[0] == 3:
            L.N_VConst(float("nan"), z)
        return rc

    I = make(pk, G, G.initial(), ls, True, fused)
    I.ARKodeSetPreconditioner(pk.fn_ptr("PSetup"), C.cast(ps, C.c_void_p))
    flag, t = I.evolve(1e-3)
    st = I.stats()
    I.free()
    assert flag == ARK_LSOLVE_FAIL, flag
    assert calls[0] == 3 and 1 <= st["nli"] <= 3 and st["nst"] == 0


def test_guard_does_not_fire_on_a_normal_solve(pk, ctx):
    """The same configurations solve normally without the NaN, with the same
    iteration counts on the fused and the N_Vector PCG."""
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=64, ny=64, kx=1.0, ky=1.0)
    nli = {}
    for fused in (False, True):
        I = make(pk, G, G.initial(), "cg", True, fused)
        flag, t = I.evolve(2e-3)
        nli[fused] = I.stats()["nli"]
        I.free()
        assert flag == 0 and t == 2e-3
    assert abs(nli[True] - nli[False]) <= max(1, 0.01 * nli[False])


def test_relative_tolerance_rejects_nan(pk, ctx):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=16, ny=16)
    I = pk.Integrator(G, "dirk", G.initial())
    assert I.ARKodeSetLSRelativeTolerance(float("nan")) == ARK_ILL_INPUT
    assert I.ARKodeSetLSRelativeTolerance(-1e-8) == ARK_ILL_INPUT
    assert I.ARKodeSetLSRelativeTolerance(0.0) == 0
    I.free()
