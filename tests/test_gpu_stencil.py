"""Problem-layer kernels on the GPU against the oracle, bit for bit: the
stencil RHS (periodic self-exchange of diffusion_2D.cpp:242-327 and the
Dirichlet/MMS variant), the Jacobi preconditioner, the initial and exact
states, and the closed-form dominant eigenvalue."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from hiplib import package

pytestmark = pytest.mark.gpu

SIZES = [(2, 2), (3, 5), (17, 13), (64, 64), (127, 130), (256, 96), (513, 257)]


@pytest.fixture(scope="module")
def pk():
    return package()


@pytest.fixture(scope="module")
def ctx(pk):
    c = pk.Context()
    yield c
    c.close()


def pair(pk, ctx, kind, **kw):
    G = pk.Problem(ctx, kind, **kw)
    P = O.OracleProblem(kind, **kw)
    return G, P


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("inhom", [0, 1])
def test_periodic_rhs_bitwise(pk, ctx, nx, ny, inhom):
    G, P = pair(pk, ctx, pk.CEDA_PERIODIC, nx=nx, ny=ny, kx=1.3, ky=0.6, inhomogeneous=inhom)
    u = np.random.default_rng(nx * ny + inhom).standard_normal(P.n)
    U, F = G.vector(u), G.vector()
    assert G.rhs(0.0, U, F) == 0
    np.testing.assert_array_equal(F.get(), P.rhs(0.0, u))


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("t", [0.0, 0.37])
def test_mms_rhs_bitwise(pk, ctx, nx, ny, t):
    G, P = pair(pk, ctx, pk.MMS_DIRICHLET, nx=nx, ny=ny, kx=1.0, ky=0.5)
    u = np.random.default_rng(nx + ny).standard_normal(P.n)
    U, F = G.vector(u), G.vector()
    assert G.rhs(t, U, F) == 0
    np.testing.assert_array_equal(F.get(), P.rhs(t, u))


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
def test_launch_shapes_are_bitwise_equal(pk, ctx, rows):
    G, P = pair(pk, ctx, pk.CEDA_PERIODIC, nx=300, ny=77, inhomogeneous=1)
    assert G.set_rows(rows) == 0
    u = np.random.default_rng(rows).standard_normal(P.n)
    U, F = G.vector(u), G.vector()
    G.rhs(0.0, U, F)
    np.testing.assert_array_equal(F.get(), P.rhs(0.0, u))
    assert G.set_rows(3) != 0


CONFIGS = [(0, 1, 0), (0, 2, 0), (0, 8, 0), (1, 1, 0), (1, 3, 0), (1, 8, 1), (1, 64, 0),
           (1, 64, 1)]


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("kind,nx,ny", [(0, 300, 77), (0, 129, 40), (1, 256, 130), (1, 77, 65)])
def test_every_launch_shape_is_bitwise(pk, ctx, cfg, kind, nx, ny):
    """RHS and fused stage kernels in every variant/shape equal the oracle's
    stencil (+ the 5-term stage combination evaluated in numpy) bit for bit."""
    import ctypes as C
    G, P = pair(pk, ctx, kind, nx=nx, ny=ny, kx=1.3, ky=0.6, inhomogeneous=1 if kind == 0 else 0)
    for which in (0, 1):
        assert G.set_config(which, *cfg) == 0
    assert G.get_config(1) == cfg
    rng = np.random.default_rng(nx * 7 + ny)
    u, a, b, g = (rng.standard_normal(P.n) for _ in range(4))
    U, Av, Bv, Gv, Out = G.vector(u), G.vector(a), G.vector(b), G.vector(g), G.vector()
    assert G.rhs(0.0, U, Out) == 0
    F = P.rhs(0.0, u)
    np.testing.assert_array_equal(Out.get(), F)
    ms = C.c_double()
    assert pk.lib().DiffusionProblem_TimeKernel(G.h, 1, U.h, Av.h, Bv.h, Gv.h, Out.h, 1, C.byref(ms)) == 0
    c = [1.0e-9, -0.5, 0.25, 1.25, -1.0e-10]  # the coefficients of the timing entry point
    ref = c[0] * F
    ref = ref + c[1] * a
    ref = ref + c[2] * b
    ref = ref + c[3] * u
    ref = ref + c[4] * g
    np.testing.assert_array_equal(Out.get(), ref)


def test_autotuner_picks_a_valid_shape(pk, ctx):
    G, P = pair(pk, ctx, 1, nx=512, ny=384)
    assert G.autotune(2) == 0
    for which in (0, 1):
        v, r, n = G.get_config(which)
        assert v in (0, 1) and r >= 1 and n in (0, 1)
    u = np.random.default_rng(3).standard_normal(P.n)
    U, F = G.vector(u), G.vector()
    G.rhs(0.1, U, F)
    np.testing.assert_array_equal(F.get(), P.rhs(0.1, u))
    assert G.set_config(0, 0, 4, 1) != 0  # nt only exists for the marching variant


def test_full_size_4096_rhs_bitwise(pk, ctx):
    G, P = pair(pk, ctx, pk.MMS_DIRICHLET, nx=4096, ny=4096, kx=1.0, ky=0.5)
    u = P.exact(0.2) + 1e-3 * np.random.default_rng(1).standard_normal(P.n)
    U, F = G.vector(u), G.vector()
    G.rhs(0.2, U, F)
    np.testing.assert_array_equal(F.get(), P.rhs(0.2, u))


@pytest.mark.parametrize("kind,inhom", [(0, 0), (0, 1), (1, 0)])
def test_jacobi_bitwise(pk, ctx, kind, inhom):
    G, P = pair(pk, ctx, kind, nx=37, ny=24, inhomogeneous=inhom)
    assert G.psetup(3.7e-3) == 0
    d = P.psetup(3.7e-3)
    r = np.random.default_rng(2).standard_normal(P.n)
    R, Z = G.vector(r), G.vector()
    assert G.psolve(R, Z) == 0
    np.testing.assert_array_equal(Z.get(), P.psolve(r))
    np.testing.assert_array_equal(Z.get(), d * r)


@pytest.mark.parametrize("kind", [0, 1])
def test_initial_exact_domeig(pk, ctx, kind):
    G, P = pair(pk, ctx, kind, nx=45, ny=31, kx=0.7, ky=2.0)
    np.testing.assert_array_equal(G.initial().get(), P.initial())
    if kind == 1:
        np.testing.assert_array_equal(G.exact(0.3).get(), P.exact(0.3))
    assert G.dom_eig() == P.dom_eig()


def test_bad_arguments_return_minus_one(pk, ctx):
    G = pk.Problem(ctx, pk.CEDA_PERIODIC, nx=8, ny=8)
    wrong = pk.Vector(ctx, n=10)
    F = G.vector()
    assert G.rhs(0.0, wrong, F) == -1
    with pytest.raises(ValueError):
        pk.Problem(ctx, pk.MMS_DIRICHLET, nx=8, ny=8, inhomogeneous=1)
    with pytest.raises(ValueError):
        pk.Problem(ctx, pk.CEDA_PERIODIC, nx=1, ny=8)
