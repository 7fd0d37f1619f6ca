"""LearnedPrec_ApplyNoAlloc, the network application that allocates nothing
(its workspace is a fixed device array), against LearnedPrec_Apply through
the C ABI: the same kernels on the same operands, so the outputs must be
equal bit for bit; two calls in a row, the second on a smaller input (the
fixed workspace keeps the first call's contents), for the symmetric and the
ReLU network and with addto; sizes beyond the fixed workspace are refused
before anything is launched."""
import ctypes as C

import numpy as np
import pytest

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


def entry(pk):
    f = pk.lib().LearnedPrec_ApplyNoAlloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    return f


def both(pk, ctx, net, nb, nx, ny, seed, addto):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal(nb * nx * ny)
    z0 = rng.standard_normal(nb * nx * ny)
    R = pk.Vector(ctx, n=r.size)
    R.set(r)
    Za, Zb = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    Za.set(z0)
    Zb.set(z0)
    L = pk.lib()
    assert L.LearnedPrec_Apply(net.h, nb, nx, ny, pk.dptr_of(R), pk.dptr_of(Za), addto) == 0
    assert entry(pk)(net.h, nb, nx, ny, pk.dptr_of(R), pk.dptr_of(Zb), addto) == 0
    za, zb = Za.get(), Zb.get()
    for v in (R, Za, Zb):
        v.destroy()
    return za, zb


@pytest.mark.parametrize("symmetric,dil", [(True, (1, 3, 9, 27)), (False, (1, 1, 1, 1))])
def test_no_alloc_apply_equals_apply(pk, ctx, symmetric, dil):
    net = pk.LearnedPrec(ctx, dil, symmetric=symmetric)
    rng = np.random.default_rng(3)
    net.set_params((0.15 * rng.standard_normal(net.nparams)).astype(np.float32))
    net.set_output_scale(0.37)
    # two calls in a row, the second smaller; then one adding to its output
    for nb, nx, ny, seed, addto in ((3, 300, 211, 1, 0), (1, 97, 64, 2, 0), (2, 130, 70, 5, 1)):
        print("no-alloc apply:", symmetric, dil, nb, nx, ny, addto, flush=True)
        za, zb = both(pk, ctx, net, nb, nx, ny, seed, addto)
        assert np.all(np.isfinite(za)) and np.abs(za).max() > 0
        assert np.array_equal(za, zb), (nb, nx, ny, np.abs(za - zb).max())
    net.release()


def test_no_alloc_apply_at_its_largest_size(pk, ctx):
    """1024 x 1024 in one image is exactly the fixed workspace."""
    net = pk.LearnedPrec(ctx, (1, 3, 9, 27), symmetric=True)
    net.init_identity(0.01, 1)
    za, zb = both(pk, ctx, net, 1, 1024, 1024, 9, 0)
    assert np.array_equal(za, zb)
    net.release()


def test_no_alloc_apply_refuses_sizes_beyond_its_workspace(pk, ctx):
    net = pk.LearnedPrec(ctx, (1, 3), symmetric=True)
    f = entry(pk)
    # checked before any launch: the (dummy) device pointers are never used
    dummy = C.c_void_p(16)
    assert f(net.h, 65, 8, 8, dummy, dummy, 0) == -2
    assert f(net.h, 1, 1025, 1024, dummy, dummy, 0) == -2
    assert f(net.h, 2, 1024, 520, dummy, dummy, 0) == -2
    assert f(net.h, 1, 0, 8, dummy, dummy, 0) == -1
    net.release()
