"""HIP N_Vector on the GPU: known answers of the reference's N_Vector test
(gkeyll_diffusion/test_nvector_gkylzero.c:97-324), bitwise element-wise parity
with the serial oracle/numpy, and reductions against exact sums."""
import ctypes as C
import math

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


def vec(pk, ctx, a):
    return pk.Vector.from_numpy(ctx, np.asarray(a, dtype=np.float64))


def test_known_answers(pk, ctx):
    L = pk.lib()
    n = 100 * 100  # num_basis x size of the reference test
    a, b, c, d = 2.0, 3.0, 1.75, 2.89
    t = vec(pk, ctx, np.arange(n, dtype=float))
    r = t.clone()
    L.N_VScale(2.0, t.h, r.h)
    assert np.all(np.abs(r.get() - 2.0 * np.arange(n)) <= 1e-10)
    L.N_VConst(173.0, r.h)
    assert np.all(r.get() == 173.0)
    v1, v2, z = t.clone(), t.clone(), t.clone()
    L.N_VConst(c, v1.h)
    L.N_VConst(d, v2.h)
    L.N_VLinearSum(a, v1.h, b, v2.h, z.h)
    assert np.all(np.abs(z.get() - (a * c + b * d)) <= 1e-10)
    L.N_VConst(-0.5, v1.h)
    L.N_VConst(0.5, v2.h)
    assert abs(L.N_VWrmsNorm(v1.h, v2.h) - 0.25) <= 1e-10
    L.N_VConst(c, v1.h)
    L.N_VConst(d, v2.h)
    L.N_VDiv(v1.h, v2.h, z.h)
    assert np.all(np.abs(z.get() - c / d) <= 1e-10)
    L.N_VConst(-1.0, v1.h)
    L.N_VAbs(v1.h, z.h)
    assert np.all(z.get() == 1.0)
    L.N_VConst(c, v1.h)
    L.N_VInv(v1.h, z.h)
    assert np.all(np.abs(z.get() - 1.0 / c) <= 1e-10)
    L.N_VConst(0.0, v1.h)
    assert L.N_VMaxNorm(v1.h) == 0.0
    L.N_VConst(c, v1.h)
    L.N_VAddConst(v1.h, d, z.h)
    assert np.all(np.abs(z.get() - (c + d)) <= 1e-10)


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 4097, 1 << 20])
def test_elementwise_bitwise(pk, ctx, n):
    L = pk.lib()
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n) + 3.0
    X, Y, Z = vec(pk, ctx, x), vec(pk, ctx, y), vec(pk, ctx, np.zeros(n))
    L.N_VLinearSum(0.3, X.h, -1.7, Y.h, Z.h)
    np.testing.assert_array_equal(Z.get(), 0.3 * x + -1.7 * y)
    L.N_VProd(X.h, Y.h, Z.h)
    np.testing.assert_array_equal(Z.get(), x * y)
    L.N_VDiv(X.h, Y.h, Z.h)
    np.testing.assert_array_equal(Z.get(), x / y)
    L.N_VScale(-2.5, X.h, Z.h)
    np.testing.assert_array_equal(Z.get(), -2.5 * x)
    L.N_VAbs(X.h, Z.h)
    np.testing.assert_array_equal(Z.get(), np.abs(x))
    L.N_VInv(Y.h, Z.h)
    np.testing.assert_array_equal(Z.get(), 1.0 / y)
    L.N_VAddConst(X.h, 0.125, Z.h)
    np.testing.assert_array_equal(Z.get(), x + 0.125)
    L.N_VCompare(0.5, X.h, Z.h)
    np.testing.assert_array_equal(Z.get(), (np.abs(x) >= 0.5).astype(float))
    # in place
    L.N_VLinearSum(1.0, X.h, 2.0, X.h, X.h)
    np.testing.assert_array_equal(X.get(), 1.0 * x + 2.0 * x)


@pytest.mark.parametrize("nv", [1, 2, 3, 4, 5, 6, 7, 8, 20, 64])
@pytest.mark.parametrize("n", [5, 4096, 100003])
def test_linear_combination_bitwise_vs_oracle(pk, ctx, nv, n):
    rng = np.random.default_rng(nv * 7 + n)
    xs = [rng.standard_normal(n) for _ in range(nv)]
    c = rng.standard_normal(nv)
    V = [vec(pk, ctx, x) for x in xs]
    Z = vec(pk, ctx, np.zeros(n))
    pk.lib().N_VLinearCombination(nv, pk.dbl_array(list(c)), pk.vec_array(V), Z.h)
    ref = np.zeros(n)
    Xp = (C.POINTER(C.c_double) * nv)(*[O.ptr(x) for x in xs])
    O.lib().orc_vlinearcombination(nv, O.ptr(np.ascontiguousarray(c)), Xp, O.ptr(ref), n)
    np.testing.assert_array_equal(Z.get(), ref)
    # aliasing z == X[0] (the LSRK stage update writes into its own input)
    pk.lib().N_VLinearCombination(nv, pk.dbl_array(list(c)), pk.vec_array(V), V[0].h)
    np.testing.assert_array_equal(V[0].get(), ref)


def test_scale_add_multi_and_dot_multi(pk, ctx):
    n, nv = 50001, 9
    rng = np.random.default_rng(11)
    x = rng.standard_normal(n)
    ys = [rng.standard_normal(n) for _ in range(nv)]
    a = rng.standard_normal(nv)
    X = vec(pk, ctx, x)
    Y = [vec(pk, ctx, y) for y in ys]
    Z = [vec(pk, ctx, np.zeros(n)) for _ in range(nv)]
    pk.lib().N_VScaleAddMulti(nv, pk.dbl_array(list(a)), X.h, pk.vec_array(Y), pk.vec_array(Z))
    for k in range(nv):
        np.testing.assert_array_equal(Z[k].get(), a[k] * x + ys[k])
    out = (C.c_double * nv)()
    pk.lib().N_VDotProdMulti(nv, X.h, pk.vec_array(Y), out)
    for k in range(nv):
        exact = math.fsum(x * ys[k])
        assert abs(out[k] - exact) <= 1e-12 * math.fsum(np.abs(x * ys[k]))


@pytest.mark.parametrize("n", [1, 7, 1 << 16, 4096 * 4096])
def test_reductions(pk, ctx, n):
    L = pk.lib()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    w = rng.random(n) + 0.5
    X, W = vec(pk, ctx, x), vec(pk, ctx, w)
    rel = 1e-12
    assert abs(L.N_VDotProd(X.h, W.h) - math.fsum(x * w)) <= rel * math.fsum(np.abs(x * w))
    assert L.N_VMaxNorm(X.h) == np.abs(x).max()
    assert L.N_VMin(X.h) == x.min()
    wr = math.sqrt(math.fsum((x * w) ** 2) / n)
    assert abs(L.N_VWrmsNorm(X.h, W.h) - wr) <= rel * wr
    assert abs(L.N_VL1Norm(X.h) - math.fsum(np.abs(x))) <= rel * math.fsum(np.abs(x))
    wl2 = math.sqrt(math.fsum((x * w) ** 2))
    assert abs(L.N_VWL2Norm(X.h, W.h) - wl2) <= rel * wl2
    # deterministic: the same reduction twice gives the same bits
    assert L.N_VDotProd(X.h, W.h) == L.N_VDotProd(X.h, W.h)


def test_invtest_constrmask_minquotient(pk, ctx):
    L = pk.lib()
    x = np.array([1.0, -2.0, 0.0, 4.0, 0.5])
    X, Z = vec(pk, ctx, x), vec(pk, ctx, np.full(5, 7.0))
    assert L.N_VInvTest(X.h, Z.h) == 0
    np.testing.assert_array_equal(Z.get(), [1.0, -0.5, 7.0, 0.25, 2.0])
    X2 = vec(pk, ctx, np.array([1.0, 2.0, 4.0]))
    Z2 = vec(pk, ctx, np.zeros(3))
    assert L.N_VInvTest(X2.h, Z2.h) == 1
    c = vec(pk, ctx, np.array([0.0, 1.0, -1.0, 2.0, -2.0]))
    xv = vec(pk, ctx, np.array([-5.0, -1.0, -1.0, 0.0, 3.0]))
    m = vec(pk, ctx, np.zeros(5))
    assert L.N_VConstrMask(c.h, xv.h, m.h) == 0
    np.testing.assert_array_equal(m.get(), [0.0, 1.0, 0.0, 1.0, 1.0])
    num = vec(pk, ctx, np.array([1.0, 2.0, 3.0]))
    den = vec(pk, ctx, np.array([2.0, 0.0, -1.0]))
    assert L.N_VMinQuotient(num.h, den.h) == -3.0


def test_seed_vector_matches_oracle(pk, ctx):
    n = 100001
    v = vec(pk, ctx, np.zeros(n))
    pk.lib().N_VSeedRandom_HipNative(v.h, 20240917)
    np.testing.assert_array_equal(v.get(), O.seed_vector(n, 20240917))


def test_make_wraps_without_owning(pk, ctx):
    L = pk.lib()
    owner = vec(pk, ctx, np.arange(10.0))
    dptr = L.N_VGetDeviceArrayPointer(owner.h)
    w = pk.Vector(ctx, handle=L.N_VMake_HipNative(10, C.cast(dptr, C.c_void_p), ctx.h))
    L.N_VScale(3.0, w.h, w.h)
    w.destroy()  # must not free the owner's memory
    np.testing.assert_array_equal(owner.get(), 3.0 * np.arange(10.0))
