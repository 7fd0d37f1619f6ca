"""Pins of the oracle's time-step adaptivity (the rules that choose the
reference path's step sequence, main.cpp:357-377; the harness always runs
--controller 2, runtests-diffusion2d.py:78) by what the controllers'
published definitions fix, not by restating their formulas:

* on the exact model error dsm = C h^(p+1) (an embedding of order p), the I
  controller is deadbeat: the first accepted step lands on
  h* = safety (bias C)^(-1/(p+1)) whatever h0 was, and rejected steps land
  on the same h*;
* every controller's closed loop on that model has its fixed point at
  bias dsm = safety^((p+1)/(k1 - k2 + k3)) (PID/PI/I) or
  safety^((p+1)/k1) (Gustafsson), with the published constants;
* the error of log h decays at the modulus of the largest root of the
  closed loop's characteristic polynomial (computed by numpy), which
  depends on each constant separately, while the order p drops out;
* the growth limits (10^4 on the first step, 20 after), the failure
  bounds (etamin 0.1, etamxf 0.3 after two failures) and the window [1, 1.5]
  in which h is held;
* the initial step heuristic on an eigenmode, where the second-derivative
  estimate is exact: h0 = 0.5 sqrt(2 / |y''|) capped by its upper bound, and
  the first attempt passes its error test on the configurations the tests
  integrate.

A wrong exponent (1/p for 1/(p+1)), a transposed constant or a sign slip
changes one of these numbers."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

PID, PI, I_CTRL, EXPGUS, IMPGUS, IMEXGUS = range(6)
SAFETY, BIAS = 0.96, 1.5
# published constants (SUNDIALS SUNAdaptController docs)
K = {PID: (0.58, 0.21, 0.1), PI: (0.8, 0.31, 0.0), I_CTRL: (1.0, 0.0, 0.0),
     EXPGUS: (0.367, 0.268, 0.0), IMPGUS: (0.98, 0.95, 0.0)}


def run(ctype, p, Cm, h0, n, ubound=1.5, q=None):
    """The oracle's controller on dsm = Cm h^q (q = p + 1 unless given)."""
    f = O.lib().orc_test_controller
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    h, dsm, acc = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
    f(ctype, p, Cm, float(p + 1 if q is None else q), h0, ubound, n, O.ptr(h), O.ptr(dsm),
      acc.ctypes.data_as(C.POINTER(C.c_int)))
    return h, dsm, acc.astype(bool)


def h_star(p, Cm):
    return SAFETY * (BIAS * Cm) ** (-1.0 / (p + 1))


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("Cm", [1e3, 1e9])
def test_i_controller_is_deadbeat(p, Cm):
    hs = h_star(p, Cm)
    for h0 in (hs / 300, hs / 7):
        h, dsm, acc = run(I_CTRL, p, Cm, h0, 4, ubound=1.0)
        assert acc.all()
        assert abs(h[1] / hs - 1) < 1e-13, (h0, h[1], hs)
        assert abs(BIAS * dsm[1] / SAFETY ** (p + 1) - 1) < 1e-12
        assert h[2] == h[1] and h[3] == h[1]  # on the fixed point: eta = 1 exactly


@pytest.mark.parametrize("p", [1, 2])
def test_i_controller_rejection_lands_on_the_same_step(p):
    Cm = 1e6
    hs = h_star(p, Cm)
    h, dsm, acc = run(I_CTRL, p, Cm, 3 * hs / SAFETY, 3, ubound=1.0)
    assert not acc[0] and acc[1]
    assert abs(h[1] / hs - 1) < 1e-13


@pytest.mark.parametrize("ctype", [PID, PI, I_CTRL, EXPGUS, IMPGUS, IMEXGUS])
@pytest.mark.parametrize("p", [1, 2])
def test_fixed_point_of_the_closed_loop(ctype, p):
    Cm = 1e7
    h, dsm, acc = run(ctype, p, Cm, 0.5 * h_star(p, Cm), 400, ubound=1.0)
    if ctype == IMEXGUS:
        # min of the explicit (0.367, 0.268) and implicit (0.95, 0.95) estimates:
        # the explicit one binds near the fixed point (safety < 1)
        expo = 0.367
    elif ctype in (EXPGUS, IMPGUS):
        expo = K[ctype][0]
    else:
        k1, k2, k3 = K[ctype]
        expo = k1 - k2 + k3
    target = SAFETY ** ((p + 1) / expo)
    assert acc[-50:].all()
    # with the window [1, 1] the controller still holds h once its proposal
    # is within 1e-6 of h: log e settles within (p+1) 1e-6 / expo of log e*
    np.testing.assert_allclose(BIAS * dsm[-10:], target, rtol=1.5e-6 * (p + 1) / expo)


def closed_loop_poles(ctype):
    """Roots of the characteristic polynomial of x_n = log(h_n / h*) on the
    model: log e_n - log e* = (p+1) x_n, so (p+1) cancels."""
    k1, k2, k3 = K[ctype]
    if ctype in (PID, PI, I_CTRL):  # x' = x - k1 x + k2 x_1 - k3 x_2
        poly = [1.0, -(1 - k1), -k2, k3]
    elif ctype == EXPGUS:  # x' = x - k1 x - k2 (x - x_1)
        poly = [1.0, -(1 - k1 - k2), -k2, 0.0]
    else:  # IMPGUS: x' = x + (x - x_1) - k1 x - k2 (x - x_1)
        poly = [1.0, -(2 - k1 - k2), (1 - k2), 0.0]
    return np.roots(poly)


@pytest.mark.parametrize("ctype", [PI, EXPGUS])
@pytest.mark.parametrize("p", [1, 2])
def test_convergence_rate_is_the_closed_loop_pole(ctype, p):
    """These controllers' dominant closed-loop pole is real and positive
    (PI 0.6657, ExpGus 0.7314): x_{n+1} / x_n tends to it."""
    Cm = 1e7
    r = closed_loop_poles(ctype)
    dom = r[np.argmax(np.abs(r))]
    assert abs(dom.imag) < 1e-12 and dom.real > 0
    h, dsm, acc = run(ctype, p, Cm, 0.5 * h_star(p, Cm), 200, ubound=1.0)
    expo = (K[ctype][0] - K[ctype][1] + K[ctype][2]) if ctype != EXPGUS else K[ctype][0]
    hfix = (SAFETY ** ((p + 1) / expo) / (BIAS * Cm)) ** (1.0 / (p + 1))
    x = np.log(h / hfix)
    # late enough for the other pole to have died out, well above the 1e-6
    # hold band
    idx = [n for n in range(5, 200) if 3e-5 < abs(x[n]) < 3e-4][:6]
    assert len(idx) >= 4, x[:40]
    # the second pole (PI -0.4657, ExpGus -0.3664) alternates around the
    # dominant one: the geometric mean over the window averages it out
    rate = (x[idx[-1] + 1] / x[idx[0]]) ** (1.0 / (idx[-1] + 1 - idx[0]))
    assert rate == pytest.approx(dom.real, rel=2e-3)


@pytest.mark.parametrize("ctype,n0,n1", [(PID, 2, 10), (IMPGUS, 2, 6)])
def test_oscillating_loops_decay_at_the_pole_modulus(ctype, n0, n1):
    """PID (dominant pole -0.4725, and 0.446 +- 0.112i) and ImpGus (poles of
    modulus sqrt(1 - k2) = 0.2236): the envelope of |x_n| shrinks by the
    largest pole modulus per step."""
    p, Cm = 2, 1e7
    r = closed_loop_poles(ctype)
    mod = np.abs(r).max()
    h, dsm, acc = run(ctype, p, Cm, 0.5 * h_star(p, Cm), 60, ubound=1.0)
    k1, k2, k3 = K[ctype]
    expo = k1 if ctype == IMPGUS else k1 - k2 + k3
    hfix = (SAFETY ** ((p + 1) / expo) / (BIAS * Cm)) ** (1.0 / (p + 1))
    x = np.log(h / hfix)
    rate = (np.abs(x[n1]) + np.abs(x[n1 + 1])) / (np.abs(x[n0]) + np.abs(x[n0 + 1]))
    assert 0.3 * mod ** (n1 - n0) < rate < 3 * mod ** (n1 - n0), (rate, mod ** (n1 - n0))


def test_growth_limits_tiny_floor_and_window():
    # errors below 1e-10 count as 1e-10: from far below, the I controller of
    # order p grows h by safety * 1e10^(1/(p+1)) (2068 for p = 2)
    p, Cm = 2, 1e6
    hs = h_star(p, Cm)
    h, dsm, acc = run(I_CTRL, p, Cm, hs * 1e-6, 2, ubound=1.0)
    assert BIAS * dsm[0] < 1e-10
    assert h[1] / h[0] == pytest.approx(SAFETY * 1e10 ** (1 / 3), rel=1e-12)
    # p = 1: the floor asks for 96000: the first step grows at most 10^4,
    # later ones at most 20
    p, Cm = 1, 1.0
    h0 = np.sqrt(1e-14 / (BIAS * Cm))
    h, dsm, acc = run(I_CTRL, p, Cm, h0, 4, ubound=1.0)
    assert h[1] / h[0] == pytest.approx(1e4, rel=1e-12)
    assert h[2] / h[1] == pytest.approx(20.0, rel=1e-12)
    assert h[3] / h[2] == pytest.approx(20.0, rel=1e-12)
    # window [1, 1.5]: a proposed growth in it keeps h
    p, Cm = 2, 1e6
    h, dsm, acc = run(I_CTRL, p, Cm, hs / 1.3, 3, ubound=1.5)
    assert h[1] == h[0] and h[2] == h[0]
    h, dsm, acc = run(I_CTRL, p, Cm, hs / 1.6, 2, ubound=1.5)
    assert h[1] / hs == pytest.approx(1.0, rel=1e-13)


@pytest.mark.parametrize("ctype", [PID, PI, I_CTRL])
def test_failure_bounds(ctype):
    p = 2
    # a huge error: the reduction is bounded by etamin = 0.1
    h, dsm, acc = run(ctype, p, 1e30, 1.0, 2, ubound=1.0)
    assert not acc[0] and h[1] == 0.1
    # a true error of order 1 where the controller assumes 3: the first
    # failure cuts h less than 0.3, the second and third by exactly
    # etamxf = 0.3, and the step then passes
    h, dsm, acc = run(ctype, p, 1.0, 10.0, 5, ubound=1.0, q=1.0)
    assert list(acc[:4]) == [False, False, False, True]
    assert 0.3 < h[1] / h[0] < 1.0
    assert h[2] / h[1] == pytest.approx(0.3, rel=1e-15)
    assert h[3] / h[2] == pytest.approx(0.3, rel=1e-15)


def eigenmode(nx, ny, kx, ky, m1, m2):
    """A periodic Fourier mode of the homogeneous reference operator and its
    eigenvalue (closed form)."""
    P = O.OracleProblem(O.CEDA, nx=nx, ny=ny, kx=kx, ky=ky)
    x = np.arange(nx)
    y = np.arange(ny)
    v = np.outer(np.cos(2 * np.pi * m2 * y / ny), np.cos(2 * np.pi * m1 * x / nx)).ravel() + 2.0
    dx = P.p.dx
    dy = P.p.dy
    lam = -4 * kx / dx**2 * np.sin(np.pi * m1 / nx) ** 2 - 4 * ky / dy**2 * np.sin(np.pi * m2 / ny) ** 2
    return P, v, lam


@pytest.mark.parametrize("rtol", [1e-4, 1e-6])
def test_initial_step_on_an_eigenmode(rtol):
    """y0 = mode + constant (the constant is in the null space): f = lam (y0 - 2),
    the difference quotient of f along f is lam f exactly, so the heuristic's
    second-derivative norm is |lam^2 (y0 - 2)|_wrms and h0 has a closed form."""
    P, y0, lam = eigenmode(32, 32, 1.0, 1.0, 3, 2)
    atol = 1e-10
    I = O.OracleIntegrator(P, O.RKC, y0)
    I.set_tolerances(rtol, atol)
    I.set_max_num_steps(1)
    I.evolve(1.0)
    f = O.lib().orc_get_h0_used
    f.restype = C.c_double
    f.argtypes = [C.c_void_p]
    h0 = f(I.h)
    ewt = 1.0 / (rtol * np.abs(y0) + atol)
    fy = lam * (y0 - 2.0)
    ydd = np.sqrt(np.mean((lam * fy * ewt) ** 2))
    hub = min(0.1 * 1.0, 1.0 / np.max(np.abs(fy) / (0.1 * np.abs(y0) + 1.0 / ewt)))
    expect = min(0.5 * np.sqrt(2.0 / ydd), hub)
    assert h0 == pytest.approx(expect, rel=1e-9)


@pytest.mark.parametrize("method", [O.RKC, O.RKL, O.DIRK])
@pytest.mark.parametrize("kind", [O.CEDA, O.MMS])
def test_first_attempt_passes_its_error_test(method, kind):
    P = O.OracleProblem(kind, nx=48, ny=40, kx=1.0, ky=0.5)
    I = O.OracleIntegrator(P, method, P.initial())
    I.set_tolerances(1e-6, 1e-10)
    I.set_max_num_steps(1)
    if method != O.DIRK:
        I.set_internal_eig(1, 100, 0.01)
    else:
        I.set_lin_solver(2000, 0.0, 0, 1)
    I.evolve(1.0)
    st = I.stats()
    f = O.lib().orc_get_last_dsm
    f.restype = C.c_double
    f.argtypes = [C.c_void_p]
    dsm = f(I.h)
    assert st["nst"] == 1 and st["netf"] == 0 and 0.0 < dsm <= 1.0
