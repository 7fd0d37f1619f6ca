"""Integrators on separate contexts driven from separate host threads (the
way the batched implicit benchmark runs many problems at once) give bitwise
the results of the same runs one after another: the library keeps no state
across contexts except the mutex-guarded launch-shape cache, and its HIP-graph
captures are thread-local."""
import threading

import numpy as np
import pytest

from hiplib import package

pytestmark = pytest.mark.gpu

CASES = [
    dict(kind="dirk", nx=48, ny=40, kx=1.0, ky=0.5, prec=True),
    dict(kind="dirk", nx=64, ny=64, kx=2.0, ky=0.3, prec="mg"),
    dict(kind="rkc", nx=96, ny=80, kx=1.0, ky=1.0, prec=None),
    dict(kind="rkl", nx=70, ny=64, kx=0.5, ky=2.0, prec=None),
    dict(kind="dirk", nx=33, ny=57, kx=1.0, ky=1.0, prec=True),
    dict(kind="dirk", nx=80, ny=72, kx=1.0, ky=1.0, prec="mg"),
]


def integrate(pk, case):
    ctx = pk.Context()
    try:
        G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=case["nx"], ny=case["ny"], kx=case["kx"], ky=case["ky"])
        if case["kind"] == "dirk":
            I = pk.Integrator(G, "dirk", G.initial(), ls="cg", maxl=5000, precond=case["prec"])
            I.analytic_jtimes()
            I.ARKodeSetOrder(3)
            I.ARKodeSetFixedStep(1e-3)
            I.ARKodeSetLSRelativeTolerance(1e-9)
            tout = 4e-3
        else:
            G.autotune(1)
            I = pk.Integrator(G, case["kind"], G.initial())
            I.ARKodeSStolerances(1e-6, 1e-10)
            assert I.LSRKStepSetDiffusionFastPath(G.h, 1) == 0
            tout = 0.02
        I.ARKodeSetMaxNumSteps(100000)
        flag, t = I.evolve(tout)
        assert flag >= 0
        st = I.stats()
        return t, I.y.get(), {k: st[k] for k in ("nst", "nfe", "nli", "netf")}
    finally:
        ctx.close()


def test_concurrent_contexts_match_sequential():
    pk = package()
    seq = [integrate(pk, c) for c in CASES]
    par = [None] * len(CASES)
    errs = []

    def work(i):
        try:
            par[i] = integrate(pk, CASES[i])
        except Exception as e:
            errs.append(repr(e))

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(CASES))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs, errs
    for i, (a, b) in enumerate(zip(seq, par)):
        assert a[0] == b[0], i
        assert a[2] == b[2], (i, a[2], b[2])
        np.testing.assert_array_equal(a[1], b[1], err_msg=str(CASES[i]))
