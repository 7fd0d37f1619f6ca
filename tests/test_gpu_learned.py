"""The learned preconditioner on the GPU: the MFMA forward and backward
kernels against the CPU oracle (BF16 roundings emulated there, FP32 vs FP64
accumulation: agreement to rounding), the Jacobi-PCG training pairs against
the system they solve, training on the GPU, and the network inside FGMRES on
the implicit path (same solution as Jacobi-PCG, fewer Krylov iterations than
the untrained network, the Jacobi-PCG fallback at the iteration cap)."""
import numpy as np
import pytest

import oracle_lib as O
from hiplib import package

pytestmark = pytest.mark.gpu

DIL = (1, 1, 1, 1)


@pytest.fixture(scope="module")
def pk():
    return package()


@pytest.fixture(scope="module")
def ctx(pk):
    c = pk.Context()
    yield c
    c.close()


def apply_A(x, nb, nx, ny, ratio):
    x = x.reshape(nb, ny, nx)
    p = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    lap = (((p[:, 1:-1, :-2] + p[:, 1:-1, 2:]) + p[:, :-2, 1:-1]) + p[:, 2:, 1:-1]) - 4.0 * x
    return (x - ratio * lap).reshape(-1)


def random_params(n, seed, scale=0.25):
    rng = np.random.default_rng(seed)
    return (scale * rng.standard_normal(n)).astype(np.float32)


@pytest.mark.parametrize("nb,nx,ny", [(1, 16, 16), (3, 37, 21), (2, 64, 9)])
def test_forward_matches_oracle(pk, ctx, nb, nx, ny):
    net = pk.LearnedPrec(ctx, DIL)
    p = random_params(net.nparams, nb * nx)
    net.set_params(p)
    net.set_output_scale(0.3)
    rng = np.random.default_rng(7)
    r = rng.standard_normal(nb * nx * ny)
    R, Z = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    R.set(r)
    assert net.apply(nb, nx, ny, R, Z) == 0
    z = Z.get()
    s = np.sqrt((r.reshape(nb, -1) ** 2).mean(axis=1))
    _, _, u = O.lp_lossgrad(DIL, p, 0.3, nb, nx, ny, 1.0, r, r, emu=1, want_grad=False)
    ref = u.reshape(nb, -1) * s[:, None]
    err = z.reshape(nb, -1) - ref
    assert np.abs(err).max() <= 1e-2 * np.abs(ref).max()
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * np.sqrt((ref ** 2).mean())
    # accumulate into an existing vector
    Z2 = pk.Vector(ctx, n=r.size)
    Z2.set(np.ones(r.size))
    assert net.apply(nb, nx, ny, R, Z2, 1) == 0
    np.testing.assert_allclose(Z2.get(), z + 1.0, rtol=0, atol=1e-12 * max(1.0, np.abs(z).max()))
    net.release()


@pytest.mark.parametrize("seed", [0, 1])
def test_loss_and_gradient_match_oracle(pk, ctx, seed):
    nb, nx, ny, ratio = 2, 33, 20, 40.0
    net = pk.LearnedPrec(ctx, DIL)
    if seed == 0:
        net.init_identity(0.05, 3)
        p = net.params()
    else:
        p = random_params(net.nparams, 11, 0.2)
        net.set_params(p)
    net.set_output_scale(1.0 / (1.0 + 4.0 * ratio))
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(nb * nx * ny)
    r = apply_A(e, nb, nx, ny, ratio)
    R, E = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    R.set(r)
    E.set(e)
    loss, g = net.loss_grad(nb, nx, ny, ratio, R, E)
    lref, gref, _ = O.lp_lossgrad(DIL, p, 1.0 / (1.0 + 4.0 * ratio), nb, nx, ny, ratio, r, e, emu=1)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-2 * np.linalg.norm(gref)
    net.release()


@pytest.mark.parametrize("n", [32, 31])  # even sides: the paired kernel, odd: cell by cell
def test_training_pairs_solve_the_system(pk, ctx, n):
    nsamp, ratio = 24, 60.0
    R, E, iters, rc = pk.training_set(ctx, nsamp, n, ratio, reltol=1e-10, seed=1234)
    assert rc == 0 and iters > 5
    r, e = R.get().reshape(nsamp, -1), E.get().reshape(nsamp, -1)
    for i in range(nsamp):
        res = apply_A(e[i], 1, n, n, ratio) - r[i]
        assert np.linalg.norm(res) <= 1.01e-10 * np.linalg.norm(r[i]), i
    # smooth fields and rough (operator-applied) ones alternate
    rough = np.array([np.abs(np.diff(r[i].reshape(n, n), axis=1)).mean() / np.abs(r[i]).mean()
                      for i in range(nsamp)])
    assert rough[1::2].mean() > 1.3 * rough[0::2].mean()


def trained_net(pk, ctx, ratio, n=32, nsamp=2048, steps=1500, batch=16, seed=0, lr=3e-4):
    """The training recipe of the benchmark: identity start with small noise,
    Adam at a small learning rate (larger rates fit the training pairs better
    but give networks that stall FGMRES)."""
    R, E, _, rc = pk.training_set(ctx, nsamp, n, ratio, reltol=1e-12, seed=1234)
    assert rc == 0
    net = pk.LearnedPrec(ctx, DIL)
    net.init_identity(8e-4, seed)
    net.set_output_scale(1.0 / (1.0 + 4.0 * ratio))
    hist = net.train(nsamp, n, ratio, R, E, steps, batch, lr, seed)
    return net, hist


def test_training_lowers_the_loss(pk, ctx):
    net, hist = trained_net(pk, ctx, 80.0, steps=600, lr=2e-3)
    assert np.all(np.isfinite(hist))
    assert hist[-30:].mean() < 0.8 * hist[:10].mean(), (hist[:10].mean(), hist[-30:].mean())
    net.release()


def dirk_run(pk, ctx, prec, tout, dt, side, net=None, maxl=400, mrestart=30, ls=None, jtimes=False, fused=False):
    G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=side, ny=side, kx=1.0, ky=1.0)
    if net is not None:
        net.attach(G)
    ls = ls or ("cg" if prec == "jacobi" else "fgmres")
    I = pk.Integrator(G, "dirk", G.initial(), ls=ls, maxl=maxl, precond="learned" if net is not None else True)
    if ls == "fgmres":
        pk.lib().SUNLinSol_SPFGMRSetMaxRestartDim(I.LS, mrestart)
    if jtimes:
        I.analytic_jtimes()
        if fused:
            I.fused_pcg()
    I.ARKodeSetOrder(3)
    I.ARKodeSetFixedStep(dt)
    I.ARKodeSetLSRelativeTolerance(1e-8)
    I.ARKodeSetMaxNumSteps(100000)
    flag, t = I.evolve(tout)
    out = dict(flag=flag, t=t, y=I.y.get(), stats=I.stats(), fallbacks=I.fallbacks())
    I.free()
    return out


def test_learned_fgmres_on_the_implicit_path(pk, ctx):
    side, dt = 64, 0.02
    dx = 1.0 / (side + 1)
    gamma_a = 1767732205903.0 / 4055673282236.0  # ARK3(2)4L[2]SA diagonal
    ratio = gamma_a * dt / dx ** 2
    net, _ = trained_net(pk, ctx, ratio)
    ident = pk.LearnedPrec(ctx, DIL)
    ident.init_identity(0.0, 0)
    ref = dirk_run(pk, ctx, "jacobi", 0.1, dt, side)
    lrn = dirk_run(pk, ctx, "learned", 0.1, dt, side, net=net)
    idn = dirk_run(pk, ctx, "learned", 0.1, dt, side, net=ident)
    for o in (ref, lrn, idn):
        assert o["flag"] == 0 and abs(o["t"] - 0.1) < 1e-12
    scale = np.abs(ref["y"]).max()
    assert np.abs(lrn["y"] - ref["y"]).max() <= 1e-6 * scale
    assert np.abs(idn["y"] - ref["y"]).max() <= 1e-6 * scale
    assert lrn["fallbacks"] == 0
    # what the trained network gains over the untrained (Jacobi-equivalent) one
    assert lrn["stats"]["nli"] < 0.7 * idn["stats"]["nli"], (lrn["stats"]["nli"], idn["stats"]["nli"])
    # and the CPU oracle's Jacobi-PCG run of the same problem
    P = O.OracleProblem(O.MMS, nx=side, ny=side, kx=1.0, ky=1.0)
    Rr = O.OracleIntegrator(P, O.DIRK, P.initial())
    Rr.set_order(3)
    Rr.set_fixed_step(dt)
    Rr.set_lin_solver(400, 0.0, 0, 1)
    Rr.set_ls_reltol(1e-8)
    fr, tr, yr = Rr.evolve(0.1)
    assert fr == 0
    assert np.abs(lrn["y"] - yr).max() <= 1e-6 * scale
    net.release()
    ident.release()


def test_fallback_to_jacobi_pcg_at_the_cap(pk, ctx):
    side, dt = 48, 0.02
    net = pk.LearnedPrec(ctx, DIL)
    net.init_identity(0.0, 0)
    ref = dirk_run(pk, ctx, "jacobi", 0.04, dt, side)
    capped = dirk_run(pk, ctx, "learned", 0.04, dt, side, net=net, maxl=3, mrestart=3)
    assert capped["flag"] == 0 and capped["fallbacks"] > 0
    scale = np.abs(ref["y"]).max()
    assert np.abs(capped["y"] - ref["y"]).max() <= 1e-6 * scale
    net.release()


def test_flexible_pcg(pk, ctx):
    """Flexible PCG: with the fixed Jacobi preconditioner it takes the PCG
    steps; with the network it reaches the same solution."""
    side, dt = 64, 0.02
    dx = 1.0 / (side + 1)
    ratio = 1767732205903.0 / 4055673282236.0 * dt / dx ** 2
    ref = dirk_run(pk, ctx, "jacobi", 0.06, dt, side, jtimes=True)
    fj = dirk_run(pk, ctx, "jacobi", 0.06, dt, side, ls="fcg", jtimes=True)
    assert fj["flag"] == 0 and abs(fj["stats"]["nli"] - ref["stats"]["nli"]) <= 3
    scale = np.abs(ref["y"]).max()
    assert np.abs(fj["y"] - ref["y"]).max() <= 1e-8 * scale
    net, _ = trained_net(pk, ctx, ratio)
    fl = dirk_run(pk, ctx, "learned", 0.06, dt, side, net=net, ls="fcg", jtimes=True)
    assert fl["flag"] == 0 and fl["fallbacks"] == 0
    assert np.abs(fl["y"] - ref["y"]).max() <= 1e-6 * scale
    net.release()


# ------------------------------------------ the symmetric network N = K^T K
SDIL = (1, 3, 9)


# the benchmark's network: dilation 27 reaches past a 16-pixel block, into
# the clamped-address / zero-select padding path on every side
BENCH_DIL = (1, 3, 9, 27)


@pytest.mark.parametrize("dil,nb,nx,ny", [((1, 2), 1, 16, 16), (SDIL, 3, 37, 21), (SDIL, 2, 64, 9),
                                          (BENCH_DIL, 1, 64, 64), (BENCH_DIL, 2, 70, 83),
                                          (BENCH_DIL, 1, 128, 96)])
def test_sym_forward_matches_oracle(pk, ctx, dil, nb, nx, ny):
    net = pk.LearnedPrec(ctx, dil, symmetric=True)
    assert pk.lib().LearnedPrec_IsSymmetric(net.h) == 1
    assert net.nparams == O.lp_num_params_sym(dil)
    p = random_params(net.nparams, nb * nx, 0.2)
    net.set_params(p)
    net.set_output_scale(0.3)
    rng = np.random.default_rng(17)
    r = rng.standard_normal(nb * nx * ny)
    R, Z = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    R.set(r)
    assert net.apply(nb, nx, ny, R, Z) == 0
    z = Z.get()
    s = np.sqrt((r.reshape(nb, -1) ** 2).mean(axis=1))
    _, _, u = O.lp_lossgrad(dil, p, 0.3, nb, nx, ny, 1.0, r, r, emu=1, want_grad=False, sym=True)
    ref = u.reshape(nb, -1) * s[:, None]
    err = z.reshape(nb, -1) - ref
    assert np.abs(err).max() <= 1e-2 * np.abs(ref).max()
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * np.sqrt((ref ** 2).mean())
    net.release()


@pytest.mark.parametrize("seed,dil,nx,ny", [(0, SDIL, 33, 20), (1, SDIL, 33, 20), (0, BENCH_DIL, 70, 64),
                                             (1, BENCH_DIL, 64, 75)])
def test_sym_loss_and_gradient_match_oracle(pk, ctx, seed, dil, nx, ny):
    nb, ratio = 2, 40.0
    net = pk.LearnedPrec(ctx, dil, symmetric=True)
    if seed == 0:
        net.init_identity(0.05, 3)
        p = net.params()
    else:
        p = random_params(net.nparams, 12, 0.15)
        net.set_params(p)
    net.set_output_scale(1.0 / (1.0 + 4.0 * ratio))
    rng = np.random.default_rng(seed + 5)
    e = rng.standard_normal(nb * nx * ny)
    r = apply_A(e, nb, nx, ny, ratio)
    R, E = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    R.set(r)
    E.set(e)
    loss, g = net.loss_grad(nb, nx, ny, ratio, R, E)
    lref, gref, _ = O.lp_lossgrad(dil, p, 1.0 / (1.0 + 4.0 * ratio), nb, nx, ny, ratio, r, e, emu=1, sym=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-2 * np.linalg.norm(gref)
    # the unused bias slots get no gradient
    off = 0
    for l in range(len(dil)):
        off += 16 * (1 if l == 0 else 16) * 9
        assert np.all(g[off:off + 16] == 0.0)
        off += 16
    net.release()


def test_sym_network_is_symmetric_and_positive(pk, ctx):
    nx, ny = 48, 40
    net = pk.LearnedPrec(ctx, SDIL, symmetric=True)
    net.set_params(random_params(net.nparams, 21, 0.2))
    rng = np.random.default_rng(22)
    xs = [rng.standard_normal(nx * ny) for _ in range(2)]
    outs = []
    for x in xs:
        X, Z = pk.Vector(ctx, n=x.size), pk.Vector(ctx, n=x.size)
        X.set(x)
        assert net.apply(1, nx, ny, X, Z) == 0
        outs.append(Z.get())
    a, b = xs[0] @ outs[1], xs[1] @ outs[0]
    # symmetric up to the BF16 roundings of the activations
    assert abs(a - b) <= 2e-2 * np.linalg.norm(xs[0]) * np.linalg.norm(outs[1])
    for x, z in zip(xs, outs):
        assert x @ z > 0.0
    net.release()


def trained_sym_net(pk, ctx, ratio, n=64, nsamp=2048, steps=1500, batch=16, seed=0, lr=3e-3):
    R, E, _, rc = pk.training_set(ctx, nsamp, n, ratio, reltol=1e-12, seed=1234)
    assert rc == 0
    net = pk.LearnedPrec(ctx, SDIL, symmetric=True)
    net.init_identity(8e-4, seed)
    net.set_output_scale(1.0 / (1.0 + 4.0 * ratio))
    hist = net.train(nsamp, n, ratio, R, E, steps, batch, lr, seed)
    return net, hist


def test_sym_learned_krylov_on_the_implicit_path(pk, ctx):
    """The trained symmetric network inside FGMRES and inside PCG reaches the
    Jacobi-PCG solution with far fewer Krylov iterations; the untrained one
    (K^T K = I, i.e. Jacobi) does not gain."""
    side, dt = 128, 0.02
    dx = 1.0 / (side + 1)
    ratio = 1767732205903.0 / 4055673282236.0 * dt / dx ** 2
    net, hist = trained_sym_net(pk, ctx, ratio)
    assert np.all(np.isfinite(hist)) and hist[-50:].mean() < 0.5 * hist[:10].mean()
    ident = pk.LearnedPrec(ctx, SDIL, symmetric=True)
    ident.init_identity(0.0, 0)
    ref = dirk_run(pk, ctx, "jacobi", 0.06, dt, side, jtimes=True)
    fg = dirk_run(pk, ctx, "learned", 0.06, dt, side, net=net, jtimes=True)
    cg = dirk_run(pk, ctx, "learned", 0.06, dt, side, net=net, ls="cg", jtimes=True)
    idn = dirk_run(pk, ctx, "learned", 0.06, dt, side, net=ident, jtimes=True)
    scale = np.abs(ref["y"]).max()
    for o in (fg, cg, idn):
        assert o["flag"] == 0 and o["fallbacks"] == 0
        assert np.abs(o["y"] - ref["y"]).max() <= 1e-6 * scale
    nj, nf, nc, ni = (o["stats"]["nli"] for o in (ref, fg, cg, idn))
    assert nf < 0.35 * nj and nc < 0.35 * nj, (nj, nf, nc)
    assert nf < 0.35 * ni, (nf, ni)
    net.release()
    ident.release()


def test_sym_network_in_fused_pcg(pk, ctx):
    """The fused PCG calls the network through the preconditioner callback:
    same Newton iterations and solution as the N_Vector PCG with the same
    network (Krylov iterations within a few, the dot products being summed in
    another order)."""
    side, dt = 128, 0.02
    dx = 1.0 / (side + 1)
    ratio = 1767732205903.0 / 4055673282236.0 * dt / dx ** 2
    net, _ = trained_sym_net(pk, ctx, ratio, steps=600)
    ref = dirk_run(pk, ctx, "learned", 0.06, dt, side, net=net, ls="cg", jtimes=True, maxl=2000)
    fus = dirk_run(pk, ctx, "learned", 0.06, dt, side, net=net, ls="cg", jtimes=True, maxl=2000, fused=True)
    for o in (ref, fus):
        assert o["flag"] == 0 and abs(o["t"] - 0.06) < 1e-12
    assert fus["stats"]["nni"] == ref["stats"]["nni"]
    assert abs(fus["stats"]["nli"] - ref["stats"]["nli"]) <= max(3, 0.05 * ref["stats"]["nli"])
    scale = np.abs(ref["y"]).max()
    assert np.abs(fus["y"] - ref["y"]).max() <= 1e-7 * scale
    net.release()


# ------------------------ the two convolution kernels of the 16-channel layers
def _set_conv(pk, kind):
    assert pk.lib().LearnedPrec_SetConvKernel(kind) == 0


@pytest.mark.parametrize("symmetric,dil", [(True, BENCH_DIL), (True, (1, 2, 6, 18)), (True, (2, 3, 33, 1)),
                                           (True, (5,)), (False, DIL)])
@pytest.mark.parametrize("nb,nx,ny", [(1, 300, 211), (2, 70, 83), (3, 37, 21), (1, 128, 96), (1, 16, 130)])
def test_row_chain_convolution_is_bitwise_the_direct_one(pk, ctx, symmetric, dil, nb, nx, ny):
    """LearnedPrec_SetConvKernel(1) (input rows of one residue class mod the
    dilation staged in LDS), (2) (as 1, and two layers per kernel in the
    symmetric network's application) and (0) (every tap read through the
    caches) must give the same network output and training gradient bit for
    bit: the same operands in the same MFMA sequence.  Shapes cover strips
    narrower than one image, images narrower than a strip, fewer rows than
    the dilation (21 < 27) and a batch of images; the dilations cover layer
    pairs whose chain offsets are 1 and 3 in either order (1,3 / 9,27 and
    their adjoints), a pair of equal dilations and the first and last layer
    in one kernel (a one-layer network), and pairs that cannot be fused (2,3;
    33 > 32) next to one that can (1,1)."""
    net = pk.LearnedPrec(ctx, dil, symmetric=symmetric)
    p = random_params(net.nparams, 31 + nx, 0.2)
    net.set_params(p)
    net.set_output_scale(0.3)
    rng = np.random.default_rng(nx * 7 + ny)
    r = rng.standard_normal(nb * nx * ny)
    e = rng.standard_normal(nb * nx * ny)
    R, E, Z = (pk.Vector(ctx, n=r.size) for _ in range(3))
    R.set(r)
    E.set(e)
    out = {}
    try:
        for kind in (0, 1, 2):
            _set_conv(pk, kind)
            assert net.apply(nb, nx, ny, R, Z) == 0
            loss, g = net.loss_grad(nb, nx, ny, 40.0, R, E)
            out[kind] = (Z.get(), loss, g)
    finally:
        _set_conv(pk, 1)
    assert np.all(np.isfinite(out[1][0])) and np.abs(out[1][0]).max() > 0
    for kind in (1, 2):
        # equal bits, including the sign of zeros
        assert np.array_equal(out[0][0].view(np.uint64), out[kind][0].view(np.uint64)), kind
        assert out[0][1] == out[kind][1]
        assert np.array_equal(out[0][2], out[kind][2])
    net.release()


@pytest.mark.parametrize("rows,ahead", [(1, 1), (5, 3), (64, 2), (16, 1), (16, 3), (3, 2)])
def test_pair_kernel_tuning_does_not_change_the_result(pk, ctx, rows, ahead):
    """The two-layer kernel's chain rows per workgroup and rows in flight
    split and schedule the work only."""
    net = pk.LearnedPrec(ctx, BENCH_DIL, symmetric=True)
    net.set_params(random_params(net.nparams, 5, 0.2))
    rng = np.random.default_rng(rows)
    nx, ny = 200, 173
    R, Z = pk.Vector(ctx, n=nx * ny), pk.Vector(ctx, n=nx * ny)
    R.set(rng.standard_normal(nx * ny))
    out = []
    _set_conv(pk, 2)
    try:
        for r, a in ((16, 2), (rows, ahead)):
            assert pk.lib().LearnedPrec_SetPairTuning(r, a) == 0
            assert net.apply(1, nx, ny, R, Z) == 0
            out.append(Z.get())
    finally:
        assert pk.lib().LearnedPrec_SetPairTuning(16, 2) == 0
        _set_conv(pk, 1)
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))
    net.release()


@pytest.mark.parametrize("sym", [False, True])
def test_chain_rows_per_workgroup_do_not_change_the_result(pk, ctx, sym):
    """Small images get shorter row segments per workgroup (more
    workgroups); the segment length splits the work only: 16 rows everywhere
    (min tasks 1), the default, and 2 rows everywhere (min tasks 2^20) give
    the same bits, forward and in the training gradient."""
    dil = BENCH_DIL if sym else (1, 2, 5, 1)
    net = pk.LearnedPrec(ctx, dil, symmetric=sym)
    net.set_params(random_params(net.nparams, 8, 0.15))
    nb, nx, ny = 2, 300, 141
    rng = np.random.default_rng(3)
    e = rng.standard_normal(nb * nx * ny)
    r = apply_A(e, nb, nx, ny, 30.0)
    R, E, Z = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    R.set(r)
    E.set(e)
    out = []
    try:
        for m in (1, 1024, 1 << 20):
            assert pk.lib().LearnedPrec_SetChainTasks(m) == 0
            assert net.apply(nb, nx, ny, R, Z) == 0
            loss, g = net.loss_grad(nb, nx, ny, 30.0, R, E)
            out.append((Z.get(), loss, g))
    finally:
        assert pk.lib().LearnedPrec_SetChainTasks(1024) == 0
    for k in (1, 2):
        assert np.array_equal(out[0][0].view(np.uint64), out[k][0].view(np.uint64)), k
        assert out[0][1] == out[k][1] and np.array_equal(out[0][2], out[k][2]), k
    net.release()


def test_conv_kernel_switch_rejects_other_values(pk):
    assert pk.lib().LearnedPrec_SetChainTasks(0) == -1
    assert pk.lib().LearnedPrec_SetChainTasks((1 << 20) + 1) == -1
    assert pk.lib().LearnedPrec_SetConvKernel(3) == -1
    assert pk.lib().LearnedPrec_SetConvKernel(-1) == -1
    assert pk.lib().LearnedPrec_SetPairTuning(0, 2) == -1
    assert pk.lib().LearnedPrec_SetPairTuning(16, 4) == -1


def kron_A(nx, ny, rx, ry):
    """I - rx Dxx - ry Dyy, x the fast axis, zero boundary, from 1-D
    second-difference matrices."""
    import scipy.sparse as sp
    d2 = lambda n: sp.diags([np.ones(n - 1), -2.0 * np.ones(n), np.ones(n - 1)], [-1, 0, 1])
    return (sp.identity(nx * ny) - rx * sp.kron(sp.identity(ny), d2(nx))
            - ry * sp.kron(d2(ny), sp.identity(nx))).tocsr()


@pytest.mark.parametrize("rx,ry,n", [(60.0, 6.0, 32), (2.0, 45.0, 32), (30.0, 3.0, 27)])
def test_anisotropic_training_pairs_solve_their_system(pk, ctx, rx, ry, n):
    nsamp = 10
    R, E, iters, rc = pk.training_set(ctx, nsamp, n, rx, reltol=1e-10, seed=99, ratio_y=ry)
    assert rc == 0 and iters > 5
    r, e = R.get().reshape(nsamp, -1), E.get().reshape(nsamp, -1)
    A = kron_A(n, n, rx, ry)
    for i in range(nsamp):
        # the recursive residual meets 1e-10; the true one within rounding of it
        assert np.linalg.norm(A @ e[i] - r[i]) <= 1.05e-10 * np.linalg.norm(r[i]), i
    R.destroy()
    E.destroy()


@pytest.mark.parametrize("sym", [False, True])
def test_anisotropic_loss_and_gradient_match_oracle(pk, ctx, sym):
    nb, nx, ny, rx, ry = 2, 40, 24, 90.0, 3.0
    dil = SDIL if sym else DIL
    net = pk.LearnedPrec(ctx, dil, symmetric=sym)
    p = random_params(net.nparams, 21, 0.15 if sym else 0.2)
    net.set_params(p)
    scale = 1.0 / (1.0 + 2.0 * rx + 2.0 * ry)
    net.set_output_scale(scale)
    rng = np.random.default_rng(4)
    e = rng.standard_normal(nb * nx * ny)
    A = kron_A(nx, ny, rx, ry)
    r = np.concatenate([A @ e.reshape(nb, -1)[b] for b in range(nb)])
    R, E = pk.Vector(ctx, n=r.size), pk.Vector(ctx, n=r.size)
    R.set(r)
    E.set(e)
    loss, g = net.loss_grad(nb, nx, ny, rx, R, E, ratio_y=ry)
    lref, gref, _ = O.lp_lossgrad(dil, p, scale, nb, nx, ny, rx, r, e, emu=1, sym=sym, ratio_y=ry)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-2 * np.linalg.norm(gref)
    # the axes matter: the transposed operator gives another loss
    lt, _ = net.loss_grad(nb, nx, ny, ry, R, E, grad=False, ratio_y=rx)
    assert abs(lt - loss) > 1e-2 * abs(loss)
    net.release()


def test_network_trained_on_the_anisotropic_operator_beats_an_isotropic_one(pk, ctx):
    """On an anisotropic problem (kx/ky = 16) a symmetric network trained on
    the problem's own operator I - rx Dxx - ry Dyy takes far fewer PCG
    iterations than the same network trained on the isotropic operator at
    the mean ratio; both reach the Jacobi-PCG solution."""
    side, dt, kx, ky, tout = 96, 0.02, 4.0, 0.25, 0.06
    gamma_a = 1767732205903.0 / 4055673282236.0
    h2 = (side + 1) ** 2
    rx, ry = gamma_a * dt * kx * h2, gamma_a * dt * ky * h2
    nets = {}
    for tag, (ax, ay) in {"aniso": (rx, ry), "iso": (0.5 * (rx + ry), 0.5 * (rx + ry))}.items():
        R, E, _, rc = pk.training_set(ctx, 2048, 64, ax, reltol=1e-12, seed=1234, ratio_y=ay)
        assert rc == 0
        net = pk.LearnedPrec(ctx, SDIL, symmetric=True)
        net.init_identity(8e-4, 0)
        net.set_output_scale(1.0 / (1.0 + 2.0 * ax + 2.0 * ay))
        hist = net.train(2048, 64, ax, R, E, 800, 16, 3e-3, 0, ratio_y=ay)
        assert np.all(np.isfinite(hist))
        R.destroy()
        E.destroy()
        nets[tag] = net

    def solve(net):
        G = pk.Problem(ctx, pk.MMS_DIRICHLET, nx=side, ny=side, kx=kx, ky=ky)
        if net is not None:
            net.set_output_scale(1.0 / (1.0 + 2.0 * rx + 2.0 * ry))
            assert net.attach(G) == 0
        I = pk.Integrator(G, "dirk", G.initial(), ls="cg", maxl=20000,
                          precond="learned" if net is not None else True)
        I.analytic_jtimes()
        I.ARKodeSetOrder(3)
        I.ARKodeSetFixedStep(dt)
        I.ARKodeSetLSRelativeTolerance(1e-8)
        flag, t = I.evolve(tout)
        out = dict(flag=flag, y=I.y.get(), nli=I.stats()["nli"])
        I.free()
        G.destroy()
        return out

    ref, an, iso = solve(None), solve(nets["aniso"]), solve(nets["iso"])
    for o in (ref, an, iso):
        assert o["flag"] == 0
    scale = np.abs(ref["y"]).max()
    assert np.abs(an["y"] - ref["y"]).max() <= 1e-6 * scale
    assert np.abs(iso["y"] - ref["y"]).max() <= 1e-6 * scale
    assert an["nli"] < 0.8 * iso["nli"], (an["nli"], iso["nli"], ref["nli"])
    assert an["nli"] < 0.5 * ref["nli"], (an["nli"], ref["nli"])
    for net in nets.values():
        net.release()
