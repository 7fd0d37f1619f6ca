"""The CPU oracle of the learned preconditioner, pinned before it is used to
check the GPU kernels: its forward pass against an independent numpy
convolution, and its reverse-mode gradient against central differences of
its own exact-FP64 loss (a wrong tap flip, a dropped ReLU mask or a wrong
loss scaling fails one of them)."""
import numpy as np
import pytest

import oracle_lib as O


def np_forward(dil, params, out_scale, nb, nx, ny, r):
    """Independent restatement: per layer, a sum over the 9 shifted copies of
    the zero-padded input (numpy slicing), ReLU between layers."""
    x = r.reshape(nb, ny, nx)
    s = np.sqrt((x ** 2).mean(axis=(1, 2), keepdims=True))
    x = (x / s)[..., None]
    off = 0
    for l, d in enumerate(dil):
        cin = 1 if l == 0 else 16
        cout = 1 if l == len(dil) - 1 else 16
        W = params[off:off + cout * cin * 9].reshape(cout, cin, 3, 3).astype(np.float64)
        off += cout * cin * 9
        b = params[off:off + cout].astype(np.float64)
        off += cout
        xp = np.pad(x, ((0, 0), (d, d), (d, d), (0, 0)))
        out = np.zeros((nb, ny, nx, cout)) + b
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                sl = xp[:, d + dy * d:d + dy * d + ny, d + dx * d:d + dx * d + nx, :]
                out += np.einsum("byxc,oc->byxo", sl, W[:, :, dy + 1, dx + 1])
        x = out if l == len(dil) - 1 else np.maximum(out, 0.0)
    return out_scale * x[..., 0].reshape(-1)


def problem(seed, dil, nb=2, nx=7, ny=6):
    rng = np.random.default_rng(seed)
    n = O.lp_num_params(dil)
    params = (0.3 * rng.standard_normal(n)).astype(np.float32)
    r = rng.standard_normal(nb * nx * ny)
    e = rng.standard_normal(nb * nx * ny) + 0.5 * r
    return params, r, e


@pytest.mark.parametrize("dil", [(1, 1), (1, 2, 1), (1, 1, 2, 1)])
def test_oracle_forward_matches_numpy(dil):
    params, r, e = problem(1, dil)
    _, _, u = O.lp_lossgrad(dil, params, 0.7, 2, 7, 6, 3.0, r, e, emu=0, want_grad=False)
    ref = np_forward(dil, params, 0.7, 2, 7, 6, r)
    assert np.abs(u - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dil", [(1, 1), (1, 2, 1), (1, 1, 2, 1)])
def test_oracle_gradient_matches_finite_differences(dil):
    params, r, e = problem(2, dil)
    ratio = 5.0
    loss, g, _ = O.lp_lossgrad(dil, params, 0.4, 2, 7, 6, ratio, r, e, emu=0)
    p64 = params.astype(np.float64)
    fd = np.zeros(len(p64))
    for k in range(len(p64)):
        h = 1e-4 * max(1.0, abs(p64[k]))
        # the oracle takes FP32 parameters: step in a way FP32 represents exactly
        pp, pm = params.copy(), params.copy()
        pp[k] = np.float32(p64[k] + h)
        pm[k] = np.float32(p64[k] - h)
        hh = float(pp[k]) - float(pm[k])
        lp_, _, _ = O.lp_lossgrad(dil, pp, 0.4, 2, 7, 6, ratio, r, e, emu=0, want_grad=False)
        lm_, _, _ = O.lp_lossgrad(dil, pm, 0.4, 2, 7, 6, ratio, r, e, emu=0, want_grad=False)
        fd[k] = (lp_ - lm_) / hh
    scale = np.abs(g).max()
    # the loss is piecewise quadratic (ReLU); a step crossing a kink is rare
    bad = np.abs(fd - g) > 1e-7 * scale
    assert bad.mean() < 0.005, (np.abs(fd - g).max() / scale, bad.sum())


def test_oracle_bf16_emulation_is_a_small_perturbation():
    dil = (1, 1, 1)
    params, r, e = problem(3, dil)
    l0, g0, u0 = O.lp_lossgrad(dil, params, 0.5, 2, 7, 6, 2.0, r, e, emu=0)
    l1, g1, u1 = O.lp_lossgrad(dil, params, 0.5, 2, 7, 6, 2.0, r, e, emu=1)
    rel = np.abs(u1 - u0).max() / np.abs(u0).max()
    assert 0.0 < rel < 0.05
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.05


# ----------------------------------------------------- the symmetric network
def dense_conv(W, d, nx, ny):
    """The dense matrix of one zero-padded 3x3 convolution (dilation d) from
    [ny][nx][cin] to [ny][nx][cout], built entry by entry."""
    cout, cin = W.shape[:2]
    M = np.zeros((ny * nx * cout, ny * nx * cin))
    for y in range(ny):
        for x in range(nx):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    ys, xs = y + dy * d, x + dx * d
                    if 0 <= ys < ny and 0 <= xs < nx:
                        for co in range(cout):
                            for ci in range(cin):
                                M[(y * nx + x) * cout + co, (ys * nx + xs) * cin + ci] += W[co, ci, dy + 1, dx + 1]
    return M


def sym_problem(seed, dil, nb=2, nx=7, ny=6):
    rng = np.random.default_rng(seed)
    n = O.lp_num_params_sym(dil)
    params = (0.3 * rng.standard_normal(n)).astype(np.float32)
    r = rng.standard_normal(nb * nx * ny)
    e = rng.standard_normal(nb * nx * ny) + 0.5 * r
    return params, r, e


def sym_matrix(dil, params, nx, ny):
    """K^T K from the dense layer matrices and numpy's transpose (independent
    of the oracle's adjoint convolution)."""
    K = None
    off = 0
    for l, d in enumerate(dil):
        cin = 1 if l == 0 else 16
        W = params[off:off + 16 * cin * 9].reshape(16, cin, 3, 3).astype(np.float64)
        off += 16 * cin * 9 + 16
        M = dense_conv(W, d, nx, ny)
        K = M if K is None else M @ K
    return K.T @ K


@pytest.mark.parametrize("dil", [(1,), (1, 2), (1, 2, 3)])
def test_sym_oracle_forward_matches_dense_adjoint(dil):
    nx, ny, nb = 7, 6, 2
    params, r, e = sym_problem(4, dil, nb, nx, ny)
    assert O.lp_num_params_sym(dil) == sum(16 * (1 if l == 0 else 16) * 9 + 16 for l in range(len(dil)))
    _, _, u = O.lp_lossgrad(dil, params, 0.7, nb, nx, ny, 3.0, r, e, emu=0, want_grad=False, sym=True)
    M = sym_matrix(dil, params, nx, ny)
    for b in range(nb):
        x = r[b * nx * ny:(b + 1) * nx * ny]
        s = np.sqrt((x ** 2).mean())
        ref = 0.7 * (M @ (x / s))
        got = u[b * nx * ny:(b + 1) * nx * ny]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_sym_oracle_is_symmetric_and_positive():
    dil = (1, 2)
    nx, ny = 9, 8
    params, _, _ = sym_problem(5, dil, 1, nx, ny)
    rng = np.random.default_rng(6)
    outs = []
    xs = [rng.standard_normal(nx * ny) for _ in range(2)]
    for x in xs:
        _, _, u = O.lp_lossgrad(dil, params, 1.0, 1, nx, ny, 1.0, x, x, emu=0, want_grad=False, sym=True)
        outs.append(u * np.sqrt((x ** 2).mean()))  # undo the normalisation: N x
    a, b = xs[0] @ outs[1], xs[1] @ outs[0]
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))
    assert xs[0] @ outs[0] > 0 and xs[1] @ outs[1] > 0


@pytest.mark.parametrize("dil", [(1,), (1, 2), (1, 3, 2)])
def test_sym_oracle_gradient_matches_finite_differences(dil):
    params, r, e = sym_problem(7, dil)
    ratio = 5.0
    loss, g, _ = O.lp_lossgrad(dil, params, 0.4, 2, 7, 6, ratio, r, e, emu=0, sym=True)
    nw = [16 * (1 if l == 0 else 16) * 9 for l in range(len(dil))]
    # the bias slots are not used: zero gradient
    off = 0
    for w in nw:
        assert np.all(g[off + w:off + w + 16] == 0.0)
        off += w + 16
    rng = np.random.default_rng(8)
    ks = rng.choice(len(params), size=min(len(params), 160), replace=False)
    for k in ks:
        h = 1e-3 * max(1.0, abs(float(params[k])))
        pp, pm = params.copy(), params.copy()
        pp[k] = np.float32(params[k] + h)
        pm[k] = np.float32(params[k] - h)
        hh = float(pp[k]) - float(pm[k])
        lp_, _, _ = O.lp_lossgrad(dil, pp, 0.4, 2, 7, 6, ratio, r, e, emu=0, want_grad=False, sym=True)
        lm_, _, _ = O.lp_lossgrad(dil, pm, 0.4, 2, 7, 6, ratio, r, e, emu=0, want_grad=False, sym=True)
        fd = (lp_ - lm_) / hh
        # smooth (polynomial) loss: central differences to O(h^2)
        assert abs(fd - g[k]) <= 1e-5 * np.abs(g).max() + 1e-6 * abs(g[k]), (k, fd, g[k])


def test_sym_oracle_bf16_emulation_is_a_small_perturbation():
    dil = (1, 2)
    params, r, e = sym_problem(9, dil)
    l0, g0, u0 = O.lp_lossgrad(dil, params, 0.5, 2, 7, 6, 2.0, r, e, emu=0, sym=True)
    l1, g1, u1 = O.lp_lossgrad(dil, params, 0.5, 2, 7, 6, 2.0, r, e, emu=1, sym=True)
    rel = np.abs(u1 - u0).max() / np.abs(u0).max()
    assert 0.0 < rel < 0.05
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.05


def kron_A(nx, ny, rx, ry):
    """I - rx Dxx - ry Dyy on an nx*ny grid (x fastest), zero boundary, built
    from 1-D second-difference matrices (independently of the oracle's
    stencil loop)."""
    import scipy.sparse as sp
    d2 = lambda n: sp.diags([np.ones(n - 1), -2.0 * np.ones(n), np.ones(n - 1)], [-1, 0, 1])
    return (sp.identity(nx * ny) - rx * sp.kron(sp.identity(ny), d2(nx))
            - ry * sp.kron(d2(ny), sp.identity(nx))).tocsr()


@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("rx,ry", [(3.0, 3.0), (40.0, 0.7), (0.25, 9.0)])
def test_oracle_anisotropic_operator_orientation(sym, rx, ry):
    """With the output scale 0 the network contributes nothing (d = -e/s), so
    the loss is e^T A e / (e.r) per image: the operator alone, checked
    against the Kronecker construction with x the fast (contiguous) axis."""
    nb, nx, ny = 2, 9, 5
    dil = (1, 2, 1)
    rng = np.random.default_rng(7)
    n = O.lp_num_params_sym(dil) if sym else O.lp_num_params(dil)
    params = (0.3 * rng.standard_normal(n)).astype(np.float32)
    r = rng.standard_normal(nb * nx * ny)
    e = rng.standard_normal(nb * nx * ny) + 0.5 * r
    loss, _, _ = O.lp_lossgrad(dil, params, 0.0, nb, nx, ny, rx, r, e, emu=0, want_grad=False, sym=sym,
                               ratio_y=ry)
    A = kron_A(nx, ny, rx, ry)
    E, R = e.reshape(nb, -1), r.reshape(nb, -1)
    want = np.mean([E[b] @ (A @ E[b]) / (E[b] @ R[b]) for b in range(nb)])
    assert abs(loss - want) <= 1e-12 * abs(want), (loss, want)


def test_oracle_anisotropic_loss_is_affine_and_meets_the_isotropic_one():
    """For a fixed network and data the loss is affine in (rx, ry) (A is), and
    the anisotropic branch tends to the isotropic one on the diagonal."""
    dil = (1, 2, 1)
    params, r, e = problem(5, dil)
    L = lambda rx, ry: O.lp_lossgrad(dil, params, 0.6, 2, 7, 6, rx, r, e, emu=0, want_grad=False,
                                     ratio_y=ry)[0]
    a = 4.0
    assert abs(L(a, a * (1 + 1e-12)) - L(a, a)) <= 1e-9 * abs(L(a, a))
    for (x0, y0) in ((1.0, 3.0), (7.0, 0.5)):
        l00, l10, l20, l01, l02 = L(x0, y0), L(2 * x0, y0), L(3 * x0, y0), L(x0, 2 * y0), L(x0, 3 * y0)
        assert abs((l20 - l10) - (l10 - l00)) <= 1e-9 * abs(l10)
        assert abs((l02 - l01) - (l01 - l00)) <= 1e-9 * abs(l01)
        # the x slope and the y slope add up to the isotropic slope
        sx, sy = (l10 - l00) / x0, (l01 - l00) / y0
        assert abs(L(0.0, 0.0) + 2.0 * (sx + sy) - L(2.0, 2.0)) <= 1e-9 * abs(L(2.0, 2.0))
