"""The CPU oracle of the multigrid preconditioner, pinned before it checks
the GPU V-cycle: it is symmetric (PCG needs a symmetric preconditioner; a
pre-smoother that is not the adjoint of the post-smoother, or a restriction
that is not P^T/4, breaks it), and as a stationary iteration it contracts
every error mode (dense brute force of I - M A on small grids)."""
import numpy as np
import pytest

import oracle_lib as O


def dense(fn, n):
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        cols.append(fn(e))
    return np.array(cols).T


@pytest.mark.parametrize("nx,ny,kx,ky,inh", [(23, 19, 1.0, 1.0, 0), (30, 18, 1.0, 1.0, 0), (30, 17, 1.0, 0.5, 0),
                                             (15, 15, 2.0, 0.3, 0)])
@pytest.mark.parametrize("gamma", [1e-3, 2e-2])
def test_vcycle_is_symmetric_and_contracts(nx, ny, kx, ky, inh, gamma):
    P = O.OracleProblem(O.MMS, nx=nx, ny=ny, kx=kx, ky=ky)
    n = P.n
    z, nl = P.mg_apply(gamma, np.ones(n))
    assert nl >= 2
    M = dense(lambda e: P.mg_apply(gamma, e)[0], n)
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    A = dense(lambda e: e - gamma * P.rhs(0.0, e) + gamma * P.rhs(0.0, np.zeros(n)), n)
    rho = np.abs(np.linalg.eigvals(np.eye(n) - M @ A)).max()
    # point smoothing: isotropic problems contract fast, anisotropic ones less
    assert rho < (0.2 if kx == ky else 0.4), rho
    # and M is positive definite (PCG)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0.0


def test_single_level_is_symmetric_gauss_seidel():
    """A grid too small to coarsen: 4 red-black pairs then 4 black-red pairs."""
    P = O.OracleProblem(O.MMS, nx=5, ny=6, kx=1.0, ky=1.0)
    gamma = 0.01
    r = np.random.default_rng(0).standard_normal(P.n)
    z, nl = P.mg_apply(gamma, r)
    assert nl == 1
    n = P.n
    A = dense(lambda e: e - gamma * P.rhs(0.0, e) + gamma * P.rhs(0.0, np.zeros(n)), n)
    x = np.zeros(n)
    red = [k for k in range(n) if ((k % 5) + (k // 5)) % 2 == 0]
    black = [k for k in range(n) if k not in red]
    def sweep(idx):
        for k in idx:
            x[k] = (r[k] - (A[k] @ x - A[k, k] * x[k])) / A[k, k]
    for _ in range(4):
        sweep(red); sweep(black)
    for _ in range(4):
        sweep(black); sweep(red)
    np.testing.assert_allclose(z, x, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("nu,bound", [(1, 0.35), (3, 0.2)])
def test_vcycle_sweep_count_option(nu, bound):
    """--pfmg_nrelax: symmetric for any sweep count; fewer sweeps contract less."""
    P = O.OracleProblem(O.MMS, nx=23, ny=19, kx=1.0, ky=1.0)
    n, gamma = P.n, 2e-2
    M = dense(lambda e: P.mg_apply(gamma, e, 2, nu)[0], n)
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    A = dense(lambda e: e - gamma * P.rhs(0.0, e) + gamma * P.rhs(0.0, np.zeros(n)), n)
    rho = np.abs(np.linalg.eigvals(np.eye(n) - M @ A)).max()
    assert rho < bound, rho
    rho2 = np.abs(np.linalg.eigvals(np.eye(n) - dense(lambda e: P.mg_apply(gamma, e, 2, 2)[0], n) @ A)).max()
    assert (rho > rho2) if nu == 1 else (rho < rho2)


def test_vcycle_nonsymmetric_relax_type():
    """--pfmg_relax 3 (red then black on both sides of the coarse correction):
    no longer symmetric, still a contraction."""
    P = O.OracleProblem(O.MMS, nx=23, ny=19, kx=1.0, ky=0.5)
    n, gamma = P.n, 2e-2
    M = dense(lambda e: P.mg_apply(gamma, e, 3, 2)[0], n)
    assert np.abs(M - M.T).max() > 1e-6 * np.abs(M).max()
    A = dense(lambda e: e - gamma * P.rhs(0.0, e) + gamma * P.rhs(0.0, np.zeros(n)), n)
    assert np.abs(np.linalg.eigvals(np.eye(n) - M @ A)).max() < 0.4


def test_single_level_nonsymmetric_is_red_black_gauss_seidel():
    """Relax type 3 on a grid too small to coarsen: 8 red-black pairs."""
    P = O.OracleProblem(O.MMS, nx=5, ny=6, kx=1.0, ky=1.0)
    gamma = 0.01
    r = np.random.default_rng(1).standard_normal(P.n)
    z, nl = P.mg_apply(gamma, r, 3, 2)
    assert nl == 1
    n = P.n
    A = dense(lambda e: e - gamma * P.rhs(0.0, e) + gamma * P.rhs(0.0, np.zeros(n)), n)
    x = np.zeros(n)
    red = [k for k in range(n) if ((k % 5) + (k // 5)) % 2 == 0]
    black = [k for k in range(n) if k not in red]
    def sweep(idx):
        for k in idx:
            x[k] = (r[k] - (A[k] @ x - A[k, k] * x[k])) / A[k, k]
    for _ in range(8):
        sweep(red); sweep(black)
    np.testing.assert_allclose(z, x, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("nx,ny,kx,ky", [(24, 20, 1.0, 1.0), (32, 16, 1.0, 0.5), (16, 24, 2.0, 0.3)])
@pytest.mark.parametrize("gamma", [1e-3, 2e-2])
def test_periodic_vcycle_is_symmetric_and_contracts(nx, ny, kx, ky, gamma):
    """The reference's periodic problem (wrap-around transfers and smoothing):
    symmetric positive definite for the homogeneous (symmetric) operator and
    a contraction of its error modes."""
    P = O.OracleProblem(O.CEDA, nx=nx, ny=ny, kx=kx, ky=ky)
    n = P.n
    z, nl = P.mg_apply(gamma, np.ones(n))
    assert nl >= 2
    M = dense(lambda e: P.mg_apply(gamma, e)[0], n)
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0.0
    A = dense(lambda e: e - gamma * P.rhs(0.0, e), n)
    rho = np.abs(np.linalg.eigvals(np.eye(n) - M @ A)).max()
    assert rho < (0.25 if kx == ky else 0.5), rho


def test_periodic_inhomogeneous_vcycle_contracts():
    """With --inhomogeneous the reference's periodic operator is not symmetric
    (the wrap faces sit at different x); the cycle still contracts."""
    P = O.OracleProblem(O.CEDA, nx=24, ny=20, kx=1.0, ky=1.0, inhomogeneous=1)
    n, gamma = P.n, 1e-2
    M = dense(lambda e: P.mg_apply(gamma, e)[0], n)
    A = dense(lambda e: e - gamma * P.rhs(0.0, e), n)
    assert np.abs(np.linalg.eigvals(np.eye(n) - M @ A)).max() < 0.3


def test_periodic_single_level_is_symmetric_gauss_seidel():
    """A periodic grid whose half side is odd is not coarsened: 4 red-black
    pairs then 4 black-red pairs with wrap-around neighbours."""
    P = O.OracleProblem(O.CEDA, nx=6, ny=10, kx=1.0, ky=0.7)
    gamma = 0.05
    r = np.random.default_rng(2).standard_normal(P.n)
    z, nl = P.mg_apply(gamma, r)
    assert nl == 1
    n = P.n
    A = dense(lambda e: e - gamma * P.rhs(0.0, e), n)
    x = np.zeros(n)
    red = [k for k in range(n) if ((k % 6) + (k // 6)) % 2 == 0]
    black = [k for k in range(n) if k not in red]
    def sweep(idx):
        for k in idx:
            x[k] = (r[k] - (A[k] @ x - A[k, k] * x[k])) / A[k, k]
    for _ in range(4):
        sweep(red); sweep(black)
    for _ in range(4):
        sweep(black); sweep(red)
    np.testing.assert_allclose(z, x, rtol=1e-13, atol=1e-15)


def test_periodic_odd_side_is_rejected():
    P = O.OracleProblem(O.CEDA, nx=15, ny=16)
    assert P.mg_apply(1e-2, np.ones(P.n))[1] == -1
