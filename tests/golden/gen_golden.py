"""Generate the golden fixtures that pin the CPU oracle.

Every fixture is computed here with numpy/scipy from closed-form mathematics,
independently of the oracle's code:

* ``periodic_fft.npz`` — exact semi-discrete solution of the homogeneous
  ceda-periodic problem (grid and IC of diffusion_2D.cpp:155-160 and
  initial.cpp:20-48): every Fourier mode of the periodic 5-point operator is an
  eigenvector, so u(t) = IFFT2(FFT2(u0) * exp(lambda_k t)) with
  lambda = -(4kx/dx^2) sin^2(theta_x/2) - (4ky/dy^2) sin^2(theta_y/2).
* ``mms_dst.npz`` — exact semi-discrete solution of the manufactured Dirichlet
  problem (diffusion_2D/README.md:11-32) with its time-dependent forcing,
  integrated exactly in the DST-I eigenbasis (Duhamel's formula with the
  integrals of exp(lambda (t - tau)) sin/cos(2 pi tau) in closed form).
* ``dirichlet_solve.npz`` — the exact solution x of (I - gamma L) x = b on a
  Dirichlet grid via the DST-I diagonalisation (checks the implicit solve).
* ``dense_ops.npz`` — tiny-grid dense matrices of the reference stencil,
  assembled entry by entry from diffusion.cpp:36-53 (including the
  inhomogeneous face positions and the periodic self-exchange of
  diffusion_2D.cpp:423-497), and of the Jacobi diagonal of
  preconditioner_jacobi.cpp:21-41.

Run from the repository root:  python tests/golden/gen_golden.py
"""
import os

import numpy as np
from scipy import fft as sfft

HERE = os.path.dirname(os.path.abspath(__file__))


def periodic_grid(nx, ny, xl=-np.pi, xu=np.pi, yl=-6.0, yu=6.0):
    dx = (xu - xl) / (nx - 1)
    dy = (yu - yl) / (ny - 1)
    x = xl + np.arange(nx) * dx
    y = yl + np.arange(ny) * dy
    return x, y, dx, dy


def periodic_ic(x, y):
    # initial.cpp:42-43
    return (1.0 + 0.3 * np.sin(2.0 * x))[None, :] / np.sqrt(5.5 * np.pi) * np.exp(-(y * y) / 5.5)[:, None]


def periodic_fft_fixture():
    out = {}
    for (nx, ny, kx, ky) in [(32, 24, 1.0, 0.5), (16, 20, 0.3, 2.0)]:
        x, y, dx, dy = periodic_grid(nx, ny)
        u0 = periodic_ic(x, y)  # shape (ny, nx), x fastest
        thx = 2.0 * np.pi * np.fft.fftfreq(nx)
        thy = 2.0 * np.pi * np.fft.fftfreq(ny)
        lam = (-(4.0 * kx / dx**2) * np.sin(thx / 2.0) ** 2)[None, :] + \
              (-(4.0 * ky / dy**2) * np.sin(thy / 2.0) ** 2)[:, None]
        uh = np.fft.fft2(u0)
        key = f"{nx}x{ny}"
        out[f"{key}_params"] = np.array([nx, ny, kx, ky])
        out[f"{key}_u0"] = u0
        for t in (0.05, 0.1):
            out[f"{key}_t{t}"] = np.real(np.fft.ifft2(uh * np.exp(lam * t)))
    np.savez(os.path.join(HERE, "periodic_fft.npz"), **out)


def dst_eigs(n, h, k):
    p = np.arange(1, n + 1)
    return -(4.0 * k / h**2) * np.sin(p * np.pi / (2.0 * (n + 1))) ** 2


def mms_tables(nx, ny, kx, ky):
    dx = 1.0 / (nx + 1)
    dy = 1.0 / (ny + 1)
    x = (np.arange(nx) + 1) * dx
    y = (np.arange(ny) + 1) * dy
    sx, sy = np.sin(np.pi * x) ** 2, np.sin(np.pi * y) ** 2
    cx = kx * (-2.0 * np.pi**2 * (np.cos(np.pi * x) ** 2 - np.sin(np.pi * x) ** 2))
    cy = ky * (-2.0 * np.pi**2 * (np.cos(np.pi * y) ** 2 - np.sin(np.pi * y) ** 2))
    return dx, dy, sx, sy, cx, cy


def dst2(a):
    return sfft.dstn(a, type=1, norm="ortho")


def mms_dst_fixture():
    out = {}
    for (nx, ny, kx, ky) in [(24, 20, 1.0, 0.7), (16, 16, 1.0, 1.0)]:
        dx, dy, sx, sy, cx, cy = mms_tables(nx, ny, kx, ky)
        lam = dst_eigs(nx, dx, kx)[None, :] + dst_eigs(ny, dy, ky)[:, None]
        S = sy[:, None] * sx[None, :]
        C = sy[:, None] * cx[None, :] + cy[:, None] * sx[None, :]
        u0 = S.copy()
        Sh, Ch, u0h = dst2(S), dst2(C), dst2(u0)
        w = 2.0 * np.pi
        key = f"{nx}x{ny}"
        out[f"{key}_params"] = np.array([nx, ny, kx, ky])
        out[f"{key}_u0"] = u0
        for t in (0.1, 0.3, 0.5):
            e = np.exp(lam * t)
            # int_0^t exp(lam (t-tau)) exp(i w tau) dtau = (exp(i w t) - exp(lam t)) / (i w - lam)
            J = (np.exp(1j * w * t) - e) / (1j * w - lam)
            Ic, Is = J.real, J.imag
            I1 = (e - 1.0) / lam
            # b(t) = T1 S + T2 C,  T1 = -pi sin(2 pi t),  T2 = (1 + cos(2 pi t))/2
            uh = e * u0h + Sh * (-np.pi) * Is + Ch * (0.5 * I1 + 0.5 * Ic)
            out[f"{key}_t{t}"] = dst2(uh)
    np.savez(os.path.join(HERE, "mms_dst.npz"), **out)


def dirichlet_solve_fixture():
    rng = np.random.default_rng(1234)
    nx, ny, kx, ky = 24, 20, 1.0, 1.0
    dx, dy = 1.0 / (nx + 1), 1.0 / (ny + 1)
    lam = dst_eigs(nx, dx, kx)[None, :] + dst_eigs(ny, dy, ky)[:, None]
    out = {"params": np.array([nx, ny, kx, ky])}
    for gamma in (1e-3, 1e-1):
        b = rng.standard_normal((ny, nx))
        x = dst2(dst2(b) / (1.0 - gamma * lam))
        out[f"g{gamma}_b"] = b
        out[f"g{gamma}_x"] = x
    np.savez(os.path.join(HERE, "dirichlet_solve.npz"), **out)


def coeff(k, inhom, s):
    return k * (1.0 + 0.99 * np.sin(s)) if inhom else k


def dense_periodic(nx, ny, kx, ky, inhom):
    """Matrix of the reference laplacian() with one rank (periodic wrap)."""
    x, y, dx, dy = periodic_grid(nx, ny)
    xl, yl = -np.pi, -6.0
    n = nx * ny
    M = np.zeros((n, n))
    for j in range(ny):
        Ds = coeff(ky, inhom, yl + (j - 0.5) * dy) / (dy * dy)
        Dn = coeff(ky, inhom, yl + (j + 0.5) * dy) / (dy * dy)
        for i in range(nx):
            Dw = coeff(kx, inhom, xl + (i - 0.5) * dx) / (dx * dx)
            De = coeff(kx, inhom, xl + (i + 0.5) * dx) / (dx * dx)
            r = j * nx + i
            M[r, r] += -((Dw + De) + (Ds + Dn))
            M[r, j * nx + (i - 1) % nx] += Dw
            M[r, j * nx + (i + 1) % nx] += De
            M[r, ((j - 1) % ny) * nx + i] += Ds
            M[r, ((j + 1) % ny) * nx + i] += Dn
    return M


def jacobi_periodic(nx, ny, kx, ky, inhom, gamma):
    x, y, dx, dy = periodic_grid(nx, ny)
    d = np.zeros((ny, nx))
    for j in range(ny):
        Ds = coeff(ky, inhom, j * dy) / (dy * dy)
        Dn = coeff(ky, inhom, (j + 1) * dy) / (dy * dy)
        for i in range(nx):
            Dw = coeff(kx, inhom, i * dx) / (dx * dx)
            De = coeff(kx, inhom, (i + 1) * dx) / (dx * dx)
            d[j, i] = 1.0 / (1.0 - gamma * (-((Dw + De) + (Ds + Dn))))
    return d


def dense_ops_fixture():
    out = {}
    for inhom in (0, 1):
        nx, ny, kx, ky = 5, 4, 1.3, 0.7
        out[f"periodic_inh{inhom}_params"] = np.array([nx, ny, kx, ky, inhom])
        out[f"periodic_inh{inhom}_M"] = dense_periodic(nx, ny, kx, ky, inhom)
        out[f"periodic_inh{inhom}_jacobi"] = jacobi_periodic(nx, ny, kx, ky, inhom, 0.01)
    np.savez(os.path.join(HERE, "dense_ops.npz"), **out)


if __name__ == "__main__":
    periodic_fft_fixture()
    mms_dst_fixture()
    dirichlet_solve_fixture()
    dense_ops_fixture()
    print("fixtures written to", HERE)
