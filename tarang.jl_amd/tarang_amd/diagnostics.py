"""Diagnostics of the trajectory checks (host side, numpy, coefficient arrays).

  energy / enstrophy   examples/2d_decaying_turbulence_gpu.jl:247-255 (E = 1/2 <u^2+v^2>,
                       Z = 1/2 <q^2> on the grid); see configs.energy_enstrophy_from_coeff
  Nusselt number       the reference has no implementation (only a comment in
                       examples/ivp/rayleigh_benard_2d.jl:24); the definition used here is
                       Nu = <uz b - kappa dz b>_V / (kappa dB / Lz), dB = b(0) - b(Lz),
                       with the x mean taken exactly (Parseval over the rfft modes) and
                       the z mean by Clenshaw-Curtis quadrature on the Gauss-Lobatto grid;
                       the conduction term is exact from the wall values of the x mean.
Coefficient arrays are in the reference layout: (Nx/2+1, Nz) complex, unnormalised
rfft along x, Chebyshev T coefficients along z.
"""
from __future__ import annotations

import numpy as np


def clenshaw_curtis_weights(n: int) -> np.ndarray:
    """Weights w_j on x_j = -cos(pi j/(n-1)), sum_j w_j f(x_j) ~ int_{-1}^{1} f (exact to degree n-1)."""
    if n == 1:
        return np.array([2.0])
    N = n - 1
    theta = np.pi * np.arange(n) / N
    w = np.zeros(n)
    for j in range(n):
        s = 0.0
        for k in range(1, N // 2 + 1):
            b = 1.0 if 2 * k == N else 2.0
            s += b * np.cos(2 * k * theta[j]) / (4 * k * k - 1)
        c = 1.0 if j in (0, N) else 2.0
        w[j] = c / N * (1 - s)
    return w


def chebyshev_to_grid(c: np.ndarray) -> np.ndarray:
    """Values at the Gauss-Lobatto points z_j = -cos(pi j/(N-1)) of sum_k c[..., k] T_k."""
    n = c.shape[-1]
    k = np.arange(n)
    j = np.arange(n)
    B = ((-1.0) ** k)[None, :] * np.cos(np.pi * np.outer(j, k) / max(n - 1, 1))
    return c @ B.T


def x_mean_of_product(F: np.ndarray, G: np.ndarray, nx: int) -> np.ndarray:
    """Mean over the x grid of f g for real f, g with unnormalised rfft coefficients F, G
    (axis 0 = kx); any trailing axes are kept."""
    w = np.full(F.shape[0], 2.0)
    w[0] = 1.0
    if nx % 2 == 0:
        w[nx // 2] = 1.0
    prod = (F * np.conj(G)).real
    return np.tensordot(w, prod, axes=(0, 0)) / (nx * nx)


def nusselt(b: np.ndarray, uz: np.ndarray, nx: int, Lz: float, kappa: float, dB: float) -> float:
    """Nu = <uz b - kappa dz b>_V / (kappa dB / Lz) from coefficient arrays (nx/2+1, nz)."""
    bz, wz = chebyshev_to_grid(b), chebyshev_to_grid(uz)       # (nxh, nz): x modes, z grid
    flux = x_mean_of_product(wz, bz, nx)                      # x mean of uz b at each z point
    wq = clenshaw_curtis_weights(b.shape[-1])
    conv = float(np.dot(wq, flux)) / 2.0                      # mean over z in [0, Lz]
    b0 = b[0].real / nx                                       # x-mean profile coefficients
    jump = float(np.sum(b0) - np.sum(b0 * (-1.0) ** np.arange(b0.size)))  # bbar(Lz) - bbar(0)
    return (conv - kappa * jump / Lz) / (kappa * dB / Lz)


def xy_mean_of_product(F: np.ndarray, G: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """Mean over the x-y grid of f g for real f, g with (nx/2+1, ny, ...) coefficients
    (rfft in x, full FFT in y, unnormalised); trailing axes are kept."""
    w = np.full(F.shape[0], 2.0)
    w[0] = 1.0
    if nx % 2 == 0:
        w[nx // 2] = 1.0
    prod = (F * np.conj(G)).real.sum(axis=1)
    return np.tensordot(w, prod, axes=(0, 0)) / (nx * nx * ny * ny)


def nusselt3(b: np.ndarray, uz: np.ndarray, nx: int, ny: int, Lz: float, kappa: float, dB: float) -> float:
    """Nu = <uz b - kappa dz b>_V / (kappa dB / Lz) from (nx/2+1, ny, nz) coefficient arrays."""
    bz, wz = chebyshev_to_grid(b), chebyshev_to_grid(uz)
    flux = xy_mean_of_product(wz, bz, nx, ny)
    wq = clenshaw_curtis_weights(b.shape[-1])
    conv = float(np.dot(wq, flux)) / 2.0
    b0 = b[0, 0].real / (nx * ny)
    jump = float(np.sum(b0) - np.sum(b0 * (-1.0) ** np.arange(b0.size)))
    return (conv - kappa * jump / Lz) / (kappa * dB / Lz)
