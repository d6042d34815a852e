"""Seeded synthetic inputs for the standard workloads of the timestep path.

Numpy generators (PCG64 via default_rng) replace Julia's MersenneTwister; the
same seed always gives the same bits, so the oracle and the GPU read
identical initial conditions.
"""
from __future__ import annotations

import math

import numpy as np


def diffusion_1d(N=64, seed=1234, kmax=8):
    """T0 = sum_{k=1..8} a_k sin(k x + phi_k), a_k ~ N(0,1)/k, phi_k ~ U(0, 2pi).
    Returns (grid, exact(t, kappa) callable)."""
    r = np.random.default_rng(seed)
    a = r.standard_normal(kmax) / np.arange(1, kmax + 1)
    phi = r.uniform(0, 2 * np.pi, kmax)
    x = 2 * np.pi * np.arange(N) / N
    k = np.arange(1, kmax + 1)

    def exact(t, kappa):
        return (a[None, :] * np.exp(-kappa * k[None, :] ** 2 * t) * np.sin(k[None, :] * x[:, None] + phi[None, :])).sum(1)

    return exact(0.0, 0.0), exact


def turbulence_2d(N=1024, seed=42, kpeak=10.0, energy=0.5, L=2 * math.pi):
    """Vorticity with random phases and shell spectrum E(k) ~ k^4 exp(-2 (k/kpeak)^2),
    normalised to total energy `energy` (E = 1/2 <u^2 + v^2>).  Grid array [ix, iy]."""
    r = np.random.default_rng(seed)
    k0 = 2 * math.pi / L
    kx = np.arange(N // 2 + 1) * k0
    ky = np.fft.fftfreq(N, 1.0 / N) * k0
    K = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    # per-mode |u_k|^2 ~ E(k)/k and |q_k|^2 = k^2 |u_k|^2 ~ k E(k)
    spec = np.where(K > 0, K * K ** 4 * np.exp(-2.0 * (K / (kpeak * k0)) ** 2), 0.0)
    qh = np.sqrt(spec) * np.exp(2j * np.pi * r.random(K.shape))
    q = np.fft.irfft(np.fft.ifft(qh, axis=1), n=N, axis=0)
    Q = np.fft.fft(np.fft.rfft(q, axis=0), axis=1)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    psi = np.where(k2 > 1e-12, -Q / np.where(k2 > 1e-12, k2, 1.0), 0.0)
    u = np.fft.irfft(np.fft.ifft(-1j * ky[None, :] * psi, axis=1), n=N, axis=0)
    v = np.fft.irfft(np.fft.ifft(1j * kx[:, None] * psi, axis=1), n=N, axis=0)
    E = 0.5 * np.mean(u ** 2 + v ** 2)
    return q * math.sqrt(energy / E)


def burgers_1d(N=256, seed=0, amp=0.1, kmax=4):
    """u0 = sin(x) + amp * N(0,1) low-mode noise (modes 1..kmax)."""
    r = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(N) / N
    u = np.sin(x)
    for k in range(1, kmax + 1):
        u = u + amp * r.standard_normal() * np.cos(k * x + r.uniform(0, 2 * np.pi))
    return u


def energy_enstrophy_from_coeff(C, N, L=2 * math.pi):
    """E = 1/2 <u^2+v^2>, Z = 1/2 <q^2> from vorticity coefficients (reference layout,
    unnormalised forward transform), by Parseval over the half spectrum."""
    C = np.asarray(C)
    k0 = 2 * math.pi / L
    kx = np.arange(C.shape[0]) * k0
    ky = np.fft.fftfreq(C.shape[1], 1.0 / C.shape[1]) * k0
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    w = np.full(C.shape[0], 2.0)
    w[0] = 1.0
    if N % 2 == 0:
        w[-1] = 1.0
    a2 = np.abs(C) ** 2 / float(N * N) ** 2
    Z = 0.5 * np.sum(w[:, None] * a2)
    E = 0.5 * np.sum(w[:, None] * np.where(k2 > 1e-12, a2 / np.where(k2 > 1e-12, k2, 1.0), 0.0))
    return E, Z
