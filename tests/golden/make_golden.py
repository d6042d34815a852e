"""Writes the golden vectors under tests/golden/ (run from the repo root).

The reference is Julia + FFTW (neither is in this image), so its published
definitions are evaluated independently with numpy/scipy and closed forms:
  - FFTW forward/backward DFT conventions  -> numpy.fft (pocketfft)
  - FFTW REDFT00 (DCT-I)                   -> scipy.fft.dct(type=1)
  - Chebyshev transform convention (transform_chebyshev.jl:21-78, test_gpu_dct1_kernels_cpu.jl:35-50)
                                           -> T_k(z) = cos(k arccos z) sampled on z_j = -cos(pi j/(N-1))
  - 3/2 padded product                     -> direct sums (no FFT, no index maps): the real
                                              band-limited interpolants (even-N Nyquist split over
                                              +-N/2) evaluated on the padded grid, multiplied, the
                                              product's DFT taken at |k| <= N/2 and -N/2 folded on +N/2
  - pad/truncate index map (nonlinear_padding.jl:164-181; test_gpu_kernels_cpu.jl:128-171)
  - diffusion decay a_k exp(-kappa k^2 t)  (test_end_to_end_pde.jl:14-61)
  - vorticity RHS for q = cos x + cos 2y   -> F = (3/2) sin x sin 2y (hand derivation)
"""
import math
import os

import numpy as np
import scipy.fft as sf

OUT = os.path.dirname(os.path.abspath(__file__))


def rng(seed):
    return np.random.default_rng(seed)


def fft_vectors():
    r = rng(1)
    out = {}
    for n in [1, 2, 3, 5, 8, 12, 16, 30, 64, 96, 127, 128, 254, 384, 510, 1536]:
        x = r.standard_normal(n) + 1j * r.standard_normal(n)
        out[f"c{n}_in"] = x
        out[f"c{n}_fwd"] = np.fft.fft(x)
        out[f"c{n}_inv"] = np.fft.ifft(x) * n  # unnormalised inverse
        xr = r.standard_normal(n)
        out[f"r{n}_in"] = xr
        out[f"r{n}_rfft"] = np.fft.rfft(xr)
        c = np.fft.rfft(xr) + 1j * r.standard_normal(n // 2 + 1) * 0.25  # non-Hermitian DC/Nyquist parts
        out[f"r{n}_hc"] = c
        out[f"r{n}_irfft"] = np.fft.irfft(c, n)
        if n > 1:
            out[f"r{n}_dct1"] = sf.dct(xr, type=1)
    np.savez_compressed(os.path.join(OUT, "fft.npz"), **out)


def cheb_vectors():
    out = {}
    for N in [2, 3, 8, 9, 17, 32, 33, 64, 128]:
        z = -np.cos(np.pi * np.arange(N) / (N - 1))
        # closed form: T_k on the Gauss-Lobatto grid has coefficient vector e_k
        T = np.stack([np.cos(k * np.arccos(np.clip(z, -1, 1))) for k in range(N)], axis=1)  # T[j, k]
        out[f"T{N}"] = T
        out[f"z{N}"] = z
    # mapped interval example: f(z) = z^3 on [0, 1]: coefficients from numpy's chebfit on the same grid
    N = 16
    xi = -np.cos(np.pi * np.arange(N) / (N - 1))
    zz = 0.5 * (xi + 1.0)
    f = zz ** 3 - 2 * zz + 0.5
    c = np.polynomial.chebyshev.chebfit(xi, f, N - 1)
    out["poly_grid"] = f
    out["poly_coeff"] = c
    np.savez_compressed(os.path.join(OUT, "cheb.npz"), **out)


def _padded_len(n):
    m = math.ceil(1.5 * n)
    return m + (m % 2)


def _split_modes(C, N):
    """{k: c_k} of the real band-limited interpolant, the even-N Nyquist split over +-N/2."""
    k = np.fft.fftfreq(N, 1.0 / N).astype(int)
    out = {}
    for i, kk in enumerate(k):
        ks = [N // 2, -N // 2] if (N % 2 == 0 and abs(kk) == N // 2) else [kk]
        for q in ks:
            out[q] = out.get(q, 0) + C[i] / len(ks)
    return out


def _interp_product_1d(ga, gb, N):
    """Direct sums, no FFT: evaluate both interpolants on the padded grid (M points),
    multiply, take the DFT of the product at |k| <= N/2 and fold -N/2 onto +N/2."""
    M = _padded_len(N)
    ma = _split_modes(np.fft.fft(ga) / N, N)
    mb = _split_modes(np.fft.fft(gb) / N, N)
    xm = 2 * np.pi * np.arange(M) / M
    fa = sum(v * np.exp(1j * k * xm) for k, v in ma.items())
    fb = sum(v * np.exp(1j * k * xm) for k, v in mb.items())
    p = fa * fb
    x = 2 * np.pi * np.arange(N) / N
    g = np.zeros(N, complex)
    for k in range(-(N // 2), N // 2 + 1):
        Pk = np.sum(p * np.exp(-1j * k * xm)) / M
        g += Pk * np.exp(1j * k * x)
    return g.real


def _interp_product_2d(ga, gb):
    """2-D version of the direct-sum construction; arrays indexed [ix, iy]."""
    Nx, Ny = ga.shape
    Mx, My = _padded_len(Nx), _padded_len(Ny)

    def modes(g):
        C = np.fft.fft2(g) / (Nx * Ny)
        out = {}
        for i in range(Nx):
            mx = _split_modes(np.eye(Nx)[i], Nx)
            for j in range(Ny):
                my = _split_modes(np.eye(Ny)[j], Ny)
                for a, wa in mx.items():
                    if wa == 0:
                        continue
                    for b, wb in my.items():
                        if wb == 0:
                            continue
                        out[(a, b)] = out.get((a, b), 0) + C[i, j] * wa * wb
        return out

    xm = 2 * np.pi * np.arange(Mx) / Mx
    ym = 2 * np.pi * np.arange(My) / My

    def evalM(m):
        f = np.zeros((Mx, My), complex)
        for (a, b), v in m.items():
            f += v * np.exp(1j * a * xm)[:, None] * np.exp(1j * b * ym)[None, :]
        return f

    p = evalM(modes(ga)) * evalM(modes(gb))
    x = 2 * np.pi * np.arange(Nx) / Nx
    y = 2 * np.pi * np.arange(Ny) / Ny
    g = np.zeros((Nx, Ny), complex)
    for a in range(-(Nx // 2), Nx // 2 + 1):
        ex = np.exp(-1j * a * xm)
        for b in range(-(Ny // 2), Ny // 2 + 1):
            Pab = np.sum(p * ex[:, None] * np.exp(-1j * b * ym)[None, :]) / (Mx * My)
            g += Pab * np.exp(1j * a * x)[:, None] * np.exp(1j * b * y)[None, :]
    return g.real


def product_vectors():
    r = rng(2)
    out = {}
    for N in [8, 12, 16]:
        ga, gb = r.standard_normal(N), r.standard_normal(N)
        out[f"p1_{N}_a"], out[f"p1_{N}_b"] = ga, gb
        out[f"p1_{N}_out"] = _interp_product_1d(ga, gb, N)
    for (Nx, Ny) in [(8, 8), (12, 8), (8, 6)]:
        ga, gb = r.standard_normal((Nx, Ny)), r.standard_normal((Nx, Ny))
        out[f"p2_{Nx}x{Ny}_a"], out[f"p2_{Nx}x{Ny}_b"] = ga, gb
        out[f"p2_{Nx}x{Ny}_out"] = _interp_product_2d(ga, gb)
    # pad/truncate index map for N=8 -> M=12: padded index of each original index
    out["padmap_8_12"] = np.array([0, 1, 2, 3, 4, 9, 10, 11])
    np.savez_compressed(os.path.join(OUT, "product.npz"), **out)


def ivp_vectors():
    out = {}
    # config-1 style initial condition: T0 = sum_k a_k sin(k x + phi_k), k = 1..8, seed 1234
    r = rng(1234)
    a = r.standard_normal(8) / np.arange(1, 9)
    phi = r.uniform(0, 2 * np.pi, 8)
    out["diff_a"], out["diff_phi"] = a, phi
    N, kappa, t = 64, 0.01, 1.0
    x = 2 * np.pi * np.arange(N) / N
    k = np.arange(1, 9)
    out["diff_T0"] = (a[None, :] * np.sin(k[None, :] * x[:, None] + phi[None, :])).sum(1)
    out["diff_T1"] = (a[None, :] * np.exp(-kappa * k[None, :] ** 2 * t) * np.sin(k[None, :] * x[:, None] + phi[None, :])).sum(1)
    # vorticity RHS closed form
    Nx = Ny = 16
    X, Y = np.meshgrid(2 * np.pi * np.arange(Nx) / Nx, 2 * np.pi * np.arange(Ny) / Ny, indexing="ij")
    out["turb_q"] = np.cos(X) + np.cos(2 * Y)
    out["turb_F"] = 1.5 * np.sin(X) * np.sin(2 * Y)
    # Burgers RHS closed form: u = sin x -> -u u_x = -sin x cos x = -sin(2x)/2
    N = 32
    xb = 2 * np.pi * np.arange(N) / N
    out["burg_u"] = np.sin(xb)
    out["burg_F"] = -0.5 * np.sin(2 * xb)
    np.savez_compressed(os.path.join(OUT, "ivp.npz"), **out)


if __name__ == "__main__":
    fft_vectors()
    cheb_vectors()
    product_vectors()
    ivp_vectors()
    print("golden vectors written to", OUT)
