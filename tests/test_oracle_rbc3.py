"""Pins of the oracle's 3-D Rayleigh-Benard channel (RealFourier x RealFourier x ChebyshevT).

  - a y-independent state evolves exactly as the (separately pinned) 2-D channel, for the
    RK stage loop and the multistep path (uy and every ky != 0 mode stay zero)
  - onset of an oblique mode (kx = ky = k_c/sqrt 2) at the textbook Ra_c = 1707.762
  - rotational invariance: modes with equal |k_h| have the same spectrum
  - SBDF2 is second order on the growth of an oblique eigenvector
  - system sizes 5N+9 at kx = ky = 0 and 5N+8 elsewhere
"""
import numpy as np
import pytest
import scipy.linalg as sl

RA_C, K_C = 1707.762, 3.1163


def zgrid(N, Lz):
    return Lz * (1 - np.cos(np.pi * np.arange(N) / (N - 1))) / 2


@pytest.mark.parametrize("ts", ["RK222", "SBDF2"])
def test_rbc3_y_independent_equals_2d(oracle, ts):
    O = oracle
    nx, ny, N, Lx, Ly, Lz = 16, 8, 12, 4.0, 3.0, 1.0
    kappa = nu = 1e5 ** -0.5
    r2 = O.RBC2D(nx, N, Lx, Lz, kappa, nu, 5e-3, ts)
    r3 = O.RBC3D(nx, ny, N, Lx, Ly, Lz, kappa, nu, 5e-3, ts)
    rng = np.random.default_rng(1)
    z = zgrid(N, Lz)
    w = (z * (Lz - z))[None, :]
    grids = {(O.RBC_B, O.RBC3_B): (Lz - z)[None, :] + 0.05 * rng.standard_normal((nx, N)) * w,
             (O.RBC_UX, O.RBC3_UX): 0.05 * rng.standard_normal((nx, N)) * w,
             (O.RBC_UZ, O.RBC3_UZ): 0.05 * rng.standard_normal((nx, N)) * w}
    for (v2, v3), g in grids.items():
        r2.set(v2, O.forward_transform(r2.d, g))
        r3.set(v3, O.forward_transform(r3.d, np.repeat(g[:, None, :], ny, axis=1)))
    r2.step(4)
    r3.step(4)
    for v2, v3 in grids:
        a = r2.get(v2) * ny          # the y transform is unnormalised
        b = r3.get(v3)
        assert np.abs(b[:, 0, :] - a).max() < 1e-14 * np.abs(a).max()
        assert np.abs(b[:, 1:, :]).max() < 1e-15 * np.abs(a).max()
    assert np.abs(r3.get(O.RBC3_UY)).max() < 1e-15


def oblique(O, Ra, N=24, m=1, n=1, ts="SBDF2", dt=1e-3, k=K_C, nx=8, ny=8):
    Lx = 2 * np.pi * m / (k / np.sqrt(2))
    Ly = 2 * np.pi * n / (k / np.sqrt(2))
    kap = Ra ** -0.5
    s = O.RBC3D(nx, ny, N, Lx, Ly, 1.0, kap, kap, dt, ts)
    mode = s.mode(m, n)
    M, L = s.matrices(mode)
    J = np.zeros_like(M)
    J[N + np.arange(N), 4 * N + np.arange(N)] = 1.0   # F_b = +uz for b0 = Lz - z
    w, V = sl.eig(-L + J, M)
    ok = np.isfinite(w)
    i = np.argmax(np.where(ok, w.real, -np.inf))
    return w[i], V[:, i], s, mode


def test_rbc3_sizes_and_oblique_onset(oracle):
    _, _, s, mode = oblique(oracle, RA_C)
    assert s.n(0) == 5 * 24 + 9 and s.n(mode) == 5 * 24 + 8
    lo = oblique(oracle, 0.99 * RA_C)[0].real
    hi = oblique(oracle, 1.01 * RA_C)[0].real
    at = oblique(oracle, RA_C)[0]
    assert lo < -1e-3 and hi > 1e-3 and abs(at.real) < 1e-5 and abs(at.imag) < 1e-10


def test_rbc3_rotational_invariance(oracle):
    """(kx, ky) = (3, 4) k0, (5, 0) k0 and (0, -5) k0 have |k_h| = 5 k0: same eigenvalues."""
    O = oracle
    s = O.RBC3D(16, 16, 12, 2 * np.pi, 2 * np.pi, 1.0, 1e-2, 1e-2, 1e-3, "SBDF2")
    spectra = []
    for m, n in ((3, 4), (5, 0), (0, 11)):   # n = 11 is ky = -5 in FFT order
        M, L = s.matrices(s.mode(m, n))
        w = sl.eig(-L, M, right=False)
        w = np.sort_complex(w[np.isfinite(w)])
        spectra.append(w)
    for w in spectra[1:]:
        assert w.shape == spectra[0].shape
        assert np.abs(w - spectra[0]).max() < 1e-9 * np.abs(spectra[0]).max()


def test_rbc3_sbdf2_growth_order(oracle):
    O = oracle
    N, Ra = 16, 3000.0
    errs = []
    for dt in (4e-2, 2e-2):
        sig, v, s, mode = oblique(O, Ra, N=N, dt=dt, k=np.pi)
        z = zgrid(N, 1.0)
        b = O.forward_transform(s.d, np.tile((1.0 - z)[None, None, :], (8, 8, 1)))
        amp = 1e-7 / np.abs(v[N:5 * N]).max()
        flds = {O.RBC3_B: b, O.RBC3_UX: np.zeros_like(b), O.RBC3_UY: np.zeros_like(b), O.RBC3_UZ: np.zeros_like(b)}
        for i, var in enumerate((O.RBC3_B, O.RBC3_UX, O.RBC3_UY, O.RBC3_UZ)):
            flds[var][1, 1, :] += amp * v[(1 + i) * N:(2 + i) * N]
            s.set(var, flds[var])
        x0 = np.concatenate([flds[v][1, 1] for v in (O.RBC3_B, O.RBC3_UX, O.RBC3_UY, O.RBC3_UZ)])
        T = 2.0
        s.step(int(round(T / dt)))
        x1 = np.concatenate([s.get(v)[1, 1] for v in (O.RBC3_B, O.RBC3_UX, O.RBC3_UY, O.RBC3_UZ)])
        r = np.vdot(x0, x1) / np.vdot(x0, x0)
        errs.append(abs(r - np.exp(sig.real * T)) / np.exp(sig.real * T))
    assert sig.real > 0.1 and errs[0] < 5e-4
    assert 3.0 < errs[0] / errs[1] < 5.2
