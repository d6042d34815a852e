"""Pins of the FP64 oracle against golden vectors, closed forms and invariants.

Each check is chosen so that a plausible mistake in the oracle fails it:
wrong DFT sign/normalisation (numpy goldens), C2R treatment of DC/Nyquist
imaginary parts (irfft goldens with non-Hermitian input), DCT-I definition
(scipy goldens), Chebyshev odd-degree flip and endpoint halving (T_k closed
forms), the 3/2 padded product's index maps and Nyquist split/fold (direct-sum
construction), the implicit diagonal operator (exact diffusion decay), the
advection RHS signs (closed form for q = cos x + cos 2y), and the RK/SBDF
tableaux (observed convergence order).
"""
import math

import numpy as np
import pytest

from tests.conftest import relerr


def test_fft_against_numpy(oracle, golden):
    g = golden("fft.npz")
    for n in [1, 2, 3, 5, 8, 12, 16, 30, 64, 96, 127, 128, 254, 384, 510, 1536]:
        x = g[f"c{n}_in"]
        assert relerr(oracle.fft(x, -1), g[f"c{n}_fwd"]) < 1e-14, n
        assert relerr(oracle.fft(x, +1), g[f"c{n}_inv"]) < 1e-14, n
        assert relerr(oracle.rfft(g[f"r{n}_in"]), g[f"r{n}_rfft"]) < 1e-14, n
        assert relerr(oracle.irfft(g[f"r{n}_hc"], n), g[f"r{n}_irfft"]) < 1e-14, n
        if n > 1:
            assert relerr(oracle.dct1(g[f"r{n}_in"]), g[f"r{n}_dct1"]) < 1e-14, n


def test_chebyshev_closed_forms(oracle, golden):
    g = golden("cheb.npz")
    for N in [2, 3, 8, 9, 17, 32, 33, 64, 128]:
        d = oracle.domain([oracle.CHEBYSHEVT], [N], bounds=[(-1.0, 1.0)])
        T = g[f"T{N}"]
        for k in range(N):
            # a Chebyshev domain is real throughout: no rfft axis
            c = oracle.forward_transform(d, T[:, k])
            e = np.zeros(N)
            e[k] = 1.0
            assert np.abs(c - e).max() < 1e-12, (N, k)
            back = oracle.backward_transform(d, e)
            assert np.abs(back - T[:, k]).max() < 1e-12, (N, k)


def test_chebyshev_mapped_polynomial(oracle, golden):
    g = golden("cheb.npz")
    d = oracle.domain([oracle.CHEBYSHEVT], [16], bounds=[(0.0, 1.0)])
    c = oracle.forward_transform(d, g["poly_grid"])
    assert relerr(c, g["poly_coeff"]) < 1e-12


def test_transforms_2d_3d_against_numpy(oracle):
    r = np.random.default_rng(5)
    # periodic 2-D: rfft along x, fft along y (transform_layout.jl:87-100)
    for (nx, ny) in [(16, 12), (8, 32)]:
        d = oracle.domain([0, 0], [nx, ny])
        g = r.standard_normal((nx, ny))
        C = oracle.forward_transform(d, g)
        assert relerr(C, np.fft.fft(np.fft.rfft(g, axis=0), axis=1)) < 1e-13
        assert relerr(oracle.backward_transform(d, C), g) < 1e-13
    # channel 2-D: rfft along x, Chebyshev along z; check against a sum of products of closed forms
    nx, nz = 16, 9
    d = oracle.domain([0, 1], [nx, nz], bounds=[(0, 2 * np.pi), (-1.0, 1.0)])
    x = 2 * np.pi * np.arange(nx) / nx
    z = -np.cos(np.pi * np.arange(nz) / (nz - 1))
    T3 = np.cos(3 * np.arccos(np.clip(z, -1, 1)))
    g = np.cos(2 * x)[:, None] * T3[None, :]
    C = oracle.forward_transform(d, g)
    E = np.zeros((nx // 2 + 1, nz), complex)
    E[2, 3] = nx / 2  # unnormalised rfft of cos(2x) puts N/2 at k=2
    assert np.abs(C - E).max() < 1e-12
    assert relerr(oracle.backward_transform(d, C), g) < 1e-13
    # 3-D periodic round trip and numpy parity
    d = oracle.domain([0, 0, 0], [8, 6, 4])
    g = r.standard_normal((8, 6, 4))
    C = oracle.forward_transform(d, g)
    assert relerr(C, np.fft.fftn(np.fft.rfft(g, axis=0), axes=(1, 2))) < 1e-13


def test_padded_product_direct_sums(oracle, golden):
    g = golden("product.npz")
    for N in [8, 12, 16]:
        d = oracle.domain([0], [N])
        out = oracle.padded_multiply(d, g[f"p1_{N}_a"], g[f"p1_{N}_b"])
        assert relerr(out, g[f"p1_{N}_out"]) < 1e-13, N
    for (Nx, Ny) in [(8, 8), (12, 8), (8, 6)]:
        d = oracle.domain([0, 0], [Nx, Ny])
        k = f"p2_{Nx}x{Ny}"
        out = oracle.padded_multiply(d, g[k + "_a"], g[k + "_b"])
        assert relerr(out, g[k + "_out"]) < 1e-13, k


def test_padded_product_nyquist_not_halved(oracle):
    # nonlinear_padding.jl:183-191: without the split/fold the Nyquist of a product
    # comes out at half amplitude.  cos(N/2 x) * 1 must return cos(N/2 x) exactly.
    N = 16
    x = 2 * np.pi * np.arange(N) / N
    d = oracle.domain([0], [N])
    out = oracle.padded_multiply(d, np.cos(N // 2 * x), np.ones(N))
    assert np.abs(out - np.cos(N // 2 * x)).max() < 1e-13


def test_padded_sizes_and_index_map(oracle, golden):
    # M = ceil(3N/2) rounded up to even (nonlinear_padding.jl:99-108)
    assert [oracle.padded_size(n) for n in (8, 10, 12, 64, 1024, 7)] == [12, 16, 18, 96, 1536, 12]
    # pad-then-truncate is the identity on the index map (test_gpu_kernels_cpu.jl:128-171)
    pm = golden("product.npz")["padmap_8_12"]
    N, M = 8, 12
    nn = N - N // 2 - 1
    ref = [i if i <= N // 2 else M - (N - i) for i in range(N)]
    assert list(pm) == ref and len(set(ref)) == N and nn == 3


def test_diffusion_exact_decay(oracle, golden):
    g = golden("ivp.npz")
    for ts in ["RK222", "RK443", "SBDF2"]:
        d = oracle.domain([0], [64])
        p = oracle.FourierIVP(d, oracle.PROB_DIFFUSION, ts, 0.01, 0.01)
        p.set_coeff(oracle.forward_transform(d, g["diff_T0"]))
        p.step(100)
        T1 = oracle.backward_transform(d, p.get_coeff())
        err = relerr(T1, g["diff_T1"])
        assert err < 2e-5, (ts, err)  # time-discretisation error only (rtol 1e-4 in test_end_to_end_pde.jl)
        assert abs(p.time - 1.0) < 1e-12


def test_turbulence_rhs_closed_form(oracle, golden):
    g = golden("ivp.npz")
    d = oracle.domain([0, 0], [16, 16])
    p = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", 0.0, 1e-3)
    F = p.rhs(oracle.forward_transform(d, g["turb_q"]))
    assert np.abs(oracle.backward_transform(d, F) - g["turb_F"]).max() < 1e-13


def test_burgers_rhs_closed_form(oracle, golden):
    g = golden("ivp.npz")
    d = oracle.domain([0], [32])
    p = oracle.FourierIVP(d, oracle.PROB_BURGERS1D, "RK222", 0.0, 1e-3)
    F = p.rhs(oracle.forward_transform(d, g["burg_u"]))
    assert np.abs(oracle.backward_transform(d, F) - g["burg_F"]).max() < 1e-13


def _burgers_run(oracle, ts, dt, tend, N=32):
    d = oracle.domain([0], [N])
    x = 2 * np.pi * np.arange(N) / N
    p = oracle.FourierIVP(d, oracle.PROB_BURGERS1D, ts, 0.05, dt)
    p.set_coeff(oracle.forward_transform(d, 0.5 * np.sin(x) + 0.2 * np.cos(2 * x)))
    p.step(int(round(tend / dt)))
    return p.get_coeff()


@pytest.mark.parametrize("ts,order", [("RK222", 2), ("RK443", 3), ("SBDF2", 2)])
def test_convergence_order(oracle, ts, order):
    # a mistyped tableau coefficient drops the observed order
    ref = _burgers_run(oracle, "RK443", 1e-4, 0.4)
    errs = [np.abs(_burgers_run(oracle, ts, dt, 0.4) - ref).max() for dt in (0.02, 0.01)]
    observed = math.log2(errs[0] / errs[1])
    assert abs(observed - order) < 0.35, (ts, errs, observed)


def _burgers_alternating(oracle, h, pairs, N=32):
    d = oracle.domain([0], [N])
    x = 2 * np.pi * np.arange(N) / N
    p = oracle.FourierIVP(d, oracle.PROB_BURGERS1D, "SBDF2", 0.05, h)
    p.set_coeff(oracle.forward_transform(d, 0.5 * np.sin(x) + 0.2 * np.cos(2 * x)))
    for _ in range(pairs):
        p.set_dt(h)
        p.step(1)
        p.set_dt(h / 2)
        p.step(1)
    return p.get_coeff(), p.time


def test_variable_step_sbdf2_second_order(oracle):
    # steps alternating h, h/2 (ratio w = 1/2 and 2): the variable-step weights keep order 2;
    # the constant-step weights (w = 1 at every step) would not converge at that order
    ref = _burgers_run(oracle, "RK443", 1e-4, 0.36)
    errs = []
    for h, pairs in ((0.02, 12), (0.01, 24)):
        c, t = _burgers_alternating(oracle, h, pairs)
        assert abs(t - 0.36) < 1e-12
        errs.append(np.abs(c - ref).max())
    observed = math.log2(errs[0] / errs[1])
    assert abs(observed - 2) < 0.35, (errs, observed)


def test_inviscid_turbulence_conserves_energy_enstrophy(oracle):
    # Galerkin-truncated 2-D Euler conserves E and Z up to time-discretisation error
    N = 32
    d = oracle.domain([0, 0], [N, N])
    r = np.random.default_rng(3)
    q = np.zeros((N, N))
    x = 2 * np.pi * np.arange(N) / N
    for kx in range(1, 5):
        for ky in range(1, 5):
            q += r.standard_normal() * np.cos(kx * x[:, None] + ky * x[None, :] + r.uniform(0, 6.28))
    p = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", 0.0, 2e-3)
    p.set_coeff(oracle.forward_transform(d, q))
    E0, Z0 = p.diagnostics()
    p.step(50)
    E1, Z1 = p.diagnostics()
    assert abs(E1 - E0) / E0 < 1e-7 and abs(Z1 - Z0) / Z0 < 1e-6, (E0, E1, Z0, Z1)
