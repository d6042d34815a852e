"""GPU transforms and the padded product through the C ABI vs the FP64 oracle.
Tolerances: 1e-12 relative on the FP64 path, 1e-5 on the FP32 path."""
import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu

RF, CH = 0, 1
TOL = {np.float64: 1e-12, np.float32: 1e-5}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kinds,sizes,bounds", [
    ([RF], [64], None),
    ([RF], [4096], None),
    ([RF, RF], [64, 32], None),
    ([RF, RF], [1024, 1024], None),
    ([RF, RF, RF], [32, 16, 8], None),
    ([RF, CH], [64, 17], [(0.0, 4.0), (0.0, 1.0)]),
    ([RF, CH], [256, 64], [(0.0, 4.0), (0.0, 1.0)]),
    ([RF, RF, CH], [32, 16, 9], [(0.0, 4.0), (0.0, 4.0), (0.0, 1.0)]),
])
def test_transform_parity(gpu, oracle, dtype, kinds, sizes, bounds):
    print(f"transform parity {kinds} {sizes} {dtype.__name__}", flush=True)
    r = np.random.default_rng(11)
    g = r.standard_normal(sizes)
    d = oracle.domain(kinds, sizes, bounds=bounds or [(0.0, 2 * np.pi)] * len(sizes))
    C_ref = oracle.forward_transform(d, g)
    p = gpu.plan(kinds, sizes, bounds=bounds, dtype=dtype)
    C = p.forward(g)
    assert C.shape == C_ref.shape
    assert relerr(C, C_ref) < TOL[dtype]
    G = p.backward(C_ref)
    assert relerr(G, oracle.backward_transform(d, C_ref)) < TOL[dtype]
    assert relerr(p.backward(C), g) < TOL[dtype] * 4
    p.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kinds,sizes", [
    ([RF, CH], [64, 5]),
    ([RF, CH], [64, 17]),
    ([RF, CH], [256, 65]),
    ([RF, CH], [96, 129]),
    ([RF, CH], [32, 769]),
    ([RF, RF, CH], [32, 16, 33]),
    ([RF, RF, CH], [64, 48, 257]),
])
def test_chebyshev_fft_form_parity(gpu, oracle, dtype, kinds, sizes):
    """ChebyshevT axes through one FFT of length 2(n-1) per column (method 2) and through the
    dense matrix (method 1), forward and backward, against the oracle's REDFT00 sums."""
    from tarang_amd import _lib
    print(f"chebyshev fft form {kinds} {sizes} {dtype.__name__}", flush=True)
    r = np.random.default_rng(17)
    g = r.standard_normal(sizes)
    bounds = [(0.0, 4.0)] * (len(sizes) - 1) + [(0.0, 1.0)]
    d = oracle.domain(kinds, sizes, bounds=bounds)
    C_ref = oracle.forward_transform(d, g)
    G_ref = oracle.backward_transform(d, C_ref)
    p = gpu.plan(kinds, sizes, bounds=bounds, dtype=dtype)
    for method in (2, 1):
        _lib.call("tarang_plan_set_cheb_method", p.h, method)
        C = p.forward(g)
        ef, eb = relerr(C, C_ref), relerr(p.backward(C_ref), G_ref)
        print(f"  method {method}: forward {ef:.2e} backward {eb:.2e}", flush=True)
        assert ef < TOL[dtype] and eb < TOL[dtype]
    p.close()


def test_chebyshev_method_checks(gpu):
    """The FFT form exists only where 2(n-1) is an LDS FFT length (n = 128: 254 = 2 127)."""
    from tarang_amd import _lib
    p = gpu.plan([RF, CH], [32, 128], bounds=[(0.0, 4.0), (0.0, 1.0)])
    with pytest.raises(_lib.TarangError):
        _lib.call("tarang_plan_set_cheb_method", p.h, 2)
    with pytest.raises(_lib.TarangError):
        _lib.call("tarang_plan_set_cheb_method", p.h, 3)
    _lib.call("tarang_plan_set_cheb_method", p.h, 1)
    p.close()


def test_transform_mode_ordering_bit_exact(gpu):
    # a single Fourier mode lands on exactly one coefficient index (FFT order on y)
    nx, ny = 16, 8
    x = 2 * np.pi * np.arange(nx) / nx
    y = 2 * np.pi * np.arange(ny) / ny
    p = gpu.plan([RF, RF], [nx, ny])
    for kx, ky in [(3, 2), (3, -2), (0, -3), (8, 0)]:
        g = np.cos(kx * x[:, None] + ky * y[None, :])
        C = p.forward(g)
        big = np.argwhere(np.abs(C) > 1e-9 * np.abs(C).max())
        iy = ky % ny
        expected = {(kx, iy)} if kx not in (0, nx // 2) else {(kx, iy), (kx, (-ky) % ny)}
        assert {tuple(b) for b in big} == expected, (kx, ky, big)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sizes", [[8], [64], [4096], [8, 8], [64, 32], [256, 256]])
def test_padded_product_parity(gpu, oracle, golden, dtype, sizes):
    print(f"padded product parity {sizes} {dtype.__name__}", flush=True)
    r = np.random.default_rng(12)
    a, b = r.standard_normal(sizes), r.standard_normal(sizes)
    d = oracle.domain([RF] * len(sizes), sizes)
    ref = oracle.padded_multiply(d, a, b)
    p = gpu.plan([RF] * len(sizes), sizes, dtype=dtype)
    out = p.multiply(a, b)
    e = relerr(out, ref)
    print(f"  rel err {e:.3e}", flush=True)
    assert e < TOL[dtype]
    p.close()


def test_padded_product_golden(gpu, golden):
    g = golden("product.npz")
    p = gpu.plan([RF, RF], [8, 8])
    out = p.multiply(g["p2_8x8_a"], g["p2_8x8_b"])
    assert relerr(out, g["p2_8x8_out"]) < 1e-12
    p1 = gpu.plan([RF], [16])
    assert relerr(p1.multiply(g["p1_16_a"], g["p1_16_b"]), g["p1_16_out"]) < 1e-12
