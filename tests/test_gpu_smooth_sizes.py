"""Fourier sizes that are not powers of two: 3 2^k and 5 2^k (3/2-padded lengths 9 2^j and
15 2^j, so the LDS FFT passes take radix 3 and 5 factors).  Transforms, the padded product,
the fused turbulence and Burgers kernels and the channel solvers against the FP64 oracle
(any length: mixed-radix Stockham), at 1e-12 (FP64) and 1e-5 (FP32)."""
import numpy as np
import pytest

from tests.conftest import relerr
from tests.test_gpu_channel import params, random_state
from tests.test_gpu_channel3d import kp, random_state3

pytestmark = pytest.mark.gpu

RF, CH = 0, 1
TOL = {np.float64: 1e-12, np.float32: 1e-5}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kinds,sizes,bounds", [
    ([RF], [96], None),
    ([RF], [1536], None),
    ([RF, RF], [96, 160], None),
    ([RF, RF], [640, 384], None),
    ([RF, CH], [80, 17], [(0.0, 4.0), (0.0, 1.0)]),
    ([RF, RF, CH], [48, 80, 9], [(0.0, 4.0), (0.0, 3.0), (0.0, 1.0)]),
])
def test_smooth_size_transforms(gpu, oracle, dtype, kinds, sizes, bounds):
    print(f"transform {kinds} {sizes} {dtype.__name__}", flush=True)
    r = np.random.default_rng(21)
    g = r.standard_normal(sizes)
    d = oracle.domain(kinds, sizes, bounds=bounds or [(0.0, 2 * np.pi)] * len(sizes))
    C_ref = oracle.forward_transform(d, g)
    p = gpu.plan(kinds, sizes, bounds=bounds, dtype=dtype)
    C = p.forward(g)
    assert C.shape == C_ref.shape
    assert relerr(C, C_ref) < TOL[dtype]
    assert relerr(p.backward(C_ref), oracle.backward_transform(d, C_ref)) < TOL[dtype]
    p.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sizes", [[96], [1536], [48, 80], [320, 192]])
def test_smooth_size_padded_product(gpu, oracle, dtype, sizes):
    print(f"padded product {sizes} {dtype.__name__}", flush=True)
    r = np.random.default_rng(22)
    a, b = r.standard_normal(sizes), r.standard_normal(sizes)
    d = oracle.domain([RF] * len(sizes), sizes)
    p = gpu.plan([RF] * len(sizes), sizes, dtype=dtype)
    e = relerr(p.multiply(a, b), oracle.padded_multiply(d, a, b))
    print(f"  rel err {e:.3e}", flush=True)
    assert e < TOL[dtype]
    p.close()


def vorticity(nx, ny, seed=8):
    """A few seeded Fourier modes of the vorticity on an nx x ny grid."""
    r = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(nx)[:, None] / nx
    y = 2 * np.pi * np.arange(ny)[None, :] / ny
    q = np.zeros((nx, ny))
    for kx, ky in [(1, 0), (0, 2), (3, -1), (2, 5), (7, 3), (11, -9)]:
        q += r.standard_normal() * np.cos(kx * x + ky * y + r.uniform(0, 2 * np.pi))
    return q


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx,ny", [(96, 160), (48, 80)])
def test_smooth_size_turbulence(gpu, oracle, dtype, nx, ny):
    """RHS and 3 RK443 steps of the vorticity equation on a non-square grid."""
    print(f"turbulence {nx} x {ny} {dtype.__name__}", flush=True)
    nu, dt = 1e-3, 2e-3
    d = oracle.domain([RF, RF], [nx, ny])
    C = oracle.forward_transform(d, vorticity(nx, ny))
    ref = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", nu, dt)
    s = gpu.fivp([nx, ny], [2 * np.pi, 2 * np.pi], 1, 1, nu, dt, dtype=dtype)
    assert relerr(s.rhs(C), ref.rhs(C)) < TOL[dtype]
    ref.set_coeff(C)
    s.set_state(C)
    ref.step(3)
    s.step(3)
    assert relerr(s.get_state(), ref.get_coeff()) < TOL[dtype] * 10
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("N", [96, 1536])
def test_smooth_size_burgers(gpu, oracle, dtype, N):
    d = oracle.domain([RF], [N])
    x = 2 * np.pi * np.arange(N) / N
    u = np.sin(x) + 0.1 * np.cos(3 * x + 0.3) + 0.05 * np.sin(7 * x)
    C = oracle.forward_transform(d, u)
    ref = oracle.FourierIVP(d, oracle.PROB_BURGERS1D, "RK222", 1e-2, 1e-3).rhs(C)
    s = gpu.fivp([N], [2 * np.pi], 2, 0, 1e-2, 1e-3, dtype=dtype)
    e = relerr(s.rhs(C), ref)
    print(f"burgers rhs N={N} {dtype.__name__} rel err {e:.3e}", flush=True)
    assert e < TOL[dtype]
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx", [96, 160])
def test_smooth_size_channel(gpu, oracle, dtype, nx):
    """2-D channel, 3 RK222 steps at nx = 96 and 160."""
    N, Lx, Lz, dt = 24, 4.0, 1.0, 5e-3
    print(f"channel nx={nx} {dtype.__name__}", flush=True)
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz, seed=13)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, "RK222")
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=0, dtype=dtype)
    for var, c in st.items():
        ref.set(var, c)
        s.set(var, c)
    ref.step(3)
    s.step(3)
    errs = [relerr(s.get(v), ref.get(v)) for v in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ)]
    print(f"  rel errs {errs}", flush=True)
    assert max(errs) < TOL[dtype], errs
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_smooth_size_channel3d(gpu, oracle, dtype):
    """3-D channel, 48 x 80 x 9, SBDF1 start and 2 SBDF2 steps."""
    nx, ny, N, Lx, Ly, Lz = 48, 80, 9, 4.0, 3.0, 1.0
    print(f"channel3d {nx}x{ny}x{N} {dtype.__name__}", flush=True)
    kappa, nu = kp()
    st, _ = random_state3(oracle, nx, ny, N, Lx, Ly, Lz, seed=14)
    ref = oracle.RBC3D(nx, ny, N, Lx, Ly, Lz, kappa, nu, 5e-3, "SBDF2")
    s = gpu.chan3(nx, ny, N, Lx, Ly, Lz, kappa, nu, 5e-3, stepper=2, dtype=dtype)
    for v, c in st.items():
        ref.set(v, c)
        s.set(v, c)
    ref.step(3)
    s.step(3)
    errs = [relerr(s.get(v), ref.get(v)) for v in (oracle.RBC3_B, oracle.RBC3_UX, oracle.RBC3_UY, oracle.RBC3_UZ)]
    print(f"  rel errs {errs}", flush=True)
    assert max(errs) < TOL[dtype], errs
    s.close()
