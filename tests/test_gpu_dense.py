"""GPU Chebyshev grid derivative and batched complex LU vs the oracle.

Derivative tolerance: max |gpu - oracle| <= tol * max |oracle|, tol = 1e-12 on
FP64 data and 1e-5 on FP32 data.  The device computes in FP64 for both storage
types (FP32 data are widened on load and rounded once on store), so on FP32
data it is compared with the oracle run on the same FP32-rounded input.  Smooth
fields at full size carry the ~n^(2p) eps rounding of any order-p spectral
derivative (the reference's DCT-I chain included); that test states its bars.
LU: pivots bit-exact against the oracle (the same zgetf2 pivot rule), factors
and solutions within 1e-12 relative.
"""
import ctypes

import numpy as np
import pytest
import torch

from tests.conftest import relerr
from tests.gpu_helpers import from_dev, ptr, to_dev

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}


def logical(t):
    return t.permute(*reversed(range(t.dim())))


def gl(n, a=-1.0, b=1.0):
    x = -np.cos(np.pi * np.arange(n) / (n - 1))
    return (b - a) / 2 * x + (b + a) / 2


def run_deriv(gpu, a, axis, order, scale, dtype, inplace=False):
    from tarang_amd.derivatives import gpu_chebyshev_deriv_
    cplx = np.iscomplexobj(a)
    tdt = {(np.float64, False): torch.float64, (np.float32, False): torch.float32,
           (np.float64, True): torch.complex128, (np.float32, True): torch.complex64}[(dtype, cplx)]
    src = to_dev(a, tdt, gpu.torch_device)
    dst = src if inplace else torch.empty_like(src)
    assert gpu_chebyshev_deriv_(logical(dst), logical(src), axis, order, scale) is True
    return from_dev(dst)


CASES = [
    ((16,), 0, 1), ((16,), 0, 2), ((12, 12), 0, 1), ((12, 12), 1, 1),
    ((40, 24, 33), 0, 1), ((40, 24, 33), 1, 2), ((40, 24, 33), 2, 1), ((40, 24, 33), 2, 3),
    ((257, 6), 0, 1), ((5, 300, 3), 1, 2), ((70, 129), 1, 1), ((3, 65), 1, 1), ((2, 2, 2), 2, 1),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape,axis,order", CASES)
def test_cheb_deriv_parity(gpu, oracle, dtype, cplx, shape, axis, order):
    print(f"cheb deriv {shape} axis {axis} order {order} {dtype.__name__} complex={cplx}", flush=True)
    r = np.random.default_rng(hash((shape, axis, order)) % 2**32)
    a = r.uniform(-1, 1, shape)
    if cplx:
        a = a + 1j * r.uniform(-1, 1, shape)
    a = a.astype(np.complex64 if cplx else np.float32) if dtype == np.float32 else a
    scale = 2.0 / 3.0
    got = run_deriv(gpu, a, axis, order, scale, dtype)
    ref = oracle.cheb_deriv(a.astype(np.complex128 if cplx else np.float64), axis, order, scale)
    assert relerr(got, ref) < TOL[dtype]


@pytest.mark.parametrize("n", [33, 200])
def test_cheb_deriv_in_place(gpu, oracle, n):
    print(f"cheb deriv in place n={n}", flush=True)
    r = np.random.default_rng(n)
    a = r.standard_normal((7, n, 5))
    got = run_deriv(gpu, a, 1, 2, 1.0, np.float64, inplace=True)
    assert relerr(got, oracle.cheb_deriv(a, 1, 2, 1.0)) < 1e-12


def test_cheb_deriv_degenerate(gpu):
    print("cheb deriv degenerate", flush=True)
    a = np.full((4, 1, 3), 2.5)
    assert np.all(run_deriv(gpu, a, 1, 1, 1.0, np.float64) == 0)  # n = 1 -> zeros
    b = np.random.default_rng(0).standard_normal((9, 8))
    assert np.array_equal(run_deriv(gpu, b, 0, 0, 5.0, np.float64), b)  # order 0 -> copy


def test_cheb_deriv_reference_polynomials(gpu):
    """test/test_derivatives_polynomial.jl:51-86,144-175 (analytic oracle)."""
    print("cheb deriv reference polynomials", flush=True)
    f = lambda z: 3 * z**4 - 2 * z**2 + z - 0.5
    z = gl(16)
    assert np.abs(run_deriv(gpu, f(z), 0, 1, 1.0, np.float64) - (12 * z**3 - 4 * z + 1)).max() < 1e-12
    assert np.abs(run_deriv(gpu, f(z), 0, 2, 1.0, np.float64) - (36 * z**2 - 4)).max() < 1e-11
    z = gl(20, 0.0, 3.0)
    assert relerr(run_deriv(gpu, f(z), 0, 1, 2 / 3, np.float64), 12 * z**3 - 4 * z + 1) < 1e-12
    assert relerr(run_deriv(gpu, f(z), 0, 2, 2 / 3, np.float64), 36 * z**2 - 4) < 1e-11
    x, zz = np.meshgrid(gl(12), gl(12), indexing="ij")
    g = x**3 * zz + 2 * x * zz**2
    assert np.abs(run_deriv(gpu, g, 0, 1, 1.0, np.float64) - (3 * x**2 * zz + 2 * zz**2)).max() < 1e-12
    assert np.abs(run_deriv(gpu, g, 1, 1, 1.0, np.float64) - (x**3 + 4 * x * zz)).max() < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cheb_deriv_full_size(gpu, oracle, dtype):
    """256 x 256 x 128 field (the 3-D channel grid), z in (0, 1): sin(x) cos(2y) p(z)
    with p of degree 7, along z, orders 1 and 2.  FP64: against the oracle and the
    analytic derivative within 1e-11 (order 1) and 5e-9 (order 2) relative -- a
    smooth field's order-p derivative carries ~n^(2p) eps of rounding in any
    method (the reference's own FP64 DCT-I chain is off by 5.5e-10 at order 2
    here).  FP32 data: against the oracle on the same FP32 input within 1e-5."""
    print(f"cheb deriv full size {dtype.__name__}", flush=True)
    nx, ny, nz = 256, 256, 128
    x = 2 * np.pi * np.arange(nx) / nx
    y = 2 * np.pi * np.arange(ny) / ny
    z = gl(nz, 0.0, 1.0)
    s = 2 * z - 1
    p = s**7 - 3 * s**4 + s
    dp = (7 * s**6 - 12 * s**3 + 1) * 2.0
    d2p = (42 * s**5 - 36 * s**2) * 4.0
    xy = np.sin(x)[:, None, None] * np.cos(2 * y)[None, :, None]
    a = (xy * p[None, None, :]).astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    src = to_dev(a, tdt, gpu.torch_device)
    dst = torch.empty_like(src)
    from tarang_amd.derivatives import gpu_chebyshev_deriv_
    for order, exact in ((1, dp), (2, d2p)):
        gpu_chebyshev_deriv_(logical(dst), logical(src), 2, order, 2.0)
        got = from_dev(dst).astype(np.float64)
        ref = oracle.cheb_deriv(a.astype(np.float64), 2, order, 2.0)
        tol = TOL[dtype] if dtype == np.float32 else (1e-11 if order == 1 else 5e-9)
        assert relerr(got, ref) < tol, order
        if dtype == np.float64:
            assert relerr(got, xy * exact[None, None, :]) < tol, order


def test_cheb_deriv_c_abi_errors(gpu):
    from tarang_amd import _lib
    print("cheb deriv C-ABI errors", flush=True)
    L = _lib.load()
    src = torch.zeros(8, device=gpu.torch_device, dtype=torch.float64)
    shape = (ctypes.c_int * 1)(8)
    assert L.tarang_cheb_deriv(gpu.ctx, ptr(src), ptr(src), 1, shape, 0, -1, 1.0, _lib.F64, 0) == _lib.ERR_INVALID
    assert L.tarang_cheb_deriv(gpu.ctx, ptr(src), ptr(src), 1, shape, 1, 1, 1.0, _lib.F64, 0) == _lib.ERR_INVALID
    assert L.tarang_cheb_deriv(gpu.ctx, ptr(src), ptr(src), 1, shape, 0, 1, 1.0, 7, 0) == _lib.ERR_INVALID


# ---------------------------------------------------------------------------
# batched LU
# ---------------------------------------------------------------------------
def well_conditioned_batch(n, nmodes):
    """test/test_batched_dense_lu.jl:18-25 (0-based indices)."""
    i = np.arange(1, n + 1)[:, None, None]
    j = np.arange(1, n + 1)[None, :, None]
    m = np.arange(1, nmodes + 1)[None, None, :]
    off = 0.1 * (i - j) + 1j * 0.05 * m
    return np.where(i == j, (n + m) + 0.5j, off).astype(np.complex128)


def lu_on_device(gpu, A):
    from tarang_amd.matsolvers import BatchedDenseLU
    At = logical(to_dev(A, torch.complex128, gpu.torch_device))
    return BatchedDenseLU(At)


def test_batched_lu_reference_cases(gpu):
    """test/test_batched_dense_lu.jl:27-65."""
    from tarang_amd.matsolvers import batched_factor_, batched_solve_, device_array
    print("batched LU reference cases", flush=True)
    n, nm = 6, 4
    A = well_conditioned_batch(n, nm)
    i = np.arange(1, n * nm + 1)
    B = (i * 0.3 + 1j * i * 0.7).reshape(nm, n).T
    expected = np.stack([np.linalg.solve(A[:, :, m], B[:, m]) for m in range(nm)], axis=1)
    s = lu_on_device(gpu, A)
    batched_factor_(s)
    X = device_array((n, nm), torch.complex128, gpu.index)
    Bt = logical(to_dev(B, torch.complex128, gpu.torch_device))
    batched_solve_(X, s, Bt)
    assert relerr(X.cpu().numpy(), expected) < 1e-12
    A[:, :, 2] = 0
    s = lu_on_device(gpu, A)
    with pytest.raises(RuntimeError, match=r"mode\(s\) \[3\] of 4"):
        batched_factor_(s)
    assert not s.factored
    with pytest.raises(RuntimeError, match="before batched_factor"):
        batched_solve_(X, s, Bt)


@pytest.mark.parametrize("n,nm", [(1, 3), (17, 8), (64, 5), (135, 12), (263, 6), (519, 2)])
def test_batched_lu_parity(gpu, oracle, n, nm):
    from tarang_amd.matsolvers import batched_factor_, batched_solve_, device_array
    print(f"batched LU parity n={n} nmodes={nm}", flush=True)
    r = np.random.default_rng(n * 100 + nm)
    A = r.standard_normal((n, n, nm)) + 1j * r.standard_normal((n, n, nm))
    A += 2 * np.sqrt(n) * np.eye(n)[:, :, None] * r.uniform(-1, 1, (1, 1, nm))  # vary the pivot pattern
    B = r.standard_normal((n, nm)) + 1j * r.standard_normal((n, nm))
    LU_ref, ipiv_ref, info_ref = oracle.batched_lu_factor(A)
    X_ref = oracle.batched_lu_solve(LU_ref, ipiv_ref, B)
    s = lu_on_device(gpu, A)
    batched_factor_(s)
    assert np.all(s.info == 0) and np.all(info_ref == 0)
    assert np.array_equal(s.pivots.cpu().numpy(), ipiv_ref)
    assert relerr(s.A.cpu().numpy(), LU_ref) < 1e-12
    X = device_array((n, nm), torch.complex128, gpu.index)
    batched_solve_(X, s, logical(to_dev(B, torch.complex128, gpu.torch_device)))
    assert relerr(X.cpu().numpy(), X_ref) < 1e-12
    # X aliasing B
    Bt = logical(to_dev(B, torch.complex128, gpu.torch_device))
    batched_solve_(Bt, s, Bt)
    assert relerr(Bt.cpu().numpy(), X_ref) < 1e-12


def test_batched_lu_zero_pivot_mid_factorization(gpu, oracle):
    """A mode with a zero column: its pivot is exactly zero when elimination reaches
    it (updates subtract multiples of zero), so info names that column (1-based) as
    LAPACK does, and the call raises naming the mode."""
    from tarang_amd.matsolvers import batched_factor_
    print("batched LU zero pivot", flush=True)
    n, nm = 20, 3
    r = np.random.default_rng(5)
    A = r.standard_normal((n, n, nm)) + 0j
    A[:, 7, 1] = 0.0  # column 8 of mode 2
    _, _, info_ref = oracle.batched_lu_factor(A)
    s = lu_on_device(gpu, A)
    with pytest.raises(RuntimeError, match=r"mode\(s\) \[2\] of 3"):
        batched_factor_(s)
    assert info_ref[1] == 8
