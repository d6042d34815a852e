"""Pins of the oracle's Chebyshev grid derivative and batched complex LU.

Derivative (oracle.cheb_deriv, restating derivatives_polynomial.jl:326-421):
  - the reference's own analytic cases (test/test_derivatives_polynomial.jl:51-86,
    144-175: f = 3z^4 - 2z^2 + z - 0.5 on [-1, 1] and on (0, 3), g = x^3 z + 2 x z^2
    along each axis of a 12 x 12 grid) at the reference's tolerances and far below;
  - T_k'(x) = k U_{k-1}(x) at the Gauss-Lobatto points for every k < N (a wrong
    recurrence index, endpoint weight or grid direction fails it);
  - the closed-form collocation matrix (Trefethen, Spectral Methods in MATLAB,
    ch. 6) including its corner entry -(2(N-1)^2 + 1)/6 on the ascending grid;
  - n <= 1 gives zeros, order 0 is the identity, complex = Re + i Im.
Batched LU (oracle.batched_lu_factor / batched_lu_solve, restating LAPACK
zgetf2/zgetrs as BatchedDenseLU calls them, batched_matsolvers.jl:78-123):
  - pivots identical to scipy.linalg.lu_factor (LAPACK zgetrf), factors equal
    to roundoff, solutions equal to numpy.linalg.solve;
  - the reference's well-conditioned batch (test/test_batched_dense_lu.jl:18-42)
    at its rtol 1e-12, and its singular-mode case (:45-58) giving info != 0
    exactly at mode 3.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from tests.conftest import relerr


def gl(n, a=-1.0, b=1.0):
    x = -np.cos(np.pi * np.arange(n) / (n - 1))
    return (b - a) / 2 * x + (b + a) / 2


def fpoly(z):
    return 3 * z**4 - 2 * z**2 + z - 0.5


def test_cheb_deriv_reference_polynomials(oracle):
    z = gl(16)
    d1 = oracle.cheb_deriv(fpoly(z), 0, 1, 1.0)
    d2 = oracle.cheb_deriv(fpoly(z), 0, 2, 1.0)
    assert np.abs(d1 - (12 * z**3 - 4 * z + 1)).max() < 1e-12
    assert np.abs(d2 - (36 * z**2 - 4)).max() < 1e-11
    z = gl(20, 0.0, 3.0)  # shifted bounds: chain-rule scale 2/(b-a)
    d1 = oracle.cheb_deriv(fpoly(z), 0, 1, 2.0 / 3.0)
    d2 = oracle.cheb_deriv(fpoly(z), 0, 2, 2.0 / 3.0)
    assert relerr(d1, 12 * z**3 - 4 * z + 1) < 1e-12
    assert relerr(d2, 36 * z**2 - 4) < 1e-11
    x, zz = np.meshgrid(gl(12), gl(12), indexing="ij")
    g = x**3 * zz + 2 * x * zz**2
    assert np.abs(oracle.cheb_deriv(g, 0, 1, 1.0) - (3 * x**2 * zz + 2 * zz**2)).max() < 1e-12
    assert np.abs(oracle.cheb_deriv(g, 1, 1, 1.0) - (x**3 + 4 * x * zz)).max() < 1e-12


def test_cheb_deriv_chebyshev_polynomials(oracle):
    N = 33
    x = gl(N)
    th = np.arccos(np.clip(x, -1, 1))
    for k in range(N):
        Tk = np.cos(k * th)
        if k == 0:
            dT = np.zeros(N)
        else:  # k U_{k-1}(x); at x = +-1, U_{k-1}(+-1) = (+-1)^(k-1) k
            with np.errstate(invalid="ignore", divide="ignore"):
                dT = k * np.sin(k * th) / np.sin(th)
            dT[0] = k * k * (-1.0) ** (k - 1)
            dT[-1] = k * k
        got = oracle.cheb_deriv(Tk, 0, 1, 1.0)
        assert np.abs(got - dT).max() <= 1e-12 * max(1.0, np.abs(dT).max()), k


def collocation_matrix(N):
    """Trefethen's cheb(N-1), re-ordered onto the ascending grid (D_asc = -D_desc)."""
    n = N - 1
    y = np.cos(np.pi * np.arange(N) / n)
    c = np.ones(N)
    c[0] = c[-1] = 2
    c = c * (-1.0) ** np.arange(N)
    X = np.tile(y, (N, 1)).T
    dX = X - X.T
    D = np.outer(c, 1 / c) / (dX + np.eye(N))
    D = D - np.diag(D.sum(axis=1))
    return -D


def test_cheb_deriv_collocation_matrix(oracle):
    for N in (2, 3, 9, 24, 65):
        D = oracle.cheb_deriv(np.eye(N), 0, 1, 1.0)  # column j = derivative of e_j
        ref = collocation_matrix(N)
        scale = np.abs(ref).max()
        assert np.abs(D - ref).max() < 1e-12 * scale, N
        assert abs(D[0, 0] + (2 * (N - 1) ** 2 + 1) / 6) < 1e-11 * scale, N
        D2 = oracle.cheb_deriv(np.eye(N), 0, 2, 0.5)  # (0.5 D)^2
        assert np.abs(D2 - 0.25 * ref @ ref).max() <= 1e-11 * max(1.0, np.abs(ref @ ref).max()), N


def test_cheb_deriv_degenerate_and_complex(oracle):
    rng = np.random.default_rng(3)
    assert np.all(oracle.cheb_deriv(np.array([2.5]), 0, 1, 1.0) == 0)
    a = rng.standard_normal((7, 5))
    assert np.array_equal(oracle.cheb_deriv(a, 1, 0, 3.0), a)
    c = rng.standard_normal((9, 6)) + 1j * rng.standard_normal((9, 6))
    got = oracle.cheb_deriv(c, 0, 2, 1.5)
    assert np.allclose(got.real, oracle.cheb_deriv(c.real, 0, 2, 1.5), rtol=0, atol=0)
    assert np.allclose(got.imag, oracle.cheb_deriv(c.imag, 0, 2, 1.5), rtol=0, atol=0)


def well_conditioned_batch(n, nmodes):
    """test/test_batched_dense_lu.jl:18-25 (indices shifted to 0-based)."""
    A = np.zeros((n, n, nmodes), np.complex128)
    for m in range(1, nmodes + 1):
        for j in range(1, n + 1):
            for i in range(1, n + 1):
                A[i - 1, j - 1, m - 1] = complex(n + m, 0.5) if i == j else complex(0.1 * (i - j), 0.05 * m)
    return A


def reference_rhs(n, nmodes):
    """B = reshape([(i*0.3) + (i*0.7)im for i in 1:n*nmodes], n, nmodes) (:29-30)."""
    i = np.arange(1, n * nmodes + 1)
    return (i * 0.3 + 1j * i * 0.7).reshape(nmodes, n).T


def test_batched_lu_reference_cases(oracle):
    n, nm = 6, 4
    A = well_conditioned_batch(n, nm)
    B = reference_rhs(n, nm)
    LU, ipiv, info = oracle.batched_lu_factor(A)
    assert np.all(info == 0)
    X = oracle.batched_lu_solve(LU, ipiv, B)
    expected = np.stack([np.linalg.solve(A[:, :, m], B[:, m]) for m in range(nm)], axis=1)
    assert relerr(X, expected) < 1e-12
    A[:, :, 2] = 0  # mode 3 (1-based) exactly singular
    _, _, info = oracle.batched_lu_factor(A)
    assert list(np.nonzero(info)[0] + 1) == [3] and info[2] == 1


@pytest.mark.parametrize("n", [1, 5, 37, 96])
def test_batched_lu_matches_lapack(oracle, n):
    rng = np.random.default_rng(n)
    nm = 6
    A = rng.standard_normal((n, n, nm)) + 1j * rng.standard_normal((n, n, nm))
    LU, ipiv, info = oracle.batched_lu_factor(A)
    assert np.all(info == 0)
    B = rng.standard_normal((n, nm)) + 1j * rng.standard_normal((n, nm))
    X = oracle.batched_lu_solve(LU, ipiv, B)
    for m in range(nm):
        lu, piv = sla.lu_factor(A[:, :, m])
        assert np.array_equal(piv + 1, ipiv[:, m]), m
        assert np.abs(lu - LU[:, :, m]).max() < 1e-12 * np.abs(lu).max()
        assert relerr(X[:, m], np.linalg.solve(A[:, :, m], B[:, m])) < 1e-11
