"""ctypes binding of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Arrays follow the
reference's column-major layout; numpy arrays passed here are Fortran-ordered
views of shape (n0, n1, ...) so that axis 0 is contiguous.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

REALFOURIER, CHEBYSHEVT = 0, 1
PROB_DIFFUSION, PROB_TURB2D, PROB_BURGERS1D = 0, 1, 2
TS = {"RK222": 0, "RK443": 1, "SBDF2": 2, "RK111": 3, "SBDF1": 4}


class Domain(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("kind", ctypes.c_int * 3),
        ("n", ctypes.c_int * 3),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("dealias", ctypes.c_double * 3),
    ]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        L.orc_fft.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.orc_rfft.argtypes = [ctypes.c_int, dp, dp]
        L.orc_irfft.argtypes = [ctypes.c_int, dp, dp]
        L.orc_dct1.argtypes = [ctypes.c_int, dp, dp]
        ip = P(ctypes.c_int)
        L.orc_cheb_deriv.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_zgetrf.argtypes = [ctypes.c_int, dp, ip]
        L.orc_zgetrf.restype = ctypes.c_int
        L.orc_zgetrs.argtypes = [ctypes.c_int, dp, ip, dp]
        L.orc_batched_zgetrf.argtypes = [ctypes.c_int, ctypes.c_int, dp, ip, ip]
        L.orc_batched_zgetrs.argtypes = [ctypes.c_int, ctypes.c_int, dp, ip, dp]
        L.orc_coeff_shape.argtypes = [P(Domain), P(ctypes.c_int)]
        L.orc_forward_transform.argtypes = [P(Domain), dp, dp]
        L.orc_backward_transform.argtypes = [P(Domain), dp, dp]
        L.orc_padded_multiply.argtypes = [P(Domain), dp, dp, dp]
        L.orc_padded_size.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_padded_size.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_fivp_create.argtypes = [P(Domain), ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_fivp_create.restype = ctypes.c_void_p
        L.orc_fivp_destroy.argtypes = [ctypes.c_void_p]
        L.orc_fivp_set_coeff.argtypes = [ctypes.c_void_p, dp]
        L.orc_fivp_get_coeff.argtypes = [ctypes.c_void_p, dp]
        L.orc_fivp_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_fivp_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_fivp_rhs.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_fivp_time.argtypes = [ctypes.c_void_p]
        L.orc_fivp_time.restype = ctypes.c_double
        L.orc_fivp_diagnostics.argtypes = [ctypes.c_void_p, dp, dp]
        vp = ctypes.c_void_p
        L.orc_rbc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_rbc_create.restype = vp
        L.orc_rbc_create_rhs_only.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_rbc_create_rhs_only.restype = vp
        L.orc_rbc_destroy.argtypes = [vp]
        L.orc_rbc_size.argtypes = [vp, ctypes.c_int]
        L.orc_rbc_set_bc.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.orc_rbc_set_bc_values.argtypes = [vp, dp, dp]
        L.orc_rbc_update_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc_get_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc_get_tau.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc_step.argtypes = [vp, ctypes.c_int]
        L.orc_rbc_set_dt.argtypes = [vp, ctypes.c_double]
        L.orc_rbc_time.argtypes = [vp]
        L.orc_rbc_time.restype = ctypes.c_double
        L.orc_rbc_rhs_fields.argtypes = [vp, dp, dp, dp]
        L.orc_rbc_matrices.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_rbc3_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_rbc3_create.restype = vp
        L.orc_rbc3_create_rhs_only.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double]
        L.orc_rbc3_create_rhs_only.restype = vp
        L.orc_rbc3_destroy.argtypes = [vp]
        L.orc_rbc3_size.argtypes = [vp, ctypes.c_int]
        L.orc_rbc3_set_bc.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.orc_rbc3_set_bc_values.argtypes = [vp, dp, dp]
        L.orc_rbc3_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc3_update_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc3_get_field.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc3_get_tau.argtypes = [vp, ctypes.c_int, dp]
        L.orc_rbc3_step.argtypes = [vp, ctypes.c_int]
        L.orc_rbc3_set_dt.argtypes = [vp, ctypes.c_double]
        L.orc_rbc3_time.argtypes = [vp]
        L.orc_rbc3_time.restype = ctypes.c_double
        L.orc_rbc3_rhs_fields.argtypes = [vp, dp, dp, dp, dp]
        L.orc_rbc3_matrices.argtypes = [vp, ctypes.c_int, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def domain(kinds, sizes, bounds=None, dealias=1.5):
    d = Domain()
    d.ndim = len(sizes)
    for a, (k, n) in enumerate(zip(kinds, sizes)):
        d.kind[a] = k
        d.n[a] = n
        lo, hi = (bounds[a] if bounds is not None else (0.0, 2 * np.pi))
        d.lo[a], d.hi[a] = lo, hi
        d.dealias[a] = dealias if np.isscalar(dealias) else dealias[a]
    return d


def coeff_shape(d):
    sh = (ctypes.c_int * 3)()
    cplx = lib().orc_coeff_shape(ctypes.byref(d), sh)
    return tuple(sh[a] for a in range(d.ndim)), bool(cplx)


def fft(x, sign=-1):
    a = np.array(x, dtype=np.complex128)
    lib().orc_fft(len(a), _dp(a), sign)
    return a


def rfft(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x) // 2 + 1, np.complex128)
    lib().orc_rfft(len(x), _dp(x), _dp(out))
    return out


def irfft(c, n):
    c = np.ascontiguousarray(c, dtype=np.complex128)
    out = np.zeros(n, np.float64)
    lib().orc_irfft(n, _dp(c), _dp(out))
    return out


def dct1(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_dct1(len(x), _dp(x), _dp(out))
    return out


def cheb_deriv(data, axis, order, scale):
    """Grid-space Chebyshev derivative of `order` along `axis` of a real or complex
    array (any shape); complex data are differentiated as Re and Im separately
    (derivatives_polynomial.jl:423-437)."""
    a = np.asarray(data)
    if np.iscomplexobj(a):
        return (cheb_deriv(a.real, axis, order, scale)
                + 1j * cheb_deriv(a.imag, axis, order, scale))
    x = np.moveaxis(np.asarray(a, np.float64), axis, 0)
    n = x.shape[0]
    cols = np.ascontiguousarray(x.reshape(n, -1).T)          # (batch, n): each column contiguous
    out = np.zeros_like(cols)
    lib().orc_cheb_deriv(n, cols.shape[0], _dp(cols), _dp(out), float(scale), int(order))
    return np.moveaxis(out.T.reshape(x.shape), 0, axis)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def batched_lu_factor(A):
    """LU-factor A[:, :, m] (n, n, nmodes complex) for every m.  Returns
    (LU, ipiv, info): LU in LAPACK getrf layout, ipiv (n, nmodes) 1-based
    interchanges, info[m] = 0 or the first zero pivot (1-based)."""
    A = np.asarray(A, np.complex128)
    n, _, nm = A.shape
    LU = np.array(A, dtype=np.complex128, order="F")
    ipiv = np.zeros((n, nm), np.int32, order="F")
    info = np.zeros(nm, np.int32)
    lib().orc_batched_zgetrf(n, nm, _dp(LU), _ip(ipiv), _ip(info))
    return LU, ipiv, info


def batched_lu_solve(LU, ipiv, B):
    """X[:, m] = A_m^{-1} B[:, m] from batched_lu_factor's output."""
    LU = np.asfortranarray(LU, np.complex128)
    ipiv = np.asfortranarray(ipiv, np.int32)
    n, _, nm = LU.shape
    X = np.array(B, dtype=np.complex128, order="F").reshape(n, nm, order="F")
    lib().orc_batched_zgetrs(n, nm, _dp(LU), _ip(ipiv), _dp(X))
    return X


def _fortran(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def forward_transform(d, grid):
    shape, cplx = coeff_shape(d)
    g = _fortran(grid, np.float64)
    out = np.zeros(shape, np.complex128 if cplx else np.float64, order="F")
    lib().orc_forward_transform(ctypes.byref(d), _dp(g), _dp(out))
    return out


def backward_transform(d, coeff):
    shape, cplx = coeff_shape(d)
    c = _fortran(coeff, np.complex128 if cplx else np.float64)
    grid = np.zeros(tuple(d.n[a] for a in range(d.ndim)), np.float64, order="F")
    lib().orc_backward_transform(ctypes.byref(d), _dp(c), _dp(grid))
    return grid


def padded_multiply(d, ga, gb):
    a = _fortran(ga, np.float64)
    b = _fortran(gb, np.float64)
    out = np.zeros_like(a, order="F")
    lib().orc_padded_multiply(ctypes.byref(d), _dp(a), _dp(b), _dp(out))
    return out


def padded_size(n, factor=1.5):
    return lib().orc_padded_size(int(n), float(factor))


class FourierIVP:
    """Pure-Fourier IVP with a diagonal implicit operator (oracle)."""

    def __init__(self, d, problem, timestepper, nu, dt):
        self.d = d
        self.shape, _ = coeff_shape(d)
        self.h = lib().orc_fivp_create(ctypes.byref(d), problem, TS[timestepper], nu, dt)

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.orc_fivp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_coeff(self, c):
        c = _fortran(c, np.complex128)
        lib().orc_fivp_set_coeff(self.h, _dp(c))

    def get_coeff(self):
        out = np.zeros(self.shape, np.complex128, order="F")
        lib().orc_fivp_get_coeff(self.h, _dp(out))
        return out

    def rhs(self, c):
        c = _fortran(c, np.complex128)
        out = np.zeros(self.shape, np.complex128, order="F")
        lib().orc_fivp_rhs(self.h, _dp(c), _dp(out))
        return out

    def step(self, n=1):
        lib().orc_fivp_step(self.h, int(n))

    def set_dt(self, dt):
        lib().orc_fivp_set_dt(self.h, float(dt))

    @property
    def time(self):
        return lib().orc_fivp_time(self.h)

    def diagnostics(self):
        E = ctypes.c_double()
        Z = ctypes.c_double()
        lib().orc_fivp_diagnostics(self.h, ctypes.byref(E), ctypes.byref(Z))
        return E.value, Z.value


RBC_P, RBC_B, RBC_UX, RBC_UZ = 0, 1, 2, 3


class RBC2D:
    """2-D Rayleigh-Benard channel, Chebyshev-tau per-mode solves (oracle)."""

    def __init__(self, nx, nz, Lx, Lz, kappa, nu, dt, timestepper="RK222", rhs_only=False):
        """rhs_only: no per-mode matrices (set / get / rhs only), for workload-size grids."""
        self.nx, self.nz = nx, nz
        self.shape = (nx // 2 + 1, nz)
        self.h = (lib().orc_rbc_create_rhs_only(nx, nz, Lx, Lz) if rhs_only
                  else lib().orc_rbc_create(nx, nz, Lx, Lz, kappa, nu, dt, TS[timestepper]))
        if not self.h:
            raise ValueError("per-mode system is not square after filtering")
        self.d = domain([REALFOURIER, CHEBYSHEVT], [nx, nz], bounds=[(0.0, Lx), (0.0, Lz)])

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.orc_rbc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def n(self, mode=1):
        return lib().orc_rbc_size(self.h, int(mode))

    def set_bc(self, b_bottom, b_top):
        """Dirichlet values of b at z = 0 and z = Lz."""
        lib().orc_rbc_set_bc(self.h, float(b_bottom), float(b_top))

    def set_bc_values(self, g0, g1):
        """Per-mode wall values of b: complex (nx/2+1) arrays (unnormalised rfft of the walls)."""
        a = np.ascontiguousarray(g0, np.complex128)
        b = np.ascontiguousarray(g1, np.complex128)
        lib().orc_rbc_set_bc_values(self.h, _dp(a), _dp(b))

    def set(self, var, coeff):
        c = _fortran(coeff, np.complex128)
        lib().orc_rbc_set_field(self.h, var, _dp(c))

    def update(self, var, coeff):
        """Replace a field keeping the clock and the multistep history."""
        c = _fortran(coeff, np.complex128)
        lib().orc_rbc_update_field(self.h, var, _dp(c))

    def get(self, var):
        out = np.zeros(self.shape, np.complex128, order="F")
        lib().orc_rbc_get_field(self.h, var, _dp(out))
        return out

    def tau(self, t):
        out = np.zeros(self.shape[0], np.complex128)
        lib().orc_rbc_get_tau(self.h, t, _dp(out))
        return out

    def step(self, n=1):
        rc = lib().orc_rbc_step(self.h, int(n))
        if rc != 0:
            raise RuntimeError("singular per-mode system")

    def set_dt(self, dt):
        lib().orc_rbc_set_dt(self.h, float(dt))

    @property
    def time(self):
        return lib().orc_rbc_time(self.h)

    def rhs(self):
        outs = [np.zeros(self.shape, np.complex128, order="F") for _ in range(3)]
        lib().orc_rbc_rhs_fields(self.h, *[_dp(o) for o in outs])
        return outs

    def matrices(self, mode):
        n = self.n(mode)
        M = np.zeros((n, n), np.complex128, order="F")
        L = np.zeros((n, n), np.complex128, order="F")
        lib().orc_rbc_matrices(self.h, mode, _dp(M), _dp(L))
        return M, L


RBC3_P, RBC3_B, RBC3_UX, RBC3_UY, RBC3_UZ = range(5)


class RBC3D:
    """3-D Rayleigh-Benard channel, Chebyshev-tau per-(kx, ky) solves (oracle)."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, kappa, nu, dt, timestepper="SBDF2", rhs_only=False):
        """rhs_only: no per-mode matrices (set / get / rhs only), for workload-size grids."""
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nx // 2 + 1, ny, nz)
        self.h = (lib().orc_rbc3_create_rhs_only(nx, ny, nz, Lx, Ly, Lz) if rhs_only
                  else lib().orc_rbc3_create(nx, ny, nz, Lx, Ly, Lz, kappa, nu, dt, TS[timestepper]))
        if not self.h:
            raise ValueError("per-mode system is not square after filtering")
        self.d = domain([REALFOURIER, REALFOURIER, CHEBYSHEVT], [nx, ny, nz],
                        bounds=[(0.0, Lx), (0.0, Ly), (0.0, Lz)])

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.orc_rbc3_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def mode(self, m, n):
        return m + (self.nx // 2 + 1) * n

    def n(self, mode):
        return lib().orc_rbc3_size(self.h, int(mode))

    def set_bc(self, b_bottom, b_top):
        lib().orc_rbc3_set_bc(self.h, float(b_bottom), float(b_top))

    def set_bc_values(self, g0, g1):
        """Per-mode wall values of b: complex (nx/2+1, ny) arrays, kx fastest."""
        a = np.asfortranarray(g0, np.complex128)
        b = np.asfortranarray(g1, np.complex128)
        lib().orc_rbc3_set_bc_values(self.h, _dp(a), _dp(b))

    def set(self, var, coeff):
        c = _fortran(coeff, np.complex128)
        lib().orc_rbc3_set_field(self.h, var, _dp(c))

    def update(self, var, coeff):
        """Replace a field keeping the clock and the multistep history."""
        c = _fortran(coeff, np.complex128)
        lib().orc_rbc3_update_field(self.h, var, _dp(c))

    def get(self, var):
        out = np.zeros(self.shape, np.complex128, order="F")
        lib().orc_rbc3_get_field(self.h, var, _dp(out))
        return out

    def tau(self, t):
        out = np.zeros(self.shape[:2], np.complex128, order="F")
        lib().orc_rbc3_get_tau(self.h, t, _dp(out))
        return out

    def step(self, n=1):
        if lib().orc_rbc3_step(self.h, int(n)) != 0:
            raise RuntimeError("singular per-mode system")

    def set_dt(self, dt):
        lib().orc_rbc3_set_dt(self.h, float(dt))

    @property
    def time(self):
        return lib().orc_rbc3_time(self.h)

    def rhs(self):
        outs = [np.zeros(self.shape, np.complex128, order="F") for _ in range(4)]
        lib().orc_rbc3_rhs_fields(self.h, *[_dp(o) for o in outs])
        return outs

    def matrices(self, mode):
        n = self.n(mode)
        M = np.zeros((n, n), np.complex128, order="F")
        L = np.zeros((n, n), np.complex128, order="F")
        lib().orc_rbc3_matrices(self.h, mode, _dp(M), _dp(L))
        return M, L
