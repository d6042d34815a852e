"""ctypes binding of libtarang_hip.so (declared in include/tarang_hip.h).

The library is loaded from tarang.jl_amd/lib/.  There is no fallback: if the
library is missing or fails to load, every GPU entry point raises.
"""
import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libtarang_hip.so")

OK, ERR_INVALID, ERR_HIP, ERR_SINGULAR, ERR_CAPACITY, ERR_STATE = range(6)
F32, F64 = 0, 1
REALFOURIER, CHEBYSHEVT = 0, 1


class TarangError(RuntimeError):
    """Raised when a library call returns a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"[tarang_hip status {code}] {msg}")
        self.code = code
        self.msg = msg


class DomainDesc(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("kind", ctypes.c_int * 3),
        ("n", ctypes.c_int * 3),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("dealias", ctypes.c_double * 3),
        ("dtype", ctypes.c_int),
    ]


class FivpDesc(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("n", ctypes.c_int * 2),
        ("L", ctypes.c_double * 2),
        ("dtype", ctypes.c_int),
        ("problem", ctypes.c_int),
        ("timestepper", ctypes.c_int),
        ("nu", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class ChanDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("Lx", ctypes.c_double),
        ("Lz", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("dtype", ctypes.c_int),
        ("timestepper", ctypes.c_int),
        ("b_bottom", ctypes.c_double),
        ("b_top", ctypes.c_double),
        ("refine_sweeps", ctypes.c_int),
        ("split_inverse", ctypes.c_int),
    ]


CHAN_P, CHAN_B, CHAN_UX, CHAN_UZ = range(4)


class Chan3Desc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("Lx", ctypes.c_double),
        ("Ly", ctypes.c_double),
        ("Lz", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("dtype", ctypes.c_int),
        ("timestepper", ctypes.c_int),
        ("b_bottom", ctypes.c_double),
        ("b_top", ctypes.c_double),
        ("refine_sweeps", ctypes.c_int),
        ("split_inverse", ctypes.c_int),
    ]


CHAN3_P, CHAN3_B, CHAN3_UX, CHAN3_UY, CHAN3_UZ = range(5)

SERVE_PERIODIC, SERVE_CHANNEL, SERVE_CHANNEL3 = range(3)


class ServeRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("fivp", FivpDesc),
        ("chan", ChanDesc),
        ("chan3", Chan3Desc),
        ("nsteps", ctypes.c_int),
        ("arrival", ctypes.c_double),
        ("ic", ctypes.c_void_p * 5),
        ("out", ctypes.c_void_p * 5),
    ]

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PP = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (every function returns int status)
SIGNATURES = {
    "tarang_version": [_PI, _PI],
    "tarang_ctx_create": [_I, _PP],
    "tarang_ctx_destroy": [_P],
    "tarang_ctx_set_stream": [_P, _P],
    "tarang_ctx_sync": [_P],
    "tarang_buf_alloc": [_P, ctypes.c_size_t, _PP],
    "tarang_buf_free": [_P, _P],
    "tarang_h2d": [_P, _P, _P, ctypes.c_size_t],
    "tarang_d2h": [_P, _P, _P, ctypes.c_size_t],
    "tarang_plan_create": [_P, ctypes.POINTER(DomainDesc), _PP],
    "tarang_plan_destroy": [_P],
    "tarang_plan_coeff_shape": [_P, _PI, _PI],
    "tarang_plan_set_cheb_method": [_P, ctypes.c_int],
    "tarang_forward_transform": [_P, _P, _P],
    "tarang_backward_transform": [_P, _P, _P],
    "tarang_padded_multiply": [_P, _P, _P, _P],
    "tarang_cheb_deriv": [_P, _P, _P, _I, _PI, _I, _I, _D, _I, _I],
    "tarang_batched_lu_factor": [_P, _P, _I, _I, _P, _PI],
    "tarang_cfl_frequency": [_P, _I, _PP, ctypes.c_size_t, _PD, _I, _PD],
    "tarang_cfl_frequency_noalloc": [_P, _I, _PP, ctypes.c_size_t, _PD, _I, _P],
    "tarang_fivp_diagnostics": [_P, _PD, _PD],
    "tarang_chan_nusselt": [_P, _PD],
    "tarang_chan3_nusselt": [_P, _PD],
    "tarang_batched_lu_solve": [_P, _P, _P, _I, _I, _P, _P],
    "tarang_fivp_create": [_P, ctypes.POINTER(FivpDesc), _PP],
    "tarang_fivp_destroy": [_P],
    "tarang_fivp_set_state": [_P, _P],
    "tarang_fivp_update_state": [_P, _P],
    "tarang_fivp_get_state": [_P, _P],
    "tarang_fivp_rhs": [_P, _P, _P],
    "tarang_fivp_aux": [_P, _P, _P, _P, _P],
    "tarang_fivp_step": [_P, _I],
    "tarang_fivp_set_graph": [_P, _I],
    "tarang_fivp_set_dt": [_P, _D],
    "tarang_fivp_time": [_P, _PD, ctypes.POINTER(ctypes.c_long)],
    "tarang_fivp_profile": [_P, _I, _PD, ctypes.POINTER(ctypes.c_long)],
    "tarang_fivp_get_tuning": [_P, _PI],
    "tarang_fivp_set_tuning": [_P, _PI],
    "tarang_fivp_kernel_bytes": [_P, _PD, _PD, _PD],
    "tarang_chan_create": [_P, ctypes.POINTER(ChanDesc), _PP],
    "tarang_chan_destroy": [_P],
    "tarang_chan_set_field": [_P, _I, _P],
    "tarang_chan_update_field": [_P, _I, _P],
    "tarang_chan_set_bc": [_P, _P, _P],
    "tarang_chan_get_field": [_P, _I, _P],
    "tarang_chan_get_tau": [_P, _I, _P],
    "tarang_chan_rhs": [_P, _P, _P, _P],
    "tarang_chan_step": [_P, _I],
    "tarang_chan_set_graph": [_P, _I],
    "tarang_chan_group_create": [_PP, _I, _PP],
    "tarang_chan_group_step": [_P, _I],
    "tarang_chan_group_set_graph": [_P, _I],
    "tarang_chan_group_destroy": [_P],
    "tarang_chan_set_dt": [_P, _D],
    "tarang_chan_time": [_P, _PD, ctypes.POINTER(ctypes.c_long)],
    "tarang_chan_profile": [_P, _I, _PD, ctypes.POINTER(ctypes.c_long)],
    "tarang_chan_kernel_bytes": [_P, _PD],
    "tarang_chan3_create": [_P, ctypes.POINTER(Chan3Desc), _PP],
    "tarang_chan3_destroy": [_P],
    "tarang_chan3_set_field": [_P, _I, _P],
    "tarang_chan3_update_field": [_P, _I, _P],
    "tarang_chan3_set_bc": [_P, _P, _P],
    "tarang_chan3_get_field": [_P, _I, _P],
    "tarang_chan3_get_tau": [_P, _I, _P],
    "tarang_chan3_rhs": [_P, _P, _P, _P, _P],
    "tarang_chan3_step": [_P, _I],
    "tarang_chan3_set_graph": [_P, _I],
    "tarang_chan3_set_dt": [_P, _D],
    "tarang_chan3_time": [_P, _PD, ctypes.POINTER(ctypes.c_long)],
    "tarang_chan3_num_keys": [_P, _PI],
    "tarang_chan3_profile": [_P, _I, _PD, ctypes.POINTER(ctypes.c_long)],
    "tarang_chan3_kernel_bytes": [_P, _PD],
    "tarang_chan3_get_tuning": [_P, _PI],
    "tarang_chan3_set_tuning": [_P, _PI],
    "tarang_serve_create": [_I, _I, ctypes.c_size_t, _I, _PP],
    "tarang_serve_destroy": [_P],
    "tarang_serve_submit": [_P, ctypes.POINTER(ServeRequest), _PI],
    "tarang_serve_run": [_P],
    "tarang_serve_record": [_P, _I, _PD, _PD, ctypes.POINTER(ctypes.c_size_t), _PI, _PI],
    "tarang_serve_record_error": [_P, _I, _PI, ctypes.c_char_p, ctypes.c_size_t],
    "tarang_serve_footprint": [ctypes.POINTER(ServeRequest), ctypes.POINTER(ctypes.c_size_t)],
}

_lib = None


def load():
    """Load the library (raises OSError if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libtarang_hip.so not found at {LIB_PATH}; build it with "
                      f"`python tarang.jl_amd/build.py` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.tarang_last_error.restype = ctypes.c_char_p
    L.tarang_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def call(name, *args):
    L = load()
    rc = getattr(L, name)(*args)
    if rc != OK:
        msg = L.tarang_last_error()
        raise TarangError(rc, msg.decode() if msg else "")
    return rc
