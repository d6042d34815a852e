"""Host-side behaviour of the derivative and batched-LU plug-ins (no GPU):
argument guards and state errors that the reference raises before any device
work (batched_matsolvers.jl:109-113, derivatives_polynomial.jl:113-140), and
the C entry points refusing NULL handles with a message."""
import ctypes

import numpy as np
import pytest
import torch


def test_batched_lu_solve_before_factor_raises():
    from tarang_amd.matsolvers import BatchedDenseLU, batched_solve_
    A = torch.zeros(6, 6, 4, dtype=torch.complex128)
    s = BatchedDenseLU(A)
    X = torch.zeros(6, 4, dtype=torch.complex128)
    B = torch.ones(6, 4, dtype=torch.complex128)
    with pytest.raises(RuntimeError, match="called before batched_factor"):
        batched_solve_(X, s, B)


def test_batched_lu_argument_guards():
    from tarang_amd.core import GPUTransformUnsupported
    from tarang_amd.matsolvers import BatchedDenseLU, batched_factor_
    with pytest.raises(ValueError):
        BatchedDenseLU(torch.zeros(6, 5, 4, dtype=torch.complex128))
    with pytest.raises(TypeError):
        BatchedDenseLU(torch.zeros(6, 6, 4, dtype=torch.complex64))
    s = BatchedDenseLU(torch.zeros(6, 6, 4, dtype=torch.complex128))
    with pytest.raises(GPUTransformUnsupported):  # never a CPU fallback
        batched_factor_(s)


def test_cheb_deriv_argument_guards():
    from tarang_amd.core import GPUTransformUnsupported
    from tarang_amd.derivatives import gpu_chebyshev_deriv_
    x = torch.zeros(8, 3, dtype=torch.float64)
    with pytest.raises(ValueError):
        gpu_chebyshev_deriv_(x, x, 0, -1, 1.0)
    with pytest.raises(GPUTransformUnsupported):
        gpu_chebyshev_deriv_(x, x, 0, 1, 1.0)


def test_c_entry_points_refuse_null_handles():
    from tarang_amd import _lib
    L = _lib.load()
    shape = (ctypes.c_int * 1)(8)
    assert L.tarang_cheb_deriv(None, None, None, 1, shape, 0, 1, 1.0, _lib.F64, 0) == _lib.ERR_INVALID
    assert b"NULL" in L.tarang_last_error()
    assert L.tarang_batched_lu_factor(None, None, 4, 2, None, None) == _lib.ERR_INVALID
    assert L.tarang_batched_lu_solve(None, None, None, 4, 2, None, None) == _lib.ERR_INVALID
    info = np.ones(3, np.int32)
    assert L.tarang_batched_lu_factor(None, None, 0, 3, None, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) \
        == _lib.ERR_INVALID
