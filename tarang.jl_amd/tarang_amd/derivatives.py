"""Chebyshev grid-space derivative on the device.

Mirrors the GPU plug-in `_gpu_chebyshev_deriv!(result, operand, data_g, axis,
order, scale)` (ext/cuda/cheb_deriv.jl:425-465) and the guards of its caller
`evaluate_chebyshev_derivative!` (src/core/operators/derivatives/
derivatives_polynomial.jl:40-140): the derivative of the given order along a
ChebyshevT axis, on the grid, scaled by 2/(b-a) per order, written straight into
the result's grid buffer.  Axes are 0-based here (1-based in Julia).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .core import ChebyshevT, GPUTransformUnsupported, ScalarField, _Device, _ptr, ensure_layout_

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

_DT = {torch.float32: (_lib.F32, 0), torch.float64: (_lib.F64, 0),
       torch.complex64: (_lib.F32, 1), torch.complex128: (_lib.F64, 1)} if torch is not None else {}


def gpu_chebyshev_deriv_(result, data_g, axis: int, order: int, scale: float) -> bool:
    """result = (scale d/dx)^order data_g along `axis` of two device tensors of the
    same logical shape and dtype (real or complex, FP32 or FP64), both with
    column-major storage.  Returns True (handled), as the reference's method does."""
    if order < 0:
        raise ValueError("derivative order must be non-negative")
    if not data_g.is_cuda or not result.is_cuda:
        raise GPUTransformUnsupported("Chebyshev derivative: data are not on a GPU; this backend has no CPU path")
    if data_g.dtype not in _DT or result.dtype != data_g.dtype or tuple(result.shape) != tuple(data_g.shape):
        raise ValueError("Chebyshev derivative: result and data must share shape and a float/complex dtype")
    nd = data_g.dim()
    if not 0 <= axis < nd:
        raise ValueError(f"axis {axis} out of range for a {nd}-d array")
    src = data_g.permute(*reversed(range(nd)))
    dst = result.permute(*reversed(range(nd)))
    if not src.is_contiguous() or not dst.is_contiguous():
        raise ValueError("Chebyshev derivative: arrays must have column-major storage")
    dt, cplx = _DT[data_g.dtype]
    shape = (_lib.ctypes.c_int * max(nd, 1))(*[int(n) for n in data_g.shape])
    dev = _Device.get(data_g.device.index or 0)
    dev.before()
    _lib.call("tarang_cheb_deriv", dev.handle, _ptr(src), _ptr(dst), nd, shape, int(axis), int(order),
              float(scale), dt, cplx)
    dev.after()
    return True


def evaluate_chebyshev_derivative_(result: ScalarField, operand: ScalarField, axis: int, order: int,
                                   layout: str = "g") -> ScalarField:
    """Derivative of `operand` along its ChebyshevT `axis` into `result`'s grid
    (evaluate_chebyshev_derivative!, derivatives_polynomial.jl:40-140)."""
    if order < 0:
        raise ValueError("derivative order must be non-negative")
    basis = operand.bases[axis]
    if not isinstance(basis, ChebyshevT):
        raise ValueError(f"axis {axis} is not a ChebyshevT axis")
    a, b = basis.bounds
    if not b > a:
        raise ValueError(f"ChebyshevT bounds must satisfy a < b, got ({a}, {b})")
    ensure_layout_(operand, "g")
    gpu_chebyshev_deriv_(result.grid_data, operand.grid_data, axis, order, 2.0 / (b - a))
    result.current_layout = "g"
    if layout == "c":
        ensure_layout_(result, "c")
    return result
