"""Batched dense LU over Fourier modes on the device.

Mirrors `BatchedDenseLU` (src/tools/batched_matsolvers.jl:57-123) and its GPU
methods (ext/cuda/batched_matsolvers.jl:80-126): every mode's stage matrix
factored and solved in one call.  Arrays use the reference's logical indexing
backed by column-major storage: `A[:, :, m]` is mode m's n x n matrix and
`B[:, m]`, `X[:, m]` its right-hand side and solution (see `device_array`).

The GPU path factors IN PLACE: after `batched_factor_(s)`, `s.A` holds the LU
factors, not the operator.  The supported lifecycle is assemble-then-factor --
overwrite `A` completely, then factor -- never factor the same buffer twice
(batched_matsolvers.jl:15-35).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .core import GPUTransformUnsupported, _Device, _ptr

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def device_array(shape, dtype=None, device=0):
    """Zero tensor of logical `shape` with column-major storage on cuda:`device`
    (the layout of a Julia array of that shape)."""
    dtype = dtype or torch.complex128
    storage = torch.zeros(tuple(reversed(shape)), dtype=dtype, device=torch.device("cuda", device))
    return storage.permute(*reversed(range(len(shape))))


def _storage(t, what):
    """The contiguous column-major storage behind a logical tensor (no copy)."""
    s = t.permute(*reversed(range(t.dim())))
    if not s.is_contiguous():
        raise ValueError(f"{what} must have column-major storage (see device_array)")
    return s


class BatchedDenseLU:
    """BatchedDenseLU(A): A is a complex128 (n, n, nmodes) device tensor."""

    def __init__(self, A):
        if A.dim() != 3 or A.shape[0] != A.shape[1]:
            raise ValueError("BatchedDenseLU: A must be (n, n, nmodes)")
        if A.dtype != torch.complex128:
            raise TypeError("BatchedDenseLU: A must be ComplexF64 (torch.complex128)")
        self.A = A
        self.pivots = None
        self.info = None
        self.factored = False


def batched_factor_(s: BatchedDenseLU) -> BatchedDenseLU:
    """LU-factor every mode in place; raises naming every singular mode
    (batched_matsolvers.jl:78-101; GPU ext/cuda/batched_matsolvers.jl:80-96)."""
    A = s.A
    if not A.is_cuda:
        raise GPUTransformUnsupported("BatchedDenseLU: the matrices are not on a GPU; "
                                      "this backend has no CPU path")
    n, _, nmodes = A.shape
    store = _storage(A, "A")
    dev = _Device.get(A.device.index or 0)
    piv = device_array((n, nmodes), torch.int32, dev.index)
    info = np.zeros(max(nmodes, 1), np.int32)
    dev.before()
    try:
        _lib.call("tarang_batched_lu_factor", dev.handle, _ptr(store), n, nmodes, _ptr(_storage(piv, "pivots")),
                  info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    except _lib.TarangError as e:
        s.factored = False
        if e.code == _lib.ERR_SINGULAR:
            raise RuntimeError(f"BatchedDenseLU (GPU): {e.msg}. No CPU fallback is attempted.") from None
        raise
    s.pivots = piv
    s.info = info[:nmodes]
    s.factored = True
    return s


def batched_solve_(X, s: BatchedDenseLU, B):
    """Solve every mode against the stored factorization; X may alias B
    (batched_matsolvers.jl:103-123; GPU ext/cuda/batched_matsolvers.jl:110-126)."""
    if not s.factored:
        raise RuntimeError("BatchedDenseLU: batched_solve! called before batched_factor!; "
                           "the factorization is stale or absent")
    n, _, nmodes = s.A.shape
    if tuple(X.shape) != (n, nmodes) or tuple(B.shape) != (n, nmodes):
        raise ValueError(f"BatchedDenseLU: X and B must be ({n}, {nmodes})")
    if X.dtype != torch.complex128 or B.dtype != torch.complex128:
        raise TypeError("BatchedDenseLU: X and B must be ComplexF64")
    dev = _Device.get(s.A.device.index or 0)
    dev.before()
    _lib.call("tarang_batched_lu_solve", dev.handle, _ptr(_storage(s.A, "A")), _ptr(_storage(s.pivots, "pivots")),
              n, nmodes, _ptr(_storage(B, "B")), _ptr(_storage(X, "X")))
    dev.after()
    return X
