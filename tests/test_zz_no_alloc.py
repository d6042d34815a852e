"""tarang_cfl_frequency_noalloc: the allocation-free variant of tarang_cfl_frequency.

Its result (a device double written on the context stream) must equal the existing
entry point's bit for bit, on two calls in a row, the second with a smaller input, so
that nothing left in the shared per-GPU scratch array leaks into a later call; also
for FP32 data, the plain maximum and the empty input.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _both(gpu, arrays, dx, code):
    from tarang_amd import _lib
    ptrs = (ctypes.c_void_p * len(arrays))(*[a.data_ptr() for a in arrays])
    dxa = (ctypes.c_double * len(arrays))(*dx) if dx is not None else None
    n = arrays[0].numel()
    torch.cuda.synchronize()
    ref = ctypes.c_double()
    _lib.call("tarang_cfl_frequency", gpu.ctx, len(arrays), ptrs, n, dxa, code, ctypes.byref(ref))
    out = torch.full((1,), -1.0, dtype=torch.float64, device=gpu.torch_device)
    _lib.call("tarang_cfl_frequency_noalloc", gpu.ctx, len(arrays), ptrs, n, dxa, code,
              ctypes.c_void_p(out.data_ptr()))
    _lib.call("tarang_ctx_sync", gpu.ctx)
    return ref.value, float(out.item())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_existing_entry_point(gpu, dtype):
    from tarang_amd import _lib
    code = _lib.F64 if dtype == np.float64 else _lib.F32
    r = np.random.default_rng(5)
    for n in (3 * 1024 * 1024 + 17, 1000):  # second call smaller: fewer partial maxima than the first
        print(f"cfl noalloc n={n} {dtype.__name__}", flush=True)
        us = [torch.as_tensor(r.standard_normal(n).astype(dtype), device=gpu.torch_device) for _ in range(3)]
        ref, got = _both(gpu, us, [0.25, 0.0123, 1.7], code)
        assert got == ref
        ref, got = _both(gpu, us[:1], None, code)  # plain maximum
        assert got == ref


def test_noalloc_empty_input(gpu):
    from tarang_amd import _lib
    print("cfl noalloc empty", flush=True)
    e = torch.zeros(0, dtype=torch.float64, device=gpu.torch_device)
    ref, got = _both(gpu, [e], [1.0], _lib.F64)
    assert got == ref == 0.0
