"""Direct C-ABI access for GPU parity tests (ctypes + torch device buffers).

Arrays cross as column-major data: a logical (n0, n1, ...) array is stored as a
torch tensor of the reversed shape, exactly like the reference's Julia arrays.
"""
import ctypes

import numpy as np
import torch

from tarang_amd import _lib


def to_dev(a, dtype, device):
    """numpy logical array (column-major semantics) -> contiguous torch tensor of reversed shape."""
    t = torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), device=device)
    return t.to(dtype).contiguous()


def from_dev(t):
    return t.detach().cpu().numpy().T


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Device:
    def __init__(self, index=0):
        self.index = index
        self.torch_device = torch.device("cuda", index)
        h = ctypes.c_void_p()
        _lib.call("tarang_ctx_create", index, ctypes.byref(h))
        self.ctx = h

    def sync(self):
        torch.cuda.synchronize(self.torch_device)
        _lib.call("tarang_ctx_sync", self.ctx)

    # ---- transform plans ----
    def plan(self, kinds, sizes, bounds=None, dtype=np.float64, dealias=1.5):
        d = _lib.DomainDesc()
        d.ndim = len(sizes)
        for a, (k, n) in enumerate(zip(kinds, sizes)):
            d.kind[a] = k
            d.n[a] = n
            lo, hi = bounds[a] if bounds else (0.0, 2 * np.pi)
            d.lo[a], d.hi[a] = lo, hi
            d.dealias[a] = dealias
        d.dtype = _lib.F32 if dtype == np.float32 else _lib.F64
        h = ctypes.c_void_p()
        _lib.call("tarang_plan_create", self.ctx, ctypes.byref(d), ctypes.byref(h))
        return Plan(self, h, sizes, dtype)

    def fivp(self, sizes, L, problem, stepper, nu, dt, dtype=np.float64):
        d = _lib.FivpDesc()
        d.ndim = len(sizes)
        d.n[0] = sizes[0]
        d.n[1] = sizes[1] if len(sizes) > 1 else 1
        d.L[0] = L[0]
        d.L[1] = L[1] if len(L) > 1 else 1.0
        d.dtype = _lib.F32 if dtype == np.float32 else _lib.F64
        d.problem = problem
        d.timestepper = stepper
        d.nu = nu
        d.dt = dt
        h = ctypes.c_void_p()
        _lib.call("tarang_fivp_create", self.ctx, ctypes.byref(d), ctypes.byref(h))
        return Fivp(self, h, sizes, dtype)


    def chan(self, nx, nz, Lx, Lz, kappa, nu, dt, stepper=0, dtype=np.float64, b_bottom=None, b_top=0.0, refine=0,
             split=None):
        """split=None: the library's recommended form (split inverses for an FP32 state)."""
        d = _lib.ChanDesc()
        d.refine_sweeps = refine
        d.split_inverse = (1 if dtype == np.float32 else 0) if split is None else split
        d.nx, d.nz, d.Lx, d.Lz = nx, nz, Lx, Lz
        d.kappa, d.nu, d.dt = kappa, nu, dt
        d.dtype = _lib.F32 if dtype == np.float32 else _lib.F64
        d.timestepper = stepper
        d.b_bottom = Lz if b_bottom is None else b_bottom
        d.b_top = b_top
        h = ctypes.c_void_p()
        _lib.call("tarang_chan_create", self.ctx, ctypes.byref(d), ctypes.byref(h))
        return Chan(self, h, nx, nz, dtype)


    def chan3(self, nx, ny, nz, Lx, Ly, Lz, kappa, nu, dt, stepper=2, dtype=np.float64, b_bottom=None, b_top=0.0,
              refine=0, split=None):
        d = _lib.Chan3Desc()
        d.refine_sweeps = refine
        d.split_inverse = (1 if dtype == np.float32 else 0) if split is None else split
        d.nx, d.ny, d.nz, d.Lx, d.Ly, d.Lz = nx, ny, nz, Lx, Ly, Lz
        d.kappa, d.nu, d.dt = kappa, nu, dt
        d.dtype = _lib.F32 if dtype == np.float32 else _lib.F64
        d.timestepper = stepper
        d.b_bottom = Lz if b_bottom is None else b_bottom
        d.b_top = b_top
        h = ctypes.c_void_p()
        _lib.call("tarang_chan3_create", self.ctx, ctypes.byref(d), ctypes.byref(h))
        return Chan3(self, h, nx, ny, nz, dtype)


class Plan:
    def __init__(self, dev, h, sizes, dtype):
        self.dev, self.h, self.sizes, self.dtype = dev, h, tuple(sizes), dtype
        self.rt = torch.float32 if dtype == np.float32 else torch.float64
        self.ct = torch.complex64 if dtype == np.float32 else torch.complex128
        self.cshape = (sizes[0] // 2 + 1,) + tuple(sizes[1:])

    def forward(self, g):
        gd = to_dev(g, self.rt, self.dev.torch_device)
        c = torch.zeros(tuple(reversed(self.cshape)), dtype=self.ct, device=self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_forward_transform", self.h, ptr(gd), ptr(c))
        self.dev.sync()
        return from_dev(c)

    def backward(self, c):
        cd = to_dev(c, self.ct, self.dev.torch_device)
        keep = cd.clone()
        g = torch.zeros(tuple(reversed(self.sizes)), dtype=self.rt, device=self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_backward_transform", self.h, ptr(cd), ptr(g))
        self.dev.sync()
        assert torch.equal(cd, keep), "backward transform modified its input"
        return from_dev(g)

    def multiply(self, a, b):
        ad = to_dev(a, self.rt, self.dev.torch_device)
        bd = to_dev(b, self.rt, self.dev.torch_device)
        o = torch.zeros_like(ad)
        self.dev.sync()
        _lib.call("tarang_padded_multiply", self.h, ptr(ad), ptr(bd), ptr(o))
        self.dev.sync()
        return from_dev(o)

    def close(self):
        _lib.call("tarang_plan_destroy", self.h)


class Fivp:
    def __init__(self, dev, h, sizes, dtype):
        self.dev, self.h, self.sizes = dev, h, tuple(sizes)
        self.ct = torch.complex64 if dtype == np.float32 else torch.complex128
        self.cshape = (sizes[0] // 2 + 1,) + tuple(sizes[1:])

    def set_state(self, c):
        cd = to_dev(c, self.ct, self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_fivp_set_state", self.h, ptr(cd))
        self.dev.sync()

    def get_state(self):
        c = torch.zeros(tuple(reversed(self.cshape)), dtype=self.ct, device=self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_fivp_get_state", self.h, ptr(c))
        self.dev.sync()
        return from_dev(c)

    def rhs(self, c):
        cd = to_dev(c, self.ct, self.dev.torch_device)
        F = torch.zeros_like(cd)
        self.dev.sync()
        _lib.call("tarang_fivp_rhs", self.h, ptr(cd), ptr(F))
        self.dev.sync()
        return from_dev(F)

    def step(self, n):
        _lib.call("tarang_fivp_step", self.h, int(n))
        self.dev.sync()

    def set_graph(self, on):
        _lib.call("tarang_fivp_set_graph", self.h, 1 if on else 0)

    def diagnostics(self):
        E, Z = ctypes.c_double(), ctypes.c_double()
        self.dev.sync()
        _lib.call("tarang_fivp_diagnostics", self.h, ctypes.byref(E), ctypes.byref(Z))
        return E.value, Z.value

    def set_dt(self, dt):
        _lib.call("tarang_fivp_set_dt", self.h, ctypes.c_double(dt))

    def close(self):
        _lib.call("tarang_fivp_destroy", self.h)


class Chan:
    """2-D Rayleigh-Benard channel solver handle (tarang_chan_*)."""

    def __init__(self, dev, h, nx, nz, dtype):
        self.dev, self.h = dev, h
        self.ct = torch.complex64 if dtype == np.float32 else torch.complex128
        self.cshape = (nx // 2 + 1, nz)

    def _empty(self):
        return torch.zeros(tuple(reversed(self.cshape)), dtype=self.ct, device=self.dev.torch_device)

    def set(self, var, c):
        cd = to_dev(c, self.ct, self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_chan_set_field", self.h, int(var), ptr(cd))
        self.dev.sync()

    def get(self, var):
        c = self._empty()
        self.dev.sync()
        _lib.call("tarang_chan_get_field", self.h, int(var), ptr(c))
        self.dev.sync()
        return from_dev(c)

    def nusselt(self):
        nu = ctypes.c_double()
        self.dev.sync()
        _lib.call("tarang_chan_nusselt", self.h, ctypes.byref(nu))
        return nu.value

    def set_bc(self, g0, g1):
        """Wall values of b: grid arrays on z = 0 and z = Lz (logical order, x fastest in storage)."""
        rt = torch.float32 if self.ct == torch.complex64 else torch.float64
        a, b = to_dev(g0, rt, self.dev.torch_device), to_dev(g1, rt, self.dev.torch_device)
        self.dev.sync()
        _lib.call(("tarang_chan3_" if isinstance(self, Chan3) else "tarang_chan_") + "set_bc", self.h, ptr(a), ptr(b))
        self.dev.sync()

    def tau(self, t):
        v = torch.zeros(self.cshape[0], dtype=self.ct, device=self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_chan_get_tau", self.h, int(t), ptr(v))
        self.dev.sync()
        return v.cpu().numpy()

    def rhs(self):
        outs = [self._empty() for _ in range(3)]
        self.dev.sync()
        _lib.call("tarang_chan_rhs", self.h, *[ptr(o) for o in outs])
        self.dev.sync()
        return [from_dev(o) for o in outs]

    def step(self, n):
        _lib.call("tarang_chan_step", self.h, int(n))
        self.dev.sync()

    def set_graph(self, on):
        _lib.call("tarang_chan_set_graph", self.h, 1 if on else 0)

    def set_dt(self, dt):
        _lib.call("tarang_chan_set_dt", self.h, float(dt))

    def profile(self, n):
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_long * 6)()
        _lib.call("tarang_chan_profile", self.h, int(n), ms, cnt)
        return list(ms), list(cnt)

    def close(self):
        _lib.call("tarang_chan_destroy", self.h)


class Chan3(Chan):
    """3-D Rayleigh-Benard channel solver handle (tarang_chan3_*)."""

    def __init__(self, dev, h, nx, ny, nz, dtype):
        self.dev, self.h = dev, h
        self.ct = torch.complex64 if dtype == np.float32 else torch.complex128
        self.cshape = (nx // 2 + 1, ny, nz)

    def set(self, var, c):
        cd = to_dev(c, self.ct, self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_chan3_set_field", self.h, int(var), ptr(cd))
        self.dev.sync()

    def get(self, var):
        c = self._empty()
        self.dev.sync()
        _lib.call("tarang_chan3_get_field", self.h, int(var), ptr(c))
        self.dev.sync()
        return from_dev(c)

    def nusselt(self):
        nu = ctypes.c_double()
        self.dev.sync()
        _lib.call("tarang_chan3_nusselt", self.h, ctypes.byref(nu))
        return nu.value

    def tau(self, t):
        v = torch.zeros(tuple(reversed(self.cshape[:2])), dtype=self.ct, device=self.dev.torch_device)
        self.dev.sync()
        _lib.call("tarang_chan3_get_tau", self.h, int(t), ptr(v))
        self.dev.sync()
        return from_dev(v)

    def rhs(self):
        outs = [self._empty() for _ in range(4)]
        self.dev.sync()
        _lib.call("tarang_chan3_rhs", self.h, *[ptr(o) for o in outs])
        self.dev.sync()
        return [from_dev(o) for o in outs]

    def step(self, n):
        _lib.call("tarang_chan3_step", self.h, int(n))
        self.dev.sync()

    def set_graph(self, on):
        _lib.call("tarang_chan3_set_graph", self.h, 1 if on else 0)

    def set_dt(self, dt):
        _lib.call("tarang_chan3_set_dt", self.h, float(dt))

    def nkeys(self):
        k = ctypes.c_int()
        _lib.call("tarang_chan3_num_keys", self.h, ctypes.byref(k))
        return k.value

    def profile(self, n):
        ms = (ctypes.c_double * 8)()
        cnt = (ctypes.c_long * 8)()
        _lib.call("tarang_chan3_profile", self.h, int(n), ms, cnt)
        return list(ms), list(cnt)

    def close(self):
        _lib.call("tarang_chan3_destroy", self.h)
