"""Adaptive CFL timestep controller.

Mirrors `CFL`, `add_velocity!`, `add_diffusivity!` and `compute_timestep`
(src/extras/flow_tools/flow_tools_cfl.jl:43-355) and `grid_spacing`
(src/core/domain.jl:564-594).  The per-point work -- max over the grid of
sum_i |u_i|/dx_i, or the maximum of a diffusivity field -- runs on the device
(`tarang_cfl_frequency`); the controller logic (safety, max_dt, change limits,
sticky-dt hysteresis) is host arithmetic on a handful of scalars, as in the
reference.  With `run_(solver, cfl=cfl)` the solver takes the returned dt every
step; an implicit solver refactors its per-mode systems whenever dt changes.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional

from . import _lib
from .core import ChebyshevT, GPUTransformUnsupported, ScalarField, VectorField, _ptr, ensure_layout_

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def grid_spacing(domain) -> List[float]:
    """Per-axis spacing: L/N on Fourier axes, the smallest Gauss-Lobatto spacing
    L (1 - cos(pi/(N-1)))/2 on a ChebyshevT axis (domain.jl:564-594)."""
    out = []
    for b in domain.bases:
        L, n = b.length, b.size
        if n < 1:
            out.append(1.0)
        elif isinstance(b, ChebyshevT):
            out.append(L if n <= 2 else L * (1 - math.cos(math.pi / (n - 1))) / 2)
        else:
            out.append(L / n)
    return out


def _device_frequency(arrays, dx, dtype_code, dev):
    ptrs = (ctypes.c_void_p * len(arrays))(*[a.data_ptr() for a in arrays])
    out = ctypes.c_double()
    dxa = (ctypes.c_double * len(arrays))(*dx) if dx is not None else None
    dev.before()
    _lib.call("tarang_cfl_frequency", dev.handle, len(arrays), ptrs, arrays[0].numel(), dxa, dtype_code,
              ctypes.byref(out))
    return out.value


class _Diffusivity:
    def __init__(self, value, domain):
        self.value = value
        self.domain = domain


class CFL:
    """CFL(solver; initial_dt=0.01, cadence=1, safety=0.4, threshold=0.1,
    max_change=2.0, min_change=0.5, max_dt=Inf)."""

    def __init__(self, solver, initial_dt: float = 0.01, cadence: int = 1, safety: float = 0.4,
                 threshold: float = 0.1, max_change: float = 2.0, min_change: float = 0.5, max_dt: float = math.inf):
        self.solver = solver
        self.initial_dt = float(initial_dt)
        self.safety = float(safety)
        self.threshold = float(threshold)
        self.max_change = float(max_change)
        self.min_change = float(min_change)
        self.max_dt = float(max_dt)
        self.cadence = int(cadence)
        self.velocities: List[VectorField] = []
        self.diffusivities: List[_Diffusivity] = []
        self.current_dt = float(initial_dt)
        self.iteration_count = 0

    def __repr__(self):
        return (f"CFL(safety={self.safety}, threshold={self.threshold}, cadence={self.cadence}, "
                f"current_dt={self.current_dt:.6g}, velocities={len(self.velocities)}, "
                f"diffusivities={len(self.diffusivities)})")

    # ---- frequencies ----
    def _velocity_frequency(self, u: VectorField) -> float:
        dom = u.components[0].domain
        dx = grid_spacing(dom)
        comps = []
        for c in u.components:
            ensure_layout_(c, "g")
            comps.append(c._g)
        code = _lib.F32 if comps[0].dtype == torch.float32 else _lib.F64
        return _device_frequency(comps, dx[:len(comps)], code, dom.device())

    def _diffusivity_domain(self, entry: _Diffusivity):
        if entry.domain is not None:
            return entry.domain
        for u in self.velocities:
            if u.components[0].domain is not None:
                return u.components[0].domain
        dom = getattr(self.solver, "field", None)
        if dom is not None and dom.domain is not None:
            return dom.domain
        raise ValueError("add_diffusivity_: cannot determine the grid for this diffusivity -- no velocity or "
                         "problem domain is available; pass domain=... explicitly")

    @staticmethod
    def _max_diffusivity(value) -> float:
        if isinstance(value, float):
            return value
        if isinstance(value, ScalarField):
            ensure_layout_(value, "g")
            t = value._g
            dev = value.domain.device()
        else:
            t = value
            if not (torch is not None and torch.is_tensor(t) and t.is_cuda):
                raise GPUTransformUnsupported("add_diffusivity_: an array diffusivity must be a device tensor")
            from .core import _Device
            dev = _Device.get(t.device.index or 0)
            t = t.contiguous()
        if t.numel() == 0:
            return 0.0
        code = _lib.F32 if t.dtype == torch.float32 else _lib.F64
        return _device_frequency([t], None, code, dev)


def add_velocity_(cfl: CFL, velocity: VectorField) -> CFL:
    cfl.velocities.append(velocity)
    return cfl


def add_diffusivity_(cfl: CFL, nu, domain=None) -> CFL:
    """Register a diffusivity (number, device array or ScalarField) for the explicit
    diffusion limit dt <= 1/(2 nu_max sum_i dx_i^-2) (flow_tools_cfl.jl:134-169)."""
    if isinstance(nu, (int, float)):
        cfl.diffusivities.append(_Diffusivity(float(nu), domain))
    elif isinstance(nu, ScalarField):
        cfl.diffusivities.append(_Diffusivity(nu, domain if domain is not None else nu.domain))
    else:
        cfl.diffusivities.append(_Diffusivity(nu, domain))
    return cfl


def compute_timestep(cfl: CFL) -> float:
    """The next dt (flow_tools_cfl.jl:240-355): safety / (largest stability frequency),
    capped by max_dt, limited to [min_change, max_change] x the current dt after the
    first real computation, and committed only when it moves by more than
    `threshold` (relative) -- every commit makes an implicit solver refactor."""
    cfl.iteration_count += 1
    if cfl.iteration_count % cfl.cadence != 0:
        return cfl.current_dt
    if not cfl.velocities and not cfl.diffusivities:
        return cfl.current_dt
    freqs = [cfl._velocity_frequency(u) for u in cfl.velocities]
    for entry in cfl.diffusivities:
        inv_dx2 = sum(1.0 / (d * d) for d in grid_spacing(cfl._diffusivity_domain(entry)) if d > 0)
        freqs.append(2.0 * max(0.0, cfl._max_diffusivity(entry.value)) * inv_dx2)
    min_dt = math.inf
    for f in freqs:
        if f > 0:
            min_dt = min(min_dt, 1.0 / f)
    proposed = cfl.initial_dt if min_dt == math.inf else cfl.safety * min_dt
    proposed = min(proposed, cfl.max_dt)
    if cfl.iteration_count > 1 and cfl.current_dt > 0 and cfl.current_dt != cfl.initial_dt:
        proposed = min(max(proposed, cfl.current_dt * cfl.min_change), cfl.current_dt * cfl.max_change)
    if cfl.current_dt > 0 and cfl.threshold > 0:
        if abs(proposed - cfl.current_dt) / cfl.current_dt <= cfl.threshold:
            return cfl.current_dt
    cfl.current_dt = proposed
    return proposed
