"""Timesteppers, InitialValueSolver, step_/run_ (the reference's solver API).

  timesteppers   src/core/timesteppers/types.jl:13-227, 533-634
  solver         src/core/solvers/solver_types.jl:220-553, 993
  step!/run!     src/core/solvers/solver_stepping.jl:24-38, 54-78, 115-206
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import Callable, Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .core import (GPUTransformUnsupported, ScalarField, VectorField, _is_gpu, _ptr, ensure_layout_)
from .problems import IVP, recognize


class TimeStepper:
    code = None
    name = ""

    def __repr__(self):
        return f"{self.name}()"


class RK111(TimeStepper):
    code, name = 3, "RK111"


class RK222(TimeStepper):
    code, name = 0, "RK222"


class RK443(TimeStepper):
    code, name = 1, "RK443"


class SBDF1(TimeStepper):
    code, name = 4, "SBDF1"


class SBDF2(TimeStepper):
    code, name = 2, "SBDF2"


# same tableaux as the global-matrix schemes (types.jl:533-634)
class DiagonalIMEX_RK222(RK222):
    name = "DiagonalIMEX_RK222"


class DiagonalIMEX_RK443(RK443):
    name = "DiagonalIMEX_RK443"


class DiagonalIMEX_SBDF2(SBDF2):
    name = "DiagonalIMEX_SBDF2"


_PROBLEM_CODE = {"diffusion": 0, "turb2d": 1, "burgers1d": 2}


class InitialValueSolver:
    """IVP solver.  On a GPU domain the whole timestep runs in libtarang_hip.so.

    Keyword arguments mirror the reference (`dt`, `matsolver`, `batched_modes`,
    ...); `use_graph` (default True) replays each step as one captured hipGraph.
    """

    def __init__(self, problem: IVP, timestepper, dt: float, matsolver=None, use_graph: bool = True, **_):
        if isinstance(timestepper, type):
            timestepper = timestepper()
        self.problem = problem
        self.timestepper = timestepper
        self.dt = float(dt)
        self.sim_time = 0.0
        self.iteration = 0
        self.stop_sim_time = math.inf
        self.stop_iteration = None
        self.stop_wall_time = math.inf
        self.start_wall_time = time.time()
        rec = recognize(problem)
        self.recognized = rec
        self.field = problem.var(rec.evolved)
        dom = self.field.domain
        if dom is None or not _is_gpu(dom.arch):
            raise GPUTransformUnsupported("InitialValueSolver here runs on arch=GPU() domains only")
        if rec.family in ("turb2d",) and dom.ndim != 2:
            raise ValueError("the vorticity system needs a 2-D PeriodicDomain")
        if rec.family == "burgers1d" and dom.ndim != 1:
            raise ValueError("Burgers needs a 1-D PeriodicDomain")
        if not all(b.kind == _lib.REALFOURIER for b in dom.bases):
            raise GPUTransformUnsupported("this equation family needs a RealFourier domain")
        if dom.ndim > 2:
            raise GPUTransformUnsupported("3-D periodic IVPs are not supported on the GPU path")
        for b in dom.bases:
            if b.dealias not in (1.5,) and rec.family != "diffusion":
                raise GPUTransformUnsupported("dealias must be 3/2 for the nonlinear GPU path")
        self.dev = dom.device()
        d = _lib.FivpDesc()
        d.ndim = dom.ndim
        for a, b in enumerate(dom.bases):
            d.n[a] = b.size
            d.L[a] = b.length
        if dom.ndim == 1:
            d.n[1], d.L[1] = 1, 1.0
        d.dtype = _lib.F32 if dom.dtype == np.float32 else _lib.F64
        d.problem = _PROBLEM_CODE[rec.family]
        d.timestepper = timestepper.code
        d.nu = rec.coeff
        d.dt = self.dt
        h = ctypes.c_void_p()
        _lib.call("tarang_fivp_create", self.dev.handle, ctypes.byref(d), ctypes.byref(h))
        self.handle = h
        _lib.call("tarang_fivp_set_graph", h, 1 if use_graph else 0)
        self._push_state()

    # -- state plumbing --
    def _push_state(self):
        ensure_layout_(self.field, "c")
        self.dev.before()
        _lib.call("tarang_fivp_set_state", self.handle, _ptr(self.field._c))
        self.dev.after()

    def _pull_state(self):
        _lib.call("tarang_fivp_get_state", self.handle, _ptr(self.field._c))
        self.field.current_layout = "c"
        if self.recognized.family == "turb2d":
            names = self.recognized.names
            psi = self.problem.var(names["psi"])
            u = self.problem.var(names["u"])
            comps = u.components if isinstance(u, VectorField) else (None, None)
            _lib.call("tarang_fivp_aux", self.handle, _ptr(self.field._c), _ptr(psi._c),
                      _ptr(comps[0]._c), _ptr(comps[1]._c))
            psi.current_layout = "c"
            comps[0].current_layout = "c"
            comps[1].current_layout = "c"
        self.dev.after()

    def evaluate_rhs(self, out: Optional[ScalarField] = None) -> ScalarField:
        """F(X) for the current state, in coefficient layout (evaluate_rhs)."""
        out = out or ScalarField(self.field.domain, "F")
        self.dev.before()
        _lib.call("tarang_fivp_rhs", self.handle, _ptr(self.field._c), _ptr(out._c))
        self.dev.after()
        out.current_layout = "c"
        return out

    def step_many(self, n: int):
        """Advance n steps in one library call (no per-step host work)."""
        if n <= 0:
            return
        if self.field.current_layout != "c" or getattr(self, "_dirty", False):
            self._push_state()
        _lib.call("tarang_fivp_step", self.handle, int(n))
        self.iteration += n
        self.sim_time += n * self.dt
        self._pull_state()

    @property
    def proceed(self) -> bool:
        if self.sim_time >= self.stop_sim_time - 1e-12 * max(1.0, abs(self.stop_sim_time)):
            return False
        if self.stop_iteration is not None and self.iteration >= self.stop_iteration:
            return False
        if time.time() - self.start_wall_time >= self.stop_wall_time:
            return False
        return True

    def set_dt(self, dt: float):
        _lib.call("tarang_fivp_set_dt", self.handle, float(dt))
        self.dt = float(dt)

    def close(self):
        if getattr(self, "handle", None):
            _lib.call("tarang_fivp_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_(solver: InitialValueSolver, dt: Optional[float] = None) -> InitialValueSolver:
    """step!(solver[, dt]) (solver_stepping.jl:24-38)."""
    if dt is not None and float(dt) != solver.dt:
        solver.set_dt(dt)
    solver.step_many(1)
    return solver


def run_(solver: InitialValueSolver, stop_time: Optional[float] = None, stop_iteration: Optional[int] = None,
         stop_wall_time: Optional[float] = None, callbacks: Sequence[Tuple[int, Callable]] = (),
         log_interval: Optional[int] = None) -> InitialValueSolver:
    """run!(solver; stop_time, stop_iteration, callbacks) (solver_stepping.jl:115-206).

    Steps between callback/log points are issued as one library call each.
    """
    if stop_time is not None:
        solver.stop_sim_time = float(stop_time)
    if stop_iteration is not None:
        solver.stop_iteration = int(stop_iteration)
    if stop_wall_time is not None:
        solver.stop_wall_time = float(stop_wall_time)
    intervals = [int(k) for k, _ in callbacks]
    if log_interval:
        intervals.append(int(log_interval))
    t0 = time.time()
    it0 = solver.iteration
    while solver.proceed:
        n_left = None
        if solver.stop_iteration is not None:
            n_left = solver.stop_iteration - solver.iteration
        if math.isfinite(solver.stop_sim_time):
            by_time = int(math.ceil((solver.stop_sim_time - solver.sim_time) / solver.dt - 1e-9))
            n_left = by_time if n_left is None else min(n_left, by_time)
        if n_left is None:
            n_left = 1
        for k in intervals:
            n_left = min(n_left, k - solver.iteration % k)
        n_left = max(1, n_left)
        solver.step_many(n_left)
        for k, fn in callbacks:
            if solver.iteration % int(k) == 0:
                fn(solver)
        if log_interval and solver.iteration % int(log_interval) == 0:
            rate = (solver.iteration - it0) / max(time.time() - t0, 1e-12)
            print(f"Iteration {solver.iteration}, t = {solver.sim_time:.6g}, rate {rate:.1f} steps/s")
    return solver
