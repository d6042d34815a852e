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
    Channel families only: `split_inverse` (default: on for an FP32 domain) stores the
    per-mode inverses as FP32 values plus 8-bit remainders, and `refine_sweeps` (0-2)
    adds FP64-residual refinement sweeps to every implicit solve.
    """

    def __init__(self, problem: IVP, timestepper, dt: float, matsolver=None, use_graph: bool = True,
                 split_inverse=None, refine_sweeps: int = 0, **_):
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
        if rec.family in ("rbc2d", "rbc3d"):
            self._channel = _ChannelBackend(self, problem, rec, timestepper, use_graph, split_inverse, refine_sweeps)
            self.dev = self._channel.dev
            self.handle = self._channel.handle
            self.field = self._channel.b
            return
        self._channel = None
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
        self._push_state(restart=True)

    # -- state plumbing --
    def _push_state(self, restart: bool = False):
        """Hand the evolved field to the library.  A new initial condition (restart)
        resets the clock and the multistep history; a state changed between steps
        keeps them, as the reference's gather does (subproblem_io.jl:1143-1152)."""
        ensure_layout_(self.field, "c")
        self.dev.before()
        _lib.call("tarang_fivp_set_state" if restart else "tarang_fivp_update_state", self.handle,
                  _ptr(self.field._c))
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
        """F(X) for the current state, in coefficient layout (evaluate_rhs).  For the
        channel family the result is the tuple (F_b, F_ux, F_uz)."""
        if self._channel is not None:
            return self._channel.rhs()
        out = out or ScalarField(self.field.domain, "F")
        ensure_layout_(self.field, "c")  # a field last written on the grid is transformed first
        self.dev.before()
        _lib.call("tarang_fivp_rhs", self.handle, _ptr(self.field._c), _ptr(out._c))
        self.dev.after()
        out.current_layout = "c"
        return out

    def step_many(self, n: int):
        """Advance n steps in one library call (no per-step host work)."""
        if n <= 0:
            return
        if self._channel is not None:
            self._channel.step(n)
            self.iteration += n
            self.sim_time += n * self.dt
            return
        # the field is handed over before every batch: an edit made in either layout since the
        # last batch (q['g'] = ..., q['c'] = ..., or in place) is seen; the clock and the
        # multistep history are kept (update_state), as the channel backend does
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

    @property
    def state(self):
        """The state fields (problem variables with spatial data, by component)."""
        from .checkpoint import state_fields
        return state_fields(self)

    def restart(self):
        """Start the timestepper afresh from the current fields: the library takes them as
        a new initial condition (clock and multistep history reset), as the reference does
        after load_state! (timestepper_state = nothing)."""
        if self._channel is not None:
            self._channel.push(restart=True)
        else:
            self._push_state(restart=True)

    def set_dt(self, dt: float):
        _lib.call(self._channel.pre + "set_dt" if self._channel is not None else "tarang_fivp_set_dt", self.handle,
                  float(dt))
        self.dt = float(dt)

    def close(self):
        if getattr(self, "handle", None):
            ch = getattr(self, "_channel", None)
            _lib.call(ch.pre + "destroy" if ch is not None else "tarang_fivp_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ChannelBackend:
    """The Rayleigh-Benard families on ChannelDomain (2-D) and RealFourier^2 x ChebyshevT
    (3-D): one tarang_chan_t / tarang_chan3_t holds the state (p, b, u, taus) and the
    factored per-mode systems.  Fields are pushed before and pulled after every batch
    of steps (scatter_inputs / gather_inputs!, subproblem_io.jl:1143-1152)."""

    def __init__(self, solver, problem, rec, timestepper, use_graph, split_inverse=None, refine_sweeps=0):
        names = rec.names
        self.problem = problem
        self.b = problem.var(names["b"])
        self.p = problem.var(names["p"])
        u = problem.var(names["u"])
        dom = self.b.domain
        self.three = rec.family == "rbc3d"
        nd = 3 if self.three else 2
        if not isinstance(u, VectorField) or len(u.components) != nd:
            raise ValueError(f"the channel family needs a {nd}-component VectorField u")
        if dom is None or not _is_gpu(dom.arch):
            raise GPUTransformUnsupported("InitialValueSolver here runs on arch=GPU() domains only")
        kinds = [bb.kind for bb in dom.bases]
        if kinds != [_lib.REALFOURIER] * (nd - 1) + [_lib.CHEBYSHEVT]:
            raise GPUTransformUnsupported("the channel family needs RealFourier horizontal axes and a ChebyshevT z axis")
        for bb in dom.bases:
            if bb.dealias != 1.5:
                raise GPUTransformUnsupported("dealias must be 3/2 for the nonlinear GPU path")
        allowed = (0, 1, 2, 4)
        if timestepper.code not in allowed:
            raise GPUTransformUnsupported(f"{timestepper.name} is not available for this channel family")
        self.taus = [problem.var(names[k]) for k in ("tau_p", "tau_b1", "tau_b2")]
        tu1, tu2 = problem.var(names["tau_u1"]), problem.var(names["tau_u2"])
        self.taus += list(tu1.components) + list(tu2.components)  # slot order of the library
        self.dev = dom.device()
        h = ctypes.c_void_p()
        if self.three:
            d = _lib.Chan3Desc()
            d.nx, d.ny, d.nz = (bb.size for bb in dom.bases)
            d.Lx, d.Ly, d.Lz = (bb.length for bb in dom.bases)
            self.pre = "tarang_chan3_"
            vars_ = (_lib.CHAN3_P, _lib.CHAN3_B, _lib.CHAN3_UX, _lib.CHAN3_UY, _lib.CHAN3_UZ)
        else:
            d = _lib.ChanDesc()
            d.nx, d.nz = dom.bases[0].size, dom.bases[1].size
            d.Lx, d.Lz = dom.bases[0].length, dom.bases[1].length
            self.pre = "tarang_chan_"
            vars_ = (_lib.CHAN_P, _lib.CHAN_B, _lib.CHAN_UX, _lib.CHAN_UZ)
        d.kappa, d.nu = rec.extra["kappa"], rec.extra["nu"]
        d.dt = solver.dt
        d.dtype = _lib.F32 if dom.dtype == np.float32 else _lib.F64
        d.timestepper = timestepper.code
        # the FP32 state's default solve form: inverses with 8-bit remainders (the library
        # refuses split inverses for an FP64 state, so an explicit request there fails loudly)
        split = (dom.dtype == np.float32) if split_inverse is None else bool(split_inverse)
        d.split_inverse = 1 if split else 0
        d.refine_sweeps = int(refine_sweeps)
        walls = (rec.extra["b_bottom"], rec.extra["b_top"])
        # array wall values (space-dependent BC) are handed over as grid arrays before every
        # batch of steps, so a change between steps is seen (the reference re-gathers its
        # ALG_F every stage, step_subproblem_rk_batched.jl:570-580)
        self.walls = walls if any(hasattr(w, "shape") for w in walls) else None
        self.wall_shape = tuple(bb.size for bb in dom.bases[:-1])
        d.b_bottom, d.b_top = (0.0, 0.0) if self.walls else walls
        _lib.call(self.pre + "create", self.dev.handle, ctypes.byref(d), ctypes.byref(h))
        self.handle = h
        _lib.call(self.pre + "set_graph", h, 1 if use_graph else 0)
        self.fields = tuple(zip(vars_, [self.p, self.b] + list(u.components)))
        self.push(restart=True)

    def push(self, restart: bool = False):
        """Fields -> library.  restart (a new initial condition) resets the clock and the
        multistep history; otherwise they are kept (gather_inputs!, subproblem_io.jl:1143)."""
        for _, f in self.fields:
            ensure_layout_(f, "c")
        self.dev.before()
        for var, f in self.fields:
            _lib.call(self.pre + ("set_field" if restart else "update_field"), self.handle, var, _ptr(f._c))
        if self.walls is not None:
            import torch
            dt = self.b._g.dtype
            arr = []
            for w in self.walls:
                a = torch.as_tensor(w) if hasattr(w, "shape") else torch.full(self.wall_shape, float(w))
                if tuple(a.shape) != self.wall_shape:
                    raise ValueError(f"wall array of shape {tuple(a.shape)}: expected {self.wall_shape}")
                # logical (x[, y]) -> column-major storage (x fastest)
                arr.append(a.to(device=self.b._g.device, dtype=dt).permute(*reversed(range(a.dim()))).contiguous())
            _lib.call(self.pre + "set_bc", self.handle, _ptr(arr[0]), _ptr(arr[1]))
            self.dev.after()  # the arrays above are freed on return
        self.dev.after()

    def pull(self):
        self.dev.before()
        for var, f in self.fields:
            _lib.call(self.pre + "get_field", self.handle, var, _ptr(f._c))
            f.current_layout = "c"
        nmodes = self.b._c.numel() // self.b._c.shape[0]  # (nx/2+1)[*ny] values per tau
        for t, f in enumerate(self.taus):
            if f._c is not None and f._c.numel() == nmodes:
                _lib.call(self.pre + "get_tau", self.handle, t, _ptr(f._c))
        self.dev.after()

    def step(self, n):
        # the fields are the state: a host-side change (grid or coefficients) is always seen
        self.push()
        _lib.call(self.pre + "step", self.handle, int(n))
        self.pull()

    def rhs(self):
        self.push()
        names = ("F_b", "F_ux", "F_uy", "F_uz") if self.three else ("F_b", "F_ux", "F_uz")
        outs = [ScalarField(self.b.domain, n) for n in names]
        self.dev.before()
        _lib.call(self.pre + "rhs", self.handle, *[_ptr(o._c) for o in outs])
        self.dev.after()
        for o in outs:
            o.current_layout = "c"
        return tuple(outs)


def step_(solver: InitialValueSolver, dt: Optional[float] = None) -> InitialValueSolver:
    """step!(solver[, dt]) (solver_stepping.jl:24-38)."""
    if dt is not None and float(dt) != solver.dt:
        solver.set_dt(dt)
    solver.step_many(1)
    return solver


def run_(solver: InitialValueSolver, stop_time: Optional[float] = None, stop_iteration: Optional[int] = None,
         stop_wall_time: Optional[float] = None, callbacks: Sequence[Tuple[int, Callable]] = (),
         log_interval: Optional[int] = None, cfl=None) -> InitialValueSolver:
    """run!(solver; stop_time, stop_iteration, callbacks, cfl) (solver_stepping.jl:115-206).

    Steps between callback/log points are issued as one library call each.  With a
    CFL controller every step first takes dt = compute_timestep(cfl)
    (solver_stepping.jl:155-159); a changed dt refactors the implicit systems.
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
    while solver.proceed and cfl is not None:
        from .flow_tools import compute_timestep
        dt = compute_timestep(cfl)
        if dt != solver.dt:
            solver.set_dt(dt)
        # compute_timestep only evaluates the velocities on every cadence-th call and returns
        # the current dt unchanged in between: those steps go to the device as one call
        # (the dt sequence is the one a step-by-step loop produces)
        n = max(1, cfl.cadence - cfl.iteration_count % cfl.cadence) if cfl.cadence > 1 else 1
        if solver.stop_iteration is not None:
            n = min(n, solver.stop_iteration - solver.iteration)
        if math.isfinite(solver.stop_sim_time):
            n = min(n, int(math.ceil((solver.stop_sim_time - solver.sim_time) / solver.dt - 1e-9)))
        for k in intervals:
            n = min(n, k - solver.iteration % k)
        n = max(1, n)
        solver.step_many(n)
        cfl.iteration_count += n - 1
        for k, fn in callbacks:
            if solver.iteration % int(k) == 0:
                fn(solver)
        if log_interval and solver.iteration % int(log_interval) == 0:
            print(f"Iteration {solver.iteration}, t = {solver.sim_time:.6g}, dt = {solver.dt:.6g}")
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
