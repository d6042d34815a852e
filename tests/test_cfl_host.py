"""Host logic of the CFL controller (flow_tools_cfl.jl:240-355) and grid_spacing
(domain.jl:564-594), with the device frequency replaced by known values: safety,
max_dt, the change limits after the first real computation, cadence, the sticky-dt
hysteresis and the constant-diffusivity limit 2 nu sum dx^-2."""
import math

import pytest

import tarang_amd as T


class FixedCFL(T.CFL):
    """Velocity frequencies fed from a list instead of the device reduction."""

    def __init__(self, freqs, **kw):
        super().__init__(None, **kw)
        self.freqs = list(freqs)
        self.velocities = [object()]

    def _velocity_frequency(self, u):
        return self.freqs.pop(0)


def test_grid_spacing():
    c = T.CartesianCoordinates("x", "z")
    d = T.Distributor(c)
    dom = T.Domain(d, (T.RealFourier(c["x"], size=64, bounds=(0.0, 4.0)),
                       T.ChebyshevT(c["z"], size=33, bounds=(0.0, 2.0))))
    dx, dz = T.grid_spacing(dom)
    assert dx == 4.0 / 64
    assert dz == 2.0 * (1 - math.cos(math.pi / 32)) / 2
    one = T.Domain(d, (T.RealFourier(c["x"], size=8, bounds=(0.0, 1.0)),
                       T.ChebyshevT(c["z"], size=2, bounds=(0.0, 3.0))))
    assert T.grid_spacing(one)[1] == 3.0


def test_safety_cap_and_limits():
    cfl = FixedCFL([10.0, 100.0, 1.0, 1.0], initial_dt=0.01, safety=0.5, threshold=0.0, max_change=2.0,
                   min_change=0.5, max_dt=0.2)
    assert T.compute_timestep(cfl) == 0.05              # 0.5 / 10
    assert T.compute_timestep(cfl) == 0.025             # 0.005 clamped to half of 0.05
    assert T.compute_timestep(cfl) == 0.05              # 0.5 clamped to twice 0.025
    assert T.compute_timestep(cfl) == 0.1


def test_max_dt_and_no_frequency():
    cfl = FixedCFL([0.0, 1e-9], initial_dt=0.3, max_dt=0.2, threshold=0.0)
    assert T.compute_timestep(cfl) == 0.2               # no positive frequency: initial_dt, then max_dt
    assert T.compute_timestep(cfl) == 0.2               # 0.4/1e-9 capped at max_dt, inside [0.1, 0.4]


def test_hysteresis_and_cadence():
    cfl = FixedCFL([8.0, 7.5, 4.0], initial_dt=0.05, safety=0.4, threshold=0.1)
    assert T.compute_timestep(cfl) == 0.05              # 0.4/8 = 0.05: no change
    assert T.compute_timestep(cfl) == 0.05              # 0.0533: inside the 10 % band, sticky
    assert T.compute_timestep(cfl) == pytest.approx(0.1)  # 0.1: committed
    cfl = FixedCFL([8.0], initial_dt=0.01, safety=0.4, threshold=0.0, cadence=3)
    assert T.compute_timestep(cfl) == 0.01 and T.compute_timestep(cfl) == 0.01
    assert T.compute_timestep(cfl) == 0.05              # third call computes


def test_constant_diffusivity_limit():
    c = T.CartesianCoordinates("x", "y")
    dom = T.Domain(T.Distributor(c), (T.RealFourier(c["x"], size=32, bounds=(0.0, 1.0)),
                                      T.RealFourier(c["y"], size=16, bounds=(0.0, 2.0))))
    cfl = T.CFL(None, initial_dt=1.0, safety=0.5, threshold=0.0)
    T.add_diffusivity_(cfl, 0.01, domain=dom)
    f = 2 * 0.01 * (32.0 ** 2 + 8.0 ** 2)
    assert T.compute_timestep(cfl) == pytest.approx(0.5 / f, rel=1e-15)
    bad = T.CFL(None)
    T.add_diffusivity_(bad, 0.01)
    with pytest.raises(ValueError, match="cannot determine the grid"):
        T.compute_timestep(bad)


class _FakeSolver:
    """The parts of InitialValueSolver that run_ uses, recording the dt of every step."""

    def __init__(self, dt):
        self.dt, self.iteration, self.sim_time = dt, 0, 0.0
        self.stop_sim_time, self.stop_iteration, self.stop_wall_time = math.inf, None, math.inf
        self.calls, self.step_dts = 0, []

    @property
    def proceed(self):
        if self.sim_time >= self.stop_sim_time - 1e-12 * max(1.0, abs(self.stop_sim_time)):
            return False
        return self.stop_iteration is None or self.iteration < self.stop_iteration

    def set_dt(self, dt):
        self.dt = float(dt)

    def step_many(self, n):
        self.calls += 1
        self.step_dts += [self.dt] * n
        self.iteration += n
        self.sim_time += n * self.dt


class _StateCFL(T.CFL):
    """Velocity frequency read from the solver's iteration, as a device reduction of the
    current state would be."""

    def __init__(self, solver, **kw):
        super().__init__(solver, **kw)
        self.velocities = [object()]

    def _velocity_frequency(self, u):
        return 10.0 + 7.0 * math.sin(0.37 * self.solver.iteration) + 0.05 * self.solver.iteration


@pytest.mark.parametrize("cadence", [1, 3, 10])
def test_run_with_cfl_matches_step_by_step(cadence):
    """run_(cfl=...) hands the steps between two CFL evaluations to the device as one call;
    the dt sequence, the stop and the callback iterations are those of the per-step loop of
    run! (solver_stepping.jl:148-180): compute_timestep, then one step."""
    kw = dict(initial_dt=0.01, cadence=cadence, safety=0.5, threshold=0.1, max_dt=0.05)
    ref = _FakeSolver(0.01)
    ref.stop_sim_time = 3.0
    rc = _StateCFL(ref, **kw)
    ref_hits = []
    while ref.proceed:
        dt = T.compute_timestep(rc)
        if dt != ref.dt:
            ref.set_dt(dt)
        ref.step_many(1)
        if ref.iteration % 7 == 0:
            ref_hits.append((ref.iteration, ref.dt))
    run = _FakeSolver(0.01)
    c = _StateCFL(run, **kw)
    hits = []
    T.run_(run, stop_time=3.0, cfl=c, callbacks=[(7, lambda s: hits.append((s.iteration, s.dt)))])
    assert run.step_dts == ref.step_dts
    assert hits == ref_hits and run.iteration == ref.iteration
    assert c.iteration_count == rc.iteration_count and c.current_dt == rc.current_dt
    assert len(set(ref.step_dts)) > 3          # the controller did move dt
    if cadence > 1:
        assert run.calls < ref.calls
