"""Checkpoint / restart of the device state (save_state / load_state!,
src/tools/solver_checkpoint.jl:88-194; NetCDF slab layout, netcdf_slab_io.jl:119-139).

- the file read back with an independent reader (scipy.io.netcdf_file): one variable
  per state field, dims <name>_d<i> with axis 1 fastest, start/count/global_shape,
  the clock attributes, data equal to the grid values;
- restart equivalence: run k steps, save, run m more (A); a fresh solver loads and
  runs m (B): a one-step scheme gives A == B to the rounding of the grid round trip
  (1e-12), for the 2-D channel (RK222) and the periodic vorticity system (RK443);
- a multistep scheme warns that it re-seeds its history; a file without the clock
  attributes is refused naming them.
"""
import os

import numpy as np
import pytest

import tarang_amd as T
from tests.conftest import relerr
from tests.test_gpu_api import rbc_problem

pytestmark = pytest.mark.gpu


def rbc_ic(b, xb, zb, Lx, Lz):
    z, x = zb.grid(), xb.grid()
    b["g"] = (Lz - z)[None, :] + 0.05 * np.cos(2 * np.pi * x / Lx)[:, None] * (z * (Lz - z))[None, :]


def test_checkpoint_file_layout(tmp_path):
    from scipy.io import netcdf_file
    print("checkpoint file layout", flush=True)
    Lx, Lz, Nx, Nz = 4.0, 1.0, 16, 10
    solver, b, u, xb, zb = rbc_problem(np.float64, Lx, Lz, Nx, Nz, 1e-2, 1e-2, dt=0.01)
    rbc_ic(b, xb, zb, Lx, Lz)
    T.run_(solver, stop_iteration=3)
    path = T.save_state(solver, str(tmp_path / "ck"))
    assert path.endswith("ck.nc") and os.path.isfile(path)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.sim_time == pytest.approx(0.03) and int(nc.iteration) == 3 and nc.dt == 0.01
        names = {f.name for f in solver.state}
        assert names <= set(nc.variables)
        v = nc.variables[b.name]
        assert v.dimensions == (f"{b.name}_d2", f"{b.name}_d1")  # axis 1 (x) fastest
        assert list(v.global_shape) == [Nx, Nz] and list(v.count) == [Nx, Nz] and list(v.start) == [0, 0]
        assert np.array_equal(np.array(v[...]).T, b["g"].cpu().numpy())
    solver.close()


def test_channel_restart_equivalence(tmp_path):
    print("channel restart equivalence", flush=True)
    Lx, Lz, Nx, Nz = 4.0, 1.0, 16, 10
    solver, b, u, xb, zb = rbc_problem(np.float64, Lx, Lz, Nx, Nz, 1e-2, 1e-2, dt=0.01)
    rbc_ic(b, xb, zb, Lx, Lz)
    T.run_(solver, stop_iteration=4)
    T.save_state(solver, str(tmp_path / "ck"))
    T.run_(solver, stop_iteration=7)
    A = [b["c"].cpu().numpy().copy(), u[1]["c"].cpu().numpy().copy()]
    solver.close()
    solver2, b2, u2, _, _ = rbc_problem(np.float64, Lx, Lz, Nx, Nz, 1e-2, 1e-2, dt=0.02)
    T.load_state_(solver2, str(tmp_path / "ck.nc"))
    assert solver2.iteration == 4 and solver2.sim_time == pytest.approx(0.04) and solver2.dt == 0.01
    T.run_(solver2, stop_iteration=7)
    assert relerr(b2["c"].cpu().numpy(), A[0]) < 1e-12
    assert relerr(u2[1]["c"].cpu().numpy(), A[1]) < 1e-12
    solver2.close()


def turb_problem(N, seed=42, nu=1e-3, dt=2e-3):
    from tests.test_gpu_fourier_ivp import turb_ic
    dom = T.PeriodicDomain(N, N, arch=T.GPU())
    q, psi, u = T.ScalarField(dom, "q"), T.ScalarField(dom, "psi"), T.VectorField(dom, "u")
    tau = T.ScalarField(dom.dist, "tau_psi", dtype=np.float64)
    problem = T.IVP([q, psi, u, tau])
    T.add_parameters_(problem, nu=nu)
    T.add_equation_(problem, "∂t(q) - nu*Δ(q) = -u⋅∇(q)")
    T.add_equation_(problem, "Δ(psi) + tau_psi - q = 0")
    T.add_equation_(problem, "u - skew(grad(psi)) = 0")
    T.add_bc_(problem, "integ(psi) = 0")
    q["g"] = turb_ic(N, seed=seed)
    return T.InitialValueSolver(problem, T.RK443(), dt=dt), q


def test_turbulence_restart_equivalence(tmp_path):
    print("turbulence restart equivalence", flush=True)
    solver, q = turb_problem(64)
    T.run_(solver, stop_iteration=5)
    T.save_state(solver, str(tmp_path / "turb"))
    T.run_(solver, stop_iteration=9)
    A = q["c"].cpu().numpy().copy()
    solver.close()
    solver2, q2 = turb_problem(64, seed=99)  # different initial data, overwritten by the load
    T.load_state_(solver2, str(tmp_path / "turb"))
    T.run_(solver2, stop_iteration=9)
    assert relerr(q2["c"].cpu().numpy(), A) < 1e-12
    solver2.close()


def test_multistep_warning_and_missing_clock(tmp_path):
    print("checkpoint warnings and errors", flush=True)
    Lx, Lz, Nx, Nz = 4.0, 1.0, 16, 10
    solver, b, u, xb, zb = rbc_problem(np.float64, Lx, Lz, Nx, Nz, 1e-2, 1e-2, dt=0.01, stepper=T.SBDF2())
    rbc_ic(b, xb, zb, Lx, Lz)
    T.run_(solver, stop_iteration=2)
    T.save_state(solver, str(tmp_path / "ms"))
    with pytest.warns(UserWarning, match="SBDF2 restart re-seeds"):
        T.load_state_(solver, str(tmp_path / "ms"))
    from scipy.io import netcdf_file
    from tarang_amd.checkpoint import _write_var
    with netcdf_file(str(tmp_path / "noclock.nc"), "w", version=2) as nc:  # every field, no clock
        for f in solver.state:
            _write_var(nc, f.name, f)
    with pytest.raises(ValueError, match="missing the required clock"):
        T.load_state_(solver, str(tmp_path / "noclock"))
    # save_field / load_field_ round trip into another field of the same domain
    T.save_field(b, str(tmp_path / "bfield"), "bsnap")
    other = T.ScalarField(b.domain, "other")
    T.load_field_(other, str(tmp_path / "bfield.nc"), "bsnap")
    assert np.array_equal(other["g"].cpu().numpy(), b["g"].cpu().numpy())
    with pytest.raises(KeyError, match="no variable 'nope'"):
        T.load_field_(other, str(tmp_path / "bfield"), "nope")
    solver.close()
