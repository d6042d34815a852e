"""Checkpoint / restart of the device state in the reference's NetCDF slab layout.

Mirrors, for one rank, save_state / load_state! (src/tools/solver_checkpoint.jl:
88-194) and save_field / load_field! (src/tools/field_netcdf_io.jl:217-272, slab
writer src/tools/netcdf_slab_io.jl:119-139).  `<stem>.nc` holds one variable per
state field: its grid data (the field brought to grid layout on the device, then
copied to the host), dimensions `<name>_d<i>` in the reference's axis order
(axis 1 varies fastest), and the slab attributes `start` (0-based), `count` and
`global_shape`.  save_state writes the clock attributes `sim_time`, `iteration`
and `dt` last, after every field.  The file is NetCDF classic with 64-bit offsets
(scipy.io.netcdf_file); integer attributes are 32-bit.

Loading restarts the timestepper: the checkpoint carries the state, not the
multistep history, so a multistep scheme re-seeds it (warned, as the reference
does, solver_checkpoint.jl:32-78).
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from .core import ScalarField, VectorField, ensure_layout_

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

_MULTISTEP_RESEED_STEPS = {"SBDF2": 1, "DiagonalIMEX_SBDF2": 1}
_CLOCK_ATTRS = ("sim_time", "iteration", "dt")


def _netcdf():
    from scipy.io import netcdf_file
    return netcdf_file


_HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"


def _open_read(path: str):
    """NetCDF classic file for reading; a NetCDF-4 (HDF5) file is refused with the way to
    convert it, since no HDF5 reader is available here."""
    with open(path, "rb") as f:
        head = f.read(8)
    if head == _HDF5_SIGNATURE:
        raise ValueError(f"{path} is a NetCDF-4 (HDF5) file; only NetCDF classic files can be read here. "
                         f"Convert it with `nccopy -k classic {path} out.nc`, or write it with "
                         f"mode = NC_CLASSIC_MODEL")
    return _netcdf()(path, "r", mmap=False)


def _target(path: str) -> str:
    stem = path[:-3] if path.endswith(".nc") else path
    if os.path.isdir(stem):
        raise ValueError(f"checkpoint: `{stem}` exists as a directory (a multi-rank slab set); "
                         f"this single-GPU writer refuses to replace it")
    return stem + ".nc"


def _source(path: str) -> str:
    if os.path.isfile(path):
        return path
    if os.path.isfile(path + ".nc"):
        return path + ".nc"
    raise FileNotFoundError(f"checkpoint: no file at `{path}` or `{path}.nc`")


def state_fields(solver):
    """The solver's state: every problem variable with spatial data (vector fields by
    component); variables without bases are skipped, as the reference skips them."""
    out = []
    for v in solver.problem.variables:
        comps = v.components if isinstance(v, VectorField) else [v]
        out += [c for c in comps if isinstance(c, ScalarField) and c.domain is not None]
    return out


def _write_var(nc, name: str, field: ScalarField):
    ensure_layout_(field, "g")
    host = field._g.detach().cpu().numpy()  # storage order: reversed logical shape
    shape = tuple(reversed(host.shape))  # logical (reference) order
    dims = [f"{name}_d{i + 1}" for i in range(len(shape))]
    for d, n in zip(dims, shape):
        nc.createDimension(d, n)
    var = nc.createVariable(name, host.dtype.char if host.dtype in (np.float32, np.float64) else "d",
                            tuple(reversed(dims)))
    var[...] = host
    var.start = np.zeros(len(shape), np.int32)
    var.count = np.array(shape, np.int32)
    var.global_shape = np.array(shape, np.int32)


def _read_var(nc, name: str, field: ScalarField, path: str):
    if name not in nc.variables or not hasattr(nc.variables[name], "global_shape"):
        slab = sorted(k for k, v in nc.variables.items() if hasattr(v, "global_shape"))
        raise KeyError(f"load_field: no variable '{name}' with slab metadata in '{path}'. Variables found: {slab}.")
    var = nc.variables[name]
    want = list(field.domain.grid_shape)
    have = [int(x) for x in np.atleast_1d(var.global_shape)]
    if have != want:
        raise ValueError(f"load_field: '{name}' was written with global shape {have} but this field is {want}.")
    data = np.array(var[...], dtype=field._g.detach().cpu().numpy().dtype)
    field._g.copy_(torch.as_tensor(data, device=field._g.device))
    field.current_layout = "g"


def save_field(field: ScalarField, filename: str, dataset_name: str = "field") -> str:
    """Write one field's grid data as variable `dataset_name` (field_netcdf_io.jl:217-253)."""
    target = _target(filename)
    if os.path.exists(target):
        os.remove(target)
    with _netcdf()(target, "w", version=2) as nc:
        _write_var(nc, dataset_name, field)
    return target


def load_field_(field: ScalarField, filename: str, dataset_name: str = "field") -> ScalarField:
    """Read variable `dataset_name` into the field's grid data (field_netcdf_io.jl:255-272)."""
    path = _source(filename)
    with _open_read(path) as nc:
        _read_var(nc, dataset_name, field, path)
    return field


def save_state(solver, path: str) -> str:
    """Write the solver state and the clock to `<stem>.nc` (solver_checkpoint.jl:88-120)."""
    fields = state_fields(solver)
    if not fields:
        raise ValueError("save_state: solver.state is empty; nothing to write.")
    for f in fields:  # all transforms before any write
        ensure_layout_(f, "g")
    target = _target(path)
    if os.path.exists(target):
        os.remove(target)
    with _netcdf()(target, "w", version=2) as nc:
        for f in fields:
            _write_var(nc, f.name, f)
        nc.sim_time = np.float64(solver.sim_time)  # numpy scalars: a bare float would be stored as FP32
        nc.iteration = np.int32(solver.iteration)
        nc.dt = np.float64(solver.dt)
    return target


def load_state_(solver, path: str):
    """Restore the state fields and the clock, then restart the timestepper from the
    loaded state (solver_checkpoint.jl:122-194)."""
    fields = state_fields(solver)
    if not fields:
        raise ValueError("load_state_: solver.state is empty; nothing to restore.")
    src = _source(path)
    with _open_read(src) as nc:
        for f in fields:
            _read_var(nc, f.name, f, src)
        attrs = nc._attributes
        missing = [k for k in _CLOCK_ATTRS if k not in attrs]
        if missing:
            raise ValueError(f"load_state_: '{path}' is missing the required clock attribute(s) "
                             f"{', '.join(missing)}. The fields loaded, but the simulation clock cannot be "
                             f"restored, so continuing would integrate for the wrong duration with "
                             f"correct-looking field values.")
        sim_time, iteration, dt = float(attrs["sim_time"]), int(attrs["iteration"]), float(attrs["dt"])
    if dt != solver.dt:
        solver.set_dt(dt)
    solver.sim_time, solver.iteration = sim_time, iteration
    solver.restart()
    scheme = getattr(solver.timestepper, "name", "")
    steps = _MULTISTEP_RESEED_STEPS.get(scheme, 0)
    if steps:
        warnings.warn(f"{scheme} restart re-seeds its multistep history: the checkpoint carries the state but "
                      f"not the stored time levels, so the first {steps} step(s) run at the seeding order. The "
                      f"run is correct but not bit-identical to an uninterrupted one.", stacklevel=2)
    return solver
