"""Adaptive CFL timestep on the device.

- tarang_cfl_frequency against numpy evaluating the same FP64 expression in the
  same order (bit-exact), FP64 and FP32 data, 1-3 components, the plain maximum,
  NaN propagation;
- the reference's CFL tests (test/test_cfl.jl:4-58) on GPU domains, rtol 1e-6 as there;
- a 2-D Rayleigh-Benard run under the controller (threshold 0, so dt changes and the
  per-mode systems are refactored every step) against the oracle stepped with the same
  dt sequence and the same grid round trip of u between steps, 1e-12 relative (FP64),
  and that sequence against the controller applied to the oracle's own state (1e-10).
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from tests.conftest import relerr

pytestmark = pytest.mark.gpu


def frequency(gpu, arrays, dx, code):
    from tarang_amd import _lib
    ptrs = (ctypes.c_void_p * len(arrays))(*[a.data_ptr() for a in arrays])
    out = ctypes.c_double()
    dxa = (ctypes.c_double * len(arrays))(*dx) if dx is not None else None
    torch.cuda.synchronize()
    _lib.call("tarang_cfl_frequency", gpu.ctx, len(arrays), ptrs, arrays[0].numel(), dxa, code, ctypes.byref(out))
    return out.value


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ncomp,n", [(1, 1), (1, 1000), (2, 65536 + 7), (3, 8 * 1024 * 1024 + 3)])
def test_cfl_frequency_bit_exact(gpu, dtype, ncomp, n):
    from tarang_amd import _lib
    print(f"cfl frequency ncomp={ncomp} n={n} {dtype.__name__}", flush=True)
    r = np.random.default_rng(n + ncomp)
    us = [r.standard_normal(n).astype(dtype) for _ in range(ncomp)]
    dx = [0.25, 0.0123, 1.7][:ncomp]
    f = np.abs(us[0].astype(np.float64)) / dx[0]
    for u, d in zip(us[1:], dx[1:]):
        f = f + np.abs(u.astype(np.float64)) / d
    code = _lib.F64 if dtype == np.float64 else _lib.F32
    ts = [torch.as_tensor(u, device=gpu.torch_device) for u in us]
    assert frequency(gpu, ts, dx, code) == f.max()
    assert frequency(gpu, ts[:1], None, code) == float(us[0].max())  # plain (signed) maximum


def test_cfl_frequency_nan_and_empty(gpu):
    from tarang_amd import _lib
    print("cfl frequency NaN / empty", flush=True)
    u = torch.zeros(5000, dtype=torch.float64, device=gpu.torch_device)
    u[4321] = float("nan")
    assert math.isnan(frequency(gpu, [u], [1.0], _lib.F64))
    e = torch.zeros(0, dtype=torch.float64, device=gpu.torch_device)
    ptrs = (ctypes.c_void_p * 1)(e.data_ptr())
    out = ctypes.c_double(-1.0)
    dx = (ctypes.c_double * 1)(1.0)
    _lib.call("tarang_cfl_frequency", gpu.ctx, 1, ptrs, 0, dx, _lib.F64, ctypes.byref(out))
    assert out.value == 0.0


def test_reference_cfl_cases(gpu):
    """test/test_cfl.jl:4-58: constant velocity on 1-D and 2-D Fourier domains."""
    import tarang_amd as T
    print("reference CFL cases", flush=True)
    for N in (16, 32):
        for L in (1.0, 2.0):
            for safety in (0.2, 0.4):
                for vel in (0.5, 3.0):
                    dom = T.PeriodicDomain(N, L=(L,), arch=T.GPU())
                    u = T.VectorField(dom, "u")
                    u[0]["g"] = np.full(N, vel)
                    cfl = T.CFL(None, initial_dt=1.0, safety=safety, cadence=1)
                    T.add_velocity_(cfl, u)
                    dt = T.compute_timestep(cfl)
                    assert dt == pytest.approx(safety * T.grid_spacing(dom)[0] / abs(vel), rel=1e-6)
    dom = T.PeriodicDomain(16, 16, L=(1.0, 1.0), arch=T.GPU())
    u = T.VectorField(dom, "u")
    u[0]["g"] = np.ones((16, 16))
    u[1]["g"] = np.ones((16, 16))
    cfl = T.CFL(None, initial_dt=1.0, safety=0.5, cadence=1)
    T.add_velocity_(cfl, u)
    dx, dy = T.grid_spacing(dom)
    assert T.compute_timestep(cfl) == pytest.approx(0.5 / (1 / dx + 1 / dy), rel=1e-6)


def cellular_state(Nx, Nz, Lx, Lz, A):
    """u = (-dpsi/dz, dpsi/dx) for psi = A sin(2 pi x/Lx) (z (Lz - z))^2: divergence-free,
    zero on both walls."""
    x = Lx * np.arange(Nx) / Nx
    z = Lz * (1 - np.cos(np.pi * np.arange(Nz) / (Nz - 1))) / 2
    k = 2 * np.pi / Lx
    g = (z * (Lz - z)) ** 2
    dg = 2 * z * (Lz - z) * (Lz - 2 * z)
    ux = -A * np.sin(k * x)[:, None] * dg[None, :]
    uz = A * k * np.cos(k * x)[:, None] * g[None, :]
    return x, z, ux, uz


def test_rbc_adaptive_dt_matches_oracle(gpu, oracle):
    import tarang_amd as T
    from tests.test_gpu_api import rbc_problem
    print("RBC under the CFL controller", flush=True)
    Lx, Lz, Nx, Nz = 4.0, 1.0, 16, 12
    kappa, nu = (1e5) ** -0.5, (1e5) ** -0.5
    solver, b, u, xb, zb = rbc_problem(np.float64, Lx, Lz, Nx, Nz, kappa, nu, dt=0.01)
    x, z, ux, uz = cellular_state(Nx, Nz, Lx, Lz, 20.0)
    rng = np.random.default_rng(7)
    bg = (Lz - z)[None, :] + 1e-2 * rng.standard_normal((Nx, Nz)) * (z * (Lz - z))[None, :]
    b["g"] = bg
    u[0]["g"] = ux
    u[1]["g"] = uz
    cfl = T.CFL(solver, initial_dt=0.01, safety=0.5, threshold=0.0, max_change=1.5, min_change=0.5)
    T.add_velocity_(cfl, u)
    dts = []
    T.run_(solver, stop_iteration=8, cfl=cfl, callbacks=[(1, lambda s: dts.append(s.dt))])
    assert solver.iteration == 8 and len(dts) == 8
    assert len(set(dts)) > 4  # dt moved, so the systems were refactored
    ref = oracle.RBC2D(Nx, Nz, Lx, Lz, kappa, nu, 0.01, "RK222")
    d = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [Nx, Nz], bounds=[(0.0, Lx), (0.0, Lz)])
    ref.set(oracle.RBC_B, oracle.forward_transform(d, bg))
    ref.set(oracle.RBC_UX, oracle.forward_transform(d, ux))
    ref.set(oracle.RBC_UZ, oracle.forward_transform(d, uz))
    dx = T.grid_spacing(b.domain)
    ref_dts, cur = [], 0.01
    for it, dt in enumerate(dts):
        # the controller on the oracle's state (flow_tools_cfl.jl:240-355 with threshold 0)
        gx = oracle.backward_transform(d, ref.get(oracle.RBC_UX))
        gz = oracle.backward_transform(d, ref.get(oracle.RBC_UZ))
        prop = 0.5 / (np.abs(gx) / dx[0] + np.abs(gz) / dx[1]).max()
        if it > 0 and cur != 0.01:
            prop = min(max(prop, cur * 0.5), cur * 1.5)
        cur = prop
        ref_dts.append(prop)
        if it > 0:
            # the controller leaves u in grid layout, so the next step gathers it back
            # through a forward transform (ensure_layout! :g then :c), which drops what the
            # grid cannot hold (the imaginary part of the Nyquist column)
            ref.update(oracle.RBC_UX, oracle.forward_transform(d, gx))
            ref.update(oracle.RBC_UZ, oracle.forward_transform(d, gz))
        ref.set_dt(dt)
        ref.step(1)
    assert relerr(np.array(dts), np.array(ref_dts)) < 1e-10
    assert relerr(b["c"].cpu().numpy(), ref.get(oracle.RBC_B)) < 1e-12
    assert relerr(u[0]["c"].cpu().numpy(), ref.get(oracle.RBC_UX)) < 1e-12
    assert relerr(u[1]["c"].cpu().numpy(), ref.get(oracle.RBC_UZ)) < 1e-12
    solver.close()


def test_rbc_sbdf2_adaptive_dt_matches_oracle(gpu, oracle):
    """SBDF2 under the controller: every committed dt makes the step variable (ratio
    w = dt/dt_prev) and refactors twice; the history must survive the per-step field
    gathers.  Same dt sequence and grid round trip on the oracle, 1e-12 relative."""
    import tarang_amd as T
    from tests.test_gpu_api import rbc_problem
    print("RBC SBDF2 under the CFL controller", flush=True)
    Lx, Lz, Nx, Nz = 4.0, 1.0, 16, 12
    kappa = nu = (1e5) ** -0.5
    solver, b, u, xb, zb = rbc_problem(np.float64, Lx, Lz, Nx, Nz, kappa, nu, dt=0.01, stepper=T.SBDF2())
    x, z, ux, uz = cellular_state(Nx, Nz, Lx, Lz, 20.0)
    bg = (Lz - z)[None, :] + 0.01 * np.cos(2 * np.pi * x / Lx)[:, None] * (z * (Lz - z))[None, :]
    b["g"] = bg
    u[0]["g"] = ux
    u[1]["g"] = uz
    cfl = T.CFL(solver, initial_dt=0.01, safety=0.5, threshold=0.0, max_change=1.5, min_change=0.5)
    T.add_velocity_(cfl, u)
    dts = []
    T.run_(solver, stop_iteration=7, cfl=cfl, callbacks=[(1, lambda s: dts.append(s.dt))])
    assert len(set(dts)) > 3
    ref = oracle.RBC2D(Nx, Nz, Lx, Lz, kappa, nu, 0.01, "SBDF2")
    d = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [Nx, Nz], bounds=[(0.0, Lx), (0.0, Lz)])
    ref.set(oracle.RBC_B, oracle.forward_transform(d, bg))
    ref.set(oracle.RBC_UX, oracle.forward_transform(d, ux))
    ref.set(oracle.RBC_UZ, oracle.forward_transform(d, uz))
    for it, dt in enumerate(dts):
        if it > 0:  # the grid round trip of u that the controller causes (see above)
            gx = oracle.backward_transform(d, ref.get(oracle.RBC_UX))
            gz = oracle.backward_transform(d, ref.get(oracle.RBC_UZ))
            ref.update(oracle.RBC_UX, oracle.forward_transform(d, gx))
            ref.update(oracle.RBC_UZ, oracle.forward_transform(d, gz))
        ref.set_dt(dt)
        ref.step(1)
    assert relerr(b["c"].cpu().numpy(), ref.get(oracle.RBC_B)) < 1e-12
    assert relerr(u[0]["c"].cpu().numpy(), ref.get(oracle.RBC_UX)) < 1e-12
    assert relerr(u[1]["c"].cpu().numpy(), ref.get(oracle.RBC_UZ)) < 1e-12
    solver.close()


def test_rbc3d_adaptive_dt_matches_oracle(gpu, oracle):
    """The 3-D channel (SBDF2) under the controller with a three-component velocity:
    per-step gathers keep the history (tarang_chan3_update_field), every dt change
    refactors the per-key systems; oracle replay with the same dt sequence, 1e-12."""
    import tarang_amd as T
    print("RBC 3-D under the CFL controller", flush=True)
    Lx, Ly, Lz, Nx, Ny, Nz = 4.0, 2.0, 1.0, 16, 8, 10
    kappa = nu = 1e-2
    coords = T.CartesianCoordinates("x", "y", "z")
    dist = T.Distributor(coords, dtype=np.float64, device=T.GPU())
    xb = T.RealFourier(coords["x"], size=Nx, bounds=(0.0, Lx), dealias=3 / 2)
    yb = T.RealFourier(coords["y"], size=Ny, bounds=(0.0, Ly), dealias=3 / 2)
    zb = T.ChebyshevT(coords["z"], size=Nz, bounds=(0.0, Lz), dealias=3 / 2)
    domain = T.Domain(dist, (xb, yb, zb))
    p, b, u = T.ScalarField(domain, "p"), T.ScalarField(domain, "b"), T.VectorField(domain, "u")
    taus = [T.ScalarField(dist, "tau_p", (), np.float64), T.ScalarField(dist, "tau_b1", (xb, yb), np.float64),
            T.ScalarField(dist, "tau_b2", (xb, yb), np.float64),
            T.VectorField(dist, coords, "tau_u1", (xb, yb), np.float64),
            T.VectorField(dist, coords, "tau_u2", (xb, yb), np.float64)]
    ex, ey, ez = T.unit_vector_fields(coords, dist)
    tau_lift = lambda A: T.lift(A, T.derivative_basis(zb, 1), -1)  # noqa: E731
    problem = T.IVP([p, b, u] + taus)
    T.add_parameters_(problem, kappa=kappa, nu=nu, Lz=Lz, ez=ez, τ_lift=tau_lift,
                      grad_u=T.grad(u) + ez * tau_lift(taus[3]), grad_b=T.grad(b) + ez * tau_lift(taus[1]))
    T.add_equation_(problem, "trace(grad_u) + tau_p = 0")
    T.add_equation_(problem, "∂t(b) - kappa*div(grad_b) + τ_lift(tau_b2) = -u⋅∇(b)")
    T.add_equation_(problem, "∂t(u) - nu*div(grad_u) + ∇(p) - b*ez + τ_lift(tau_u2) = -u⋅∇(u)")
    for bc in ("b(z=0) = Lz", "u(z=0) = 0", "b(z=Lz) = 0", "u(z=Lz) = 0", "integ(p) = 0"):
        T.add_bc_(problem, bc)
    solver = T.InitialValueSolver(problem, T.SBDF2(), dt=0.01)
    x = Lx * np.arange(Nx) / Nx
    z = zb.grid()
    g, dg = (z * (Lz - z)) ** 2, 2 * z * (Lz - z) * (Lz - 2 * z)
    k = 2 * np.pi / Lx
    A = 10.0
    ux = -A * np.sin(k * x)[:, None, None] * np.ones(Ny)[None, :, None] * dg[None, None, :]
    uz = A * k * np.cos(k * x)[:, None, None] * np.ones(Ny)[None, :, None] * g[None, None, :]
    bg = np.broadcast_to((Lz - z)[None, None, :], (Nx, Ny, Nz)).copy()
    b["g"] = bg
    u[0]["g"] = ux
    u[2]["g"] = uz
    cfl = T.CFL(solver, initial_dt=0.01, safety=0.5, threshold=0.0, max_change=1.5, min_change=0.5)
    T.add_velocity_(cfl, u)
    dts = []
    T.run_(solver, stop_iteration=5, cfl=cfl, callbacks=[(1, lambda s: dts.append(s.dt))])
    assert len(set(dts)) > 2
    ref = oracle.RBC3D(Nx, Ny, Nz, Lx, Ly, Lz, kappa, nu, 0.01, "SBDF2")
    d = ref.d
    ref.set(oracle.RBC3_B, oracle.forward_transform(d, bg))
    ref.set(oracle.RBC3_UX, oracle.forward_transform(d, ux))
    ref.set(oracle.RBC3_UZ, oracle.forward_transform(d, uz))
    for it, dt in enumerate(dts):
        if it > 0:  # the controller's grid round trip of u
            for v in (oracle.RBC3_UX, oracle.RBC3_UY, oracle.RBC3_UZ):
                ref.update(v, oracle.forward_transform(d, oracle.backward_transform(d, ref.get(v))))
        ref.set_dt(dt)
        ref.step(1)
    assert relerr(b["c"].cpu().numpy(), ref.get(oracle.RBC3_B)) < 1e-12
    refs = [ref.get(v) for v in (oracle.RBC3_UX, oracle.RBC3_UY, oracle.RBC3_UZ)]
    scale = max(np.abs(r).max() for r in refs)  # uy stays ~0 for this y-independent flow
    for i, r in enumerate(refs):
        assert np.abs(u[i]["c"].cpu().numpy() - r).max() < 1e-12 * scale
    solver.close()
