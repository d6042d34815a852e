"""Trajectory diagnostics reduced on the device (tarang_fivp_diagnostics, tarang_chan_nusselt,
tarang_chan3_nusselt).

- On one state: the device reductions equal the host (numpy) formulas of
  tarang_amd.configs / tarang_amd.diagnostics to 1e-12 relative (FP64 accumulation, both
  dtypes of the state).
- Over a fixed horizon against the FP64 oracle stepped from the same initial condition,
  the diagnostic of the oracle's state computed on the host:
    2-D turbulence 128^2, RK443, 200 steps: E and Z within 1e-10 (FP64) / 1e-5 (FP32);
    2-D Rayleigh-Benard 64 x 32, RK222, 500 steps: Nu within 1e-10 (FP64) / 1e-6 (FP32);
    3-D Rayleigh-Benard 16 x 16 x 17, SBDF2, 200 steps: Nu within 1e-10 (FP64) / 1e-6 (FP32).
"""
import numpy as np
import pytest

from tarang_amd import configs, diagnostics

pytestmark = pytest.mark.gpu

TURB, RK443, RK222, SBDF2 = 1, 1, 0, 2
BOUND = {np.float64: {"EZ": 1e-10, "Nu": 1e-10}, np.float32: {"EZ": 1e-5, "Nu": 1e-6}}


def zgrid(N, Lz):
    return Lz * (1 - np.cos(np.pi * np.arange(N) / (N - 1))) / 2


def rbc_ic(O, nx, N, Lz=1.0, ny=None, seed=3, amp=0.05):
    rng = np.random.default_rng(seed)
    z = zgrid(N, Lz)
    if ny is None:
        d = O.domain([O.REALFOURIER, O.CHEBYSHEVT], [nx, N], bounds=[(0.0, 4.0), (0.0, Lz)])
        shape, zz = (nx, N), (None, slice(None))
    else:
        d = O.domain([O.REALFOURIER, O.REALFOURIER, O.CHEBYSHEVT], [nx, ny, N],
                     bounds=[(0.0, 4.0), (0.0, 4.0), (0.0, Lz)])
        shape, zz = (nx, ny, N), (None, None, slice(None))
    wall = (z * (Lz - z))[zz]
    b = (Lz - z)[zz] + amp * rng.standard_normal(shape) * wall
    u = [amp * rng.standard_normal(shape) * wall for _ in range(2 if ny is None else 3)]
    return d, np.asarray(O.forward_transform(d, b)), [np.asarray(O.forward_transform(d, v)) for v in u]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_diagnostics_equal_host_formulas(gpu, oracle, dtype):
    print(f"device diagnostics vs host formulas {dtype.__name__}", flush=True)
    N = 256
    s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, RK443, 1e-4, 1e-3, dtype=dtype)
    C = np.fft.fft(np.fft.rfft(configs.turbulence_2d(N, seed=5, kpeak=10.0), axis=0), axis=1)
    s.set_state(C)
    E, Z = s.diagnostics()
    Eh, Zh = configs.energy_enstrophy_from_coeff(s.get_state().astype(np.complex128), N)
    assert abs(E - Eh) <= 1e-12 * Eh and abs(Z - Zh) <= 1e-12 * Zh, (E, Eh, Z, Zh)
    s.close()
    kappa = nu = 1e5 ** -0.5
    d, B, (U, W) = rbc_ic(oracle, 64, 32)
    c = gpu.chan(64, 32, 4.0, 1.0, kappa, nu, 5e-3, stepper=RK222, dtype=dtype)
    c.set(oracle.RBC_B, B)
    c.set(oracle.RBC_UX, U)
    c.set(oracle.RBC_UZ, W)
    Nu = c.nusselt()
    Nuh = diagnostics.nusselt(c.get(oracle.RBC_B).astype(complex), c.get(oracle.RBC_UZ).astype(complex), 64, 1.0,
                              kappa, 1.0)
    assert abs(Nu - Nuh) <= 1e-12 * abs(Nuh), (Nu, Nuh)
    c.close()
    d, B, (U, V, W) = rbc_ic(oracle, 16, 17, ny=16)
    c3 = gpu.chan3(16, 16, 17, 4.0, 4.0, 1.0, kappa, nu, 5e-3, dtype=dtype)
    for var, v in ((oracle.RBC3_B, B), (oracle.RBC3_UX, U), (oracle.RBC3_UY, V), (oracle.RBC3_UZ, W)):
        c3.set(var, v)
    Nu = c3.nusselt()
    Nuh = diagnostics.nusselt3(c3.get(oracle.RBC3_B).astype(complex), c3.get(oracle.RBC3_UZ).astype(complex),
                               16, 16, 1.0, kappa, 1.0)
    assert abs(Nu - Nuh) <= 1e-12 * abs(Nuh), (Nu, Nuh)
    c3.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_turbulence_energy_enstrophy_over_horizon(gpu, oracle, dtype):
    N, nu, dt, steps = 128, 1e-3, 1e-3, 200
    print(f"turbulence E, Z over {steps} steps {dtype.__name__}", flush=True)
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, configs.turbulence_2d(N, seed=8, kpeak=8.0))
    ref = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", nu, dt)
    ref.set_coeff(C)
    ref.step(steps)
    Er, Zr = configs.energy_enstrophy_from_coeff(ref.get_coeff(), N)
    s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, RK443, nu, dt, dtype=dtype)
    s.set_state(C)
    s.step(steps)
    E, Z = s.diagnostics()
    eE, eZ = abs(E - Er) / Er, abs(Z - Zr) / Zr
    print(f"  E rel err {eE:.3e}, Z rel err {eZ:.3e}", flush=True)
    assert eE < BOUND[dtype]["EZ"] and eZ < BOUND[dtype]["EZ"], (eE, eZ)
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rbc2d_nusselt_over_horizon(gpu, oracle, dtype):
    nx, N, steps, dt = 64, 32, 500, 5e-3
    print(f"2-D Rayleigh-Benard Nu over {steps} steps {dtype.__name__}", flush=True)
    kappa = nu = 1e5 ** -0.5
    d, B, (U, W) = rbc_ic(oracle, nx, N)
    ref = oracle.RBC2D(nx, N, 4.0, 1.0, kappa, nu, dt, "RK222")
    c = gpu.chan(nx, N, 4.0, 1.0, kappa, nu, dt, stepper=RK222, dtype=dtype)
    for var, v in ((oracle.RBC_B, B), (oracle.RBC_UX, U), (oracle.RBC_UZ, W)):
        ref.set(var, v)
        c.set(var, v)
    ref.step(steps)
    c.step(steps)
    Nur = diagnostics.nusselt(ref.get(oracle.RBC_B), ref.get(oracle.RBC_UZ), nx, 1.0, kappa, 1.0)
    Nu = c.nusselt()
    e = abs(Nu - Nur) / abs(Nur)
    print(f"  Nu {Nu:.10f} (oracle {Nur:.10f}), rel err {e:.3e}", flush=True)
    assert e < BOUND[dtype]["Nu"], e
    c.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rbc3d_nusselt_over_horizon(gpu, oracle, dtype):
    nx, N, steps, dt = 16, 17, 200, 5e-3
    print(f"3-D Rayleigh-Benard Nu over {steps} steps {dtype.__name__}", flush=True)
    kappa = nu = 1e5 ** -0.5
    d, B, (U, V, W) = rbc_ic(oracle, nx, N, ny=nx)
    ref = oracle.RBC3D(nx, nx, N, 4.0, 4.0, 1.0, kappa, nu, dt, "SBDF2")
    c = gpu.chan3(nx, nx, N, 4.0, 4.0, 1.0, kappa, nu, dt, dtype=dtype)
    for var, v in ((oracle.RBC3_B, B), (oracle.RBC3_UX, U), (oracle.RBC3_UY, V), (oracle.RBC3_UZ, W)):
        ref.set(var, v)
        c.set(var, v)
    ref.step(steps)
    c.step(steps)
    Nur = diagnostics.nusselt3(ref.get(oracle.RBC3_B), ref.get(oracle.RBC3_UZ), nx, nx, 1.0, kappa, 1.0)
    Nu = c.nusselt()
    e = abs(Nu - Nur) / abs(Nur)
    print(f"  Nu {Nu:.10f} (oracle {Nur:.10f}), rel err {e:.3e}", flush=True)
    assert e < BOUND[dtype]["Nu"], e
    c.close()
