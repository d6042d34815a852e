"""Space- and time-dependent wall values of b (tarang_chan_set_bc, tarang_chan3_set_bc):
the GPU solvers against the oracle fed the same per-mode values (the unnormalised rfft of
the wall arrays, numpy), 1e-12 relative in FP64 (1e-5 for FP32 data, the channel tests'
bar), with the walls changed between steps; and, for SBDF2, whose new state is one
implicit solve, the tau-method invariant that b equals the wall arrays on z = 0 and
z = Lz after every step (1e-10 of max |b|).  The IMEX RK final update
M X = M X0 + dt sum(bE F - bI LX) does not re-impose the wall rows (the oracle drifts
by ~1e-8 there too), so RK is held to the oracle only."""
import numpy as np
import pytest

import tarang_amd as T
from tests.conftest import relerr
from tests.test_gpu_channel import TS, params, random_state, zgrid

pytestmark = pytest.mark.gpu


def walls(nx, Lx, t, ny=None, Ly=None):
    x = Lx * np.arange(nx) / nx
    if ny is None:
        g0 = 1.0 + 0.3 * np.cos(2 * np.pi * x / Lx + t) + 0.1 * np.sin(4 * np.pi * x / Lx)
        g1 = 0.2 * np.sin(2 * np.pi * x / Lx - t)
        return g0, g1
    y = Ly * np.arange(ny) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    g0 = 1.0 + 0.3 * np.cos(2 * np.pi * X / Lx + t) * np.cos(2 * np.pi * Y / Ly)
    g1 = 0.2 * np.sin(2 * np.pi * (X / Lx + Y / Ly) - t)
    return g0, g1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts", ["RK222", "SBDF2"])
def test_channel_wall_arrays(gpu, oracle, ts, dtype):
    nx, N, Lx, Lz, dt = 32, 24, 4.0, 1.0, 5e-3
    print(f"channel wall arrays {ts} {dtype.__name__}", flush=True)
    kappa, nu = params(Ra=1e4)
    st = random_state(oracle, nx, N, Lx, Lz, seed=5, raw=0.0)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, ts)
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS[ts], dtype=dtype)
    for var, c in st.items():
        ref.set(var, c)
        s.set(var, c)
    d = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [nx, N], bounds=[(0.0, Lx), (0.0, Lz)])
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for t in (0.0, 0.7):  # walls changed between steps
        g0, g1 = (g.astype(dtype).astype(np.float64) for g in walls(nx, Lx, t))
        s.set_bc(g0, g1)
        ref.set_bc_values(np.fft.rfft(g0), np.fft.rfft(g1))
        for _ in range(3):
            ref.step(1)
            s.step(1)
            if ts == "SBDF2":
                grid = oracle.backward_transform(d, s.get(oracle.RBC_B).astype(np.complex128))
                scale = np.abs(grid).max()
                assert np.abs(grid[:, 0] - g0).max() < (1e-10 if dtype == np.float64 else 1e-5) * scale
                assert np.abs(grid[:, -1] - g1).max() < (1e-10 if dtype == np.float64 else 1e-5) * scale
        for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
            assert relerr(s.get(var), ref.get(var)) < tol * (1 if var == oracle.RBC_B else 10), (t, var)
    s.close()


def test_channel3d_wall_arrays(gpu, oracle):
    nx, ny, N, Lx, Ly, Lz, dt = 16, 8, 12, 4.0, 2.0, 1.0, 5e-3
    print("channel3d wall arrays", flush=True)
    kappa, nu = params(Ra=1e4)
    d = oracle.domain([oracle.REALFOURIER, oracle.REALFOURIER, oracle.CHEBYSHEVT], [nx, ny, N],
                      bounds=[(0.0, Lx), (0.0, Ly), (0.0, Lz)])
    z = zgrid(N, Lz)
    rng = np.random.default_rng(9)
    b0 = (Lz - z)[None, None, :] + 0.02 * rng.standard_normal((nx, ny, N)) * (z * (Lz - z))[None, None, :]
    ref = oracle.RBC3D(nx, ny, N, Lx, Ly, Lz, kappa, nu, dt, "SBDF2")
    s = gpu.chan3(nx, ny, N, Lx, Ly, Lz, kappa, nu, dt, stepper=TS["SBDF2"])
    c = oracle.forward_transform(d, b0)
    ref.set(oracle.RBC3_B, c)
    s.set(oracle.RBC3_B, c)
    for t in (0.0, 0.5):
        g0, g1 = walls(nx, Lx, t, ny, Ly)
        s.set_bc(g0, g1)
        ref.set_bc_values(np.fft.fft(np.fft.rfft(g0, axis=0), axis=1), np.fft.fft(np.fft.rfft(g1, axis=0), axis=1))
        ref.step(3)
        s.step(3)
        grid = oracle.backward_transform(d, s.get(oracle.RBC3_B))
        scale = np.abs(grid).max()
        assert np.abs(grid[:, :, 0] - g0).max() < 1e-10 * scale
        assert np.abs(grid[:, :, -1] - g1).max() < 1e-10 * scale
        for var in (oracle.RBC3_B, oracle.RBC3_UX, oracle.RBC3_UY, oracle.RBC3_UZ):
            assert relerr(s.get(var), ref.get(var)) < 1e-12, (t, var)
    s.close()


def test_wall_array_bc_through_api(oracle):
    """b(z=0) = gbot with gbot a grid array parameter (a space-dependent BC, the reference's
    ArrayOperator); the array is modified in place between run_ calls and the change is seen."""
    from tests.test_gpu_api import rbc_problem
    print("wall array BC through the API", flush=True)
    Lx, Lz, Nx, Nz, dt = 4.0, 1.0, 16, 10, 0.01
    kappa = nu = 1e-2
    g0, g1 = walls(Nx, Lx, 0.0)
    solver, b, u, xb, zb = rbc_problem(np.float64, Lx, Lz, Nx, Nz, kappa, nu, dt=dt, bcs=("gbot", "gtop"),
                                       params={"gbot": g0, "gtop": g1})
    z = zb.grid()
    bg = np.outer(g0, 1 - z / Lz) + np.outer(g1, z / Lz)
    b["g"] = bg
    T.run_(solver, stop_iteration=3)
    g0[:] = walls(Nx, Lx, 0.4)[0]  # in place: the parameter holds the same array
    T.run_(solver, stop_iteration=6)
    ref = oracle.RBC2D(Nx, Nz, Lx, Lz, kappa, nu, dt, "RK222")
    d = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [Nx, Nz], bounds=[(0.0, Lx), (0.0, Lz)])
    ref.set(oracle.RBC_B, oracle.forward_transform(d, bg))
    ref.set_bc_values(np.fft.rfft(walls(Nx, Lx, 0.0)[0]), np.fft.rfft(g1))
    ref.step(3)
    ref.set_bc_values(np.fft.rfft(g0), np.fft.rfft(g1))
    ref.step(3)
    assert relerr(b["c"].cpu().numpy(), ref.get(oracle.RBC_B)) < 1e-12
    assert relerr(u[1]["c"].cpu().numpy(), ref.get(oracle.RBC_UZ)) < 1e-12
    solver.close()
