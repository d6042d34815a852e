"""Several 2-D channel problems stepped as one group (tarang_chan_group_*): every implicit
stage solves the per-mode tau systems of all members in one launch (uneven nx and nz, their
own dt, Rayleigh numbers and boxes).  Each member must give the same bits as the same
problem stepped alone, and the FP64 members match the oracle to 1e-12.
"""
import ctypes

import numpy as np
import pytest

from tarang_amd import _lib
from tests.conftest import relerr
from tests.test_gpu_channel import TS, params, random_state

pytestmark = pytest.mark.gpu

# uneven members: nx, nz, Lx, Lz, Ra, dt
MEMBERS = [(64, 17, 4.0, 1.0, 1e5, 2e-3), (128, 33, 2.0, 1.5, 3e5, 1e-3), (32, 24, 4.0, 1.0, 1e4, 3e-3)]


def build(gpu, oracle, dtype, ts, graph=True):
    out = []
    for i, (nx, N, Lx, Lz, Ra, dt) in enumerate(MEMBERS):
        kappa, nu = params(Ra)
        s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS[ts], dtype=dtype)
        s.set_graph(graph)
        for var, c in random_state(oracle, nx, N, Lx, Lz, seed=5 + i).items():
            s.set(var, c)
        out.append(s)
    return out


def make_group(members):
    arr = (ctypes.c_void_p * len(members))(*[m.h.value for m in members])
    g = ctypes.c_void_p()
    _lib.call("tarang_chan_group_create", arr, len(members), ctypes.byref(g))
    return g


def fields(oracle, s):
    return [s.get(v) for v in (oracle.RBC_P, oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ)] + [s.tau(t) for t in range(7)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts", ["RK222", "SBDF2"])
def test_group_same_bits_as_alone(gpu, oracle, ts, dtype):
    """3 + 2 group steps (graph replays, with a dt change of one member in between) against
    each member stepped alone by the same sequence."""
    print(f"channel group {ts} {dtype.__name__}", flush=True)
    alone = build(gpu, oracle, dtype, ts)
    grouped = build(gpu, oracle, dtype, ts)
    g = make_group(grouped)
    for s in alone:
        s.step(3)
    _lib.call("tarang_chan_group_step", g, 3)
    alone[1].set_dt(0.5 * MEMBERS[1][5])
    grouped[1].set_dt(0.5 * MEMBERS[1][5])
    for s in alone:
        s.step(2)
    _lib.call("tarang_chan_group_step", g, 2)
    gpu.sync()
    for i, (a, b) in enumerate(zip(alone, grouped)):
        for x, y in zip(fields(oracle, a), fields(oracle, b)):
            assert np.array_equal(x, y), f"member {i}"
        ta, tb = ctypes.c_double(), ctypes.c_double()
        ia, ib = ctypes.c_long(), ctypes.c_long()
        _lib.call("tarang_chan_time", a.h, ctypes.byref(ta), ctypes.byref(ia))
        _lib.call("tarang_chan_time", b.h, ctypes.byref(tb), ctypes.byref(ib))
        assert (ta.value, ia.value) == (tb.value, ib.value) and ib.value == 5
    # members may still step alone after group steps
    alone[0].step(1)
    grouped[0].step(1)
    for x, y in zip(fields(oracle, alone[0]), fields(oracle, grouped[0])):
        assert np.array_equal(x, y)
    _lib.call("tarang_chan_group_destroy", g)
    for s in alone + grouped:
        s.close()


def test_group_eager_matches_graph_and_oracle(gpu, oracle):
    """FP64 RK222: eager group steps equal graph group steps, and every member matches the
    oracle stepping that member alone to 1e-12."""
    eager = build(gpu, oracle, np.float64, "RK222", graph=False)
    graph = build(gpu, oracle, np.float64, "RK222")
    ge, gg = make_group(eager), make_group(graph)
    _lib.call("tarang_chan_group_set_graph", ge, 0)
    _lib.call("tarang_chan_group_step", ge, 4)
    _lib.call("tarang_chan_group_step", gg, 4)
    gpu.sync()
    for i, (a, b) in enumerate(zip(eager, graph)):
        for x, y in zip(fields(oracle, a), fields(oracle, b)):
            assert np.array_equal(x, y), f"member {i}"
    for i, (nx, N, Lx, Lz, Ra, dt) in enumerate(MEMBERS):
        kappa, nu = params(Ra)
        ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, "RK222")
        for var, c in random_state(oracle, nx, N, Lx, Lz, seed=5 + i).items():
            ref.set(var, c)
        ref.step(4)
        errs = [relerr(graph[i].get(v), ref.get(v)) for v in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ)]
        print(f"  member {i} ({nx} x {N}): rel errs {errs}", flush=True)
        assert max(errs) < 1e-12, (i, errs)
    for g in (ge, gg):
        _lib.call("tarang_chan_group_destroy", g)
    for s in eager + graph:
        s.close()


def test_group_rejects_mismatched_members(gpu, oracle):
    kappa, nu = params()
    a = gpu.chan(32, 17, 4.0, 1.0, kappa, nu, 1e-3, stepper=TS["RK222"], dtype=np.float64)
    b = gpu.chan(32, 17, 4.0, 1.0, kappa, nu, 1e-3, stepper=TS["SBDF2"], dtype=np.float64)
    c = gpu.chan(32, 17, 4.0, 1.0, kappa, nu, 1e-3, stepper=TS["RK222"], dtype=np.float32)
    for bad in ([a, b], [a, c], [a, a]):
        with pytest.raises(_lib.TarangError):
            make_group(bad)
    with pytest.raises(_lib.TarangError):
        _lib.call("tarang_chan_group_create", (ctypes.c_void_p * 1)(a.h.value), 0, ctypes.byref(ctypes.c_void_p()))
    for s in (a, b, c):
        s.close()
