"""Pure-Fourier IVPs on the GPU vs the FP64 oracle (RHS, single steps, trajectories)."""
import numpy as np
import pytest

from tarang_amd import configs
from tests.conftest import relerr

pytestmark = pytest.mark.gpu

DIFF, TURB, BURG = 0, 1, 2
TS = {"RK222": 0, "RK443": 1, "SBDF2": 2}
TOL = {np.float64: 1e-12, np.float32: 1e-5}


def turb_ic(N, seed=7, kpeak=4.0):
    """Seeded random-phase vorticity, E(k) ~ k^4 exp(-2 (k/kpeak)^2), E = 0.5."""
    return configs.turbulence_2d(N, seed=seed, kpeak=kpeak)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("N", [16, 64, 256, 512, 1024])  # padded lengths 24, 96, 384, 768, 1536
def test_turbulence_rhs_parity(gpu, oracle, dtype, N):
    print(f"turbulence rhs N={N} {dtype.__name__}", flush=True)
    d = oracle.domain([0, 0], [N, N])
    q = turb_ic(N)
    C = oracle.forward_transform(d, q)
    ref = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", 1e-4, 1e-3).rhs(C)
    s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS["RK443"], 1e-4, 1e-3, dtype=dtype)
    F = s.rhs(C)
    e = relerr(F, ref)
    print(f"  rel err {e:.3e}", flush=True)
    assert e < TOL[dtype]
    s.close()


def test_turbulence_rhs_closed_form(gpu, oracle, golden):
    g = golden("ivp.npz")
    d = oracle.domain([0, 0], [16, 16])
    s = gpu.fivp([16, 16], [2 * np.pi, 2 * np.pi], TURB, TS["RK443"], 0.0, 1e-3)
    F = s.rhs(oracle.forward_transform(d, g["turb_q"]))
    assert np.abs(oracle.backward_transform(d, F) - g["turb_F"]).max() < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("N", [32, 4096])
def test_burgers_rhs_parity(gpu, oracle, dtype, N):
    d = oracle.domain([0], [N])
    x = 2 * np.pi * np.arange(N) / N
    u = np.sin(x) + 0.1 * np.cos(3 * x + 0.3) + 0.05 * np.sin(7 * x)
    C = oracle.forward_transform(d, u)
    ref = oracle.FourierIVP(d, oracle.PROB_BURGERS1D, "RK222", 1e-2, 1e-3).rhs(C)
    s = gpu.fivp([N], [2 * np.pi], BURG, TS["RK222"], 1e-2, 1e-3, dtype=dtype)
    e = relerr(s.rhs(C), ref)
    print(f"burgers rhs N={N} {dtype.__name__} rel err {e:.3e}", flush=True)
    assert e < TOL[dtype]
    s.close()


@pytest.mark.parametrize("ts", ["RK222", "RK443", "SBDF2"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_turbulence_steps_parity(gpu, oracle, ts, dtype):
    N, nu, dt, nsteps = 64, 1e-3, 2e-3, 5
    print(f"turbulence steps {ts} {dtype.__name__}", flush=True)
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, turb_ic(N))
    ref = oracle.FourierIVP(d, oracle.PROB_TURB2D, ts, nu, dt)
    ref.set_coeff(C)
    s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS[ts], nu, dt, dtype=dtype)
    s.set_state(C)
    ref.step(1)
    s.step(1)
    assert relerr(s.get_state(), ref.get_coeff()) < TOL[dtype], "single step"
    ref.step(nsteps - 1)
    s.step(nsteps - 1)
    assert relerr(s.get_state(), ref.get_coeff()) < TOL[dtype] * 10, "multi step"
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_turbulence_benchmark_configuration_steps(gpu, oracle, dtype):
    """The benchmarked configuration itself: 1024^2, nu = 1e-4, dt = 1e-3, RK443, the
    seeded E(k) ~ k^4 exp(-2 (k/10)^2) initial vorticity with E = 0.5; two steps through
    the captured step graph against the oracle."""
    N, nu, dt = 1024, 1e-4, 1e-3
    print(f"turbulence benchmark configuration {dtype.__name__}", flush=True)
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, configs.turbulence_2d(N, seed=42, kpeak=10.0, energy=0.5))
    ref = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", nu, dt)
    ref.set_coeff(C)
    s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS["RK443"], nu, dt, dtype=dtype)
    s.set_state(C)
    ref.step(2)
    s.step(2)
    assert relerr(s.get_state(), ref.get_coeff()) < TOL[dtype]
    s.close()


def test_graph_replay_matches_eager(gpu, oracle):
    N = 64
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, turb_ic(N))
    out = []
    for graph in (False, True):
        s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS["RK443"], 1e-3, 2e-3)
        s.set_graph(graph)
        s.set_state(C)
        s.step(3)
        out.append(s.get_state())
        s.close()
    assert np.array_equal(out[0], out[1])


def test_sbdf2_dt_change_graph_matches_eager_and_oracle(gpu, oracle):
    """SBDF2 with one dt change, then several steps at the new dt: the step after the change
    runs at the ratio w = dt2/dt1, the following ones at w = 1 again.  The step graph must be
    recaptured for both (a graph baked at w != 1 replayed at w = 1 gives the wrong derivative)."""
    N, nu, dt1, dt2 = 64, 1e-3, 2e-3, 4e-3
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, turb_ic(N))
    ref = oracle.FourierIVP(d, oracle.PROB_TURB2D, "SBDF2", nu, dt1)
    ref.set_coeff(C)
    ref.step(3)
    ref.set_dt(dt2)
    ref.step(4)
    out = []
    for graph in (False, True):
        s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS["SBDF2"], nu, dt1)
        s.set_graph(graph)
        s.set_state(C)
        s.step(3)
        s.set_dt(dt2)
        s.step(4)
        out.append(s.get_state())
        s.close()
    assert np.array_equal(out[0], out[1]), "graph replay differs from eager stepping"
    assert relerr(out[1], ref.get_coeff()) < 1e-12


@pytest.mark.parametrize("ts", ["RK222", "RK443", "SBDF2"])
def test_diffusion_exact_decay(gpu, oracle, golden, ts):
    g = golden("ivp.npz")
    d = oracle.domain([0], [64])
    s = gpu.fivp([64], [2 * np.pi], DIFF, TS[ts], 0.01, 0.01)
    s.set_state(oracle.forward_transform(d, g["diff_T0"]))
    s.step(100)
    C = s.get_state()
    ref = oracle.FourierIVP(d, oracle.PROB_DIFFUSION, ts, 0.01, 0.01)
    ref.set_coeff(oracle.forward_transform(d, g["diff_T0"]))
    ref.step(100)
    assert relerr(C, ref.get_coeff()) < 1e-12
    assert relerr(oracle.backward_transform(d, C), g["diff_T1"]) < 2e-5


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_burgers_trajectory(gpu, oracle, dtype):
    N, nu, dt = 256, 1e-2, 1e-3
    d = oracle.domain([0], [N])
    x = 2 * np.pi * np.arange(N) / N
    r = np.random.default_rng(9)
    u0 = np.sin(x) + 0.1 * sum(r.standard_normal() * np.cos(k * x + r.uniform(0, 6.3)) for k in range(1, 5))
    C = oracle.forward_transform(d, u0)
    ref = oracle.FourierIVP(d, oracle.PROB_BURGERS1D, "RK222", nu, dt)
    ref.set_coeff(C)
    ref.step(200)
    s = gpu.fivp([N], [2 * np.pi], BURG, TS["RK222"], nu, dt, dtype=dtype)
    s.set_state(C)
    s.step(200)
    assert relerr(s.get_state(), ref.get_coeff()) < (1e-12 if dtype == np.float64 else 1e-5)
    s.close()


def test_full_size_turbulence_fp32_vs_fp64(gpu, oracle):
    """1024^2 (the benchmark grid): size-independent checks -- FP32 tracks FP64,
    energy and enstrophy decay, one RHS matches the oracle."""
    N = 1024
    q = turb_ic(N, kpeak=10.0)
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, q)
    ref_F = oracle.FourierIVP(d, oracle.PROB_TURB2D, "RK443", 1e-4, 1e-3).rhs(C)
    states = {}
    for dtype in (np.float64, np.float32):
        s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS["RK443"], 1e-4, 1e-3, dtype=dtype)
        if dtype == np.float64:
            assert relerr(s.rhs(C), ref_F) < 1e-12
        s.set_state(C)
        s.step(20)
        states[dtype] = s.get_state()
        s.close()
    assert relerr(states[np.float32], states[np.float64]) < 1e-4
    w = np.full(C.shape[0], 2.0)
    w[0] = w[-1] = 1.0  # half-spectrum weights of a real field
    Z0 = np.sum(w[:, None] * np.abs(C) ** 2)
    Z1 = np.sum(w[:, None] * np.abs(states[np.float64]) ** 2)
    assert Z1 < Z0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_every_tuning_variant_is_bit_identical(gpu, oracle, dtype):
    """The autotuner only changes tile width / workgroup size: every variant must
    produce the same bits as the default."""
    import ctypes
    from tarang_amd import _lib
    N = 256
    d = oracle.domain([0, 0], [N, N])
    C = oracle.forward_transform(d, turb_ic(N))
    s = gpu.fivp([N, N], [2 * np.pi, 2 * np.pi], TURB, TS["RK443"], 1e-4, 1e-3, dtype=dtype)
    chosen = (ctypes.c_int * 3)()
    _lib.call("tarang_fivp_get_tuning", s.h, chosen)
    print("autotuned variants", list(chosen), flush=True)
    base = None
    for yb in range(4):
        for xp in range(2):
            for yf in range(4):
                _lib.call("tarang_fivp_set_tuning", s.h, (ctypes.c_int * 3)(yb, xp, yf))
                F = s.rhs(C)
                if base is None:
                    base = F
                else:
                    assert np.array_equal(F, base), (yb, xp, yf)
    s.close()
