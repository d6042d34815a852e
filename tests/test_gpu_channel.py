"""2-D Rayleigh-Benard channel on the GPU (tarang_chan_*) vs the FP64 oracle.

Tolerances (max-norm, relative to the largest oracle entry): FP64 1e-12 for a
right-hand side and single steps, FP32 1e-5 (x10 where a Chebyshev series of
length N is summed in FP32).  At the 1024 x 256 workload the oracle cannot hold
its dense per-mode matrices, so the check there is mode-by-mode: a small-amplitude
state confined to modes {0, 1, 3} evolves under the same per-mode systems and the
same (alias-free) product interactions on a 16 x 256 oracle grid.
"""
import numpy as np
import pytest

from tarang_amd import _lib
from tests.conftest import relerr

pytestmark = pytest.mark.gpu

TS = {"RK222": 0, "RK443": 1, "SBDF2": 2, "SBDF1": 4}
TOL = {np.float64: 1e-12, np.float32: 1e-5}


def params(Ra=1e5, Pr=1.0):
    return (Ra * Pr) ** -0.5, (Ra / Pr) ** -0.5


def zgrid(N, Lz):
    return Lz * (1 - np.cos(np.pi * np.arange(N) / (N - 1))) / 2


def random_state(O, nx, N, Lx, Lz, seed=3, amp=0.05, raw=1e-3):
    """Conduction profile plus seeded noise vanishing at the walls; `raw` adds unprojected
    coefficient noise (imaginary parts at kx = 0 and kx = Nx/2 included)."""
    rng = np.random.default_rng(seed)
    d = O.domain([O.REALFOURIER, O.CHEBYSHEVT], [nx, N], bounds=[(0.0, Lx), (0.0, Lz)])
    z = zgrid(N, Lz)
    wall = (z * (Lz - z))[None, :]
    grids = {O.RBC_B: (Lz - z)[None, :] + amp * rng.standard_normal((nx, N)) * wall,
             O.RBC_UX: amp * rng.standard_normal((nx, N)) * wall,
             O.RBC_UZ: amp * rng.standard_normal((nx, N)) * wall}
    out = {}
    for var, g in grids.items():
        c = O.forward_transform(d, np.broadcast_to(g, (nx, N)).copy())
        if raw:
            c = c + raw * (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
        out[var] = np.asarray(c)
    return out


def projected(c, nx):
    c = c.copy()
    c[0] = c[0].real
    c[nx // 2] = c[nx // 2].real
    return c


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx,N", [(16, 12), (64, 33), (128, 64)])
def test_channel_rhs_parity(gpu, oracle, dtype, nx, N):
    print(f"channel rhs nx={nx} N={N} {dtype.__name__}", flush=True)
    Lx, Lz = 4.0, 1.0
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, 1e-2)
    for var, c in st.items():
        ref.set(var, projected(c, nx))
    Fr = ref.rhs()
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, 1e-2, dtype=dtype)
    for var, c in st.items():
        s.set(var, c)
    Fg = s.rhs()
    tol = TOL[dtype]
    errs = [relerr(a, b) for a, b in zip(Fg, Fr)]
    print(f"  rel errs {errs}", flush=True)
    assert max(errs) < tol
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts", ["RK222", "RK443"])
def test_channel_steps_parity(gpu, oracle, ts, dtype):
    nx, N, Lx, Lz, dt = 32, 24, 4.0, 1.0, 5e-3
    print(f"channel steps {ts} {dtype.__name__}", flush=True)
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz, seed=11)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, ts)
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS[ts], dtype=dtype)
    for var, c in st.items():
        ref.set(var, c)
        s.set(var, c)
    tol = TOL[dtype]
    for nsteps in (1, 5):
        ref.step(nsteps)
        s.step(nsteps)
        errs = [relerr(s.get(var), ref.get(var)) for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ)]
        print(f"  after {nsteps}: rel errs {errs}", flush=True)
        assert max(errs) < tol, (nsteps, errs)
        assert np.abs(s.get(oracle.RBC_P)).max() == 0.0
        for t in range(7):
            assert np.abs(s.tau(t)).max() == 0.0
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts", ["RK222", "SBDF2"])
def test_channel_large_nz_parity(gpu, oracle, ts, dtype):
    """nz = 450: u-p systems of 678 unknowns, whose inverses are formed 8 columns per
    workgroup (16 do not fit in LDS), and four solution rows per thread in the solve."""
    nx, N, Lx, Lz, dt = 8, 450, 4.0, 1.0, 5e-3
    print(f"channel large nz {ts} {dtype.__name__}", flush=True)
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz, seed=17)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, ts)
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS[ts], dtype=dtype)
    for var, c in st.items():
        ref.set(var, c)
        s.set(var, c)
    ref.step(2)
    s.step(2)
    # at nz = 450 the u-p systems have condition numbers ~1e10: an FP64 LU solve and an
    # FP64 explicit inverse each differ from the exact solution by up to ~3e-12 relative
    # (checked against iteratively refined solves), so the two paths agree to ~1e-11.
    # FP32 at this size: 1e-4 (measured 6e-5 with split inverses): the FP32 Chebyshev sums
    # over 450 terms and the growth of the z derivatives with Nz put the FP32-state floor
    # above 1e-5 here; the workload sizes (Nz <= 256) meet 1e-5
    tol = {np.float64: 1e-10, np.float32: 1e-4}[dtype]
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        assert relerr(s.get(var), ref.get(var)) < tol, var
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_channel_benchmark_x_size_steps(gpu, oracle, dtype):
    """The benchmarked 2-D configuration's x size and parameters (Nx = 1024, so the
    x pass runs at M = 1536; Lx = 4, Ra = 2e6, Pr = 1, RK222, dt = 0.01) at Nz = 64,
    where the oracle's per-mode factorisations take about a second: two steps."""
    nx, N, Lx, Lz, dt = 1024, 64, 4.0, 1.0, 1e-2
    print(f"channel benchmark x size {dtype.__name__}", flush=True)
    kappa, nu = params(Ra=2e6, Pr=1.0)
    st = random_state(oracle, nx, N, Lx, Lz, seed=17)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, "RK222")
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS["RK222"], dtype=dtype)
    for var, c in st.items():
        ref.set(var, c)
        s.set(var, c)
    ref.step(2)
    s.step(2)
    tol = TOL[dtype]
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        assert relerr(s.get(var), ref.get(var)) < tol, var
    s.close()


def test_channel_graph_matches_eager(gpu, oracle):
    nx, N, Lx, Lz, dt = 32, 16, 4.0, 1.0, 5e-3
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz, seed=5)
    out = []
    for graph in (False, True):
        s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt)
        s.set_graph(graph)
        for var, c in st.items():
            s.set(var, c)
        s.step(3)
        out.append([s.get(v) for v in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ)])
        s.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_channel_dt_change_refactors(gpu, oracle):
    """A new dt refactors the per-mode systems (factor pairing on dt change)."""
    nx, N, Lx, Lz = 16, 16, 4.0, 1.0
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz, seed=8)
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, 1e-2)
    r1 = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, 1e-2)
    for var, c in st.items():
        s.set(var, c)
        r1.set(var, c)
    s.step(2)
    r1.step(2)
    r2 = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, 4e-3)
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        r2.set(var, r1.get(var))
    s.set_dt(4e-3)
    s.step(3)
    r2.step(3)
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        assert relerr(s.get(var), r2.get(var)) < 1e-12
    s.close()


def test_channel_conduction_steady_and_bc(gpu, oracle):
    nx, N, Lx, Lz = 64, 32, 4.0, 1.0
    kappa, nu = params()
    d = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [nx, N], bounds=[(0.0, Lx), (0.0, Lz)])
    z = zgrid(N, Lz)
    prof = 2.0 - 1.5 * z
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, 1e-2, b_bottom=2.0, b_top=0.5)
    s.set(oracle.RBC_B, oracle.forward_transform(d, np.tile(prof, (nx, 1))))
    s.step(4)
    b = oracle.backward_transform(d, s.get(oracle.RBC_B))
    assert np.abs(b - prof[None, :]).max() < 1e-13
    assert np.abs(s.get(oracle.RBC_UZ)).max() < 1e-13
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_channel_full_size_modewise(gpu, oracle, dtype):
    """1024 x 256 at the channel workload's parameters (Ra = 2e6, Pr = 1, RK222, dt = 0.01)
    vs a 16 x 256 oracle on modes {0, 1, 3}."""
    nx, N, Lx, Lz, dt, ts = 1024, 256, 4.0, 1.0, 1e-2, "RK222"
    print(f"channel full size {dtype.__name__}", flush=True)
    kappa, nu = params(2e6)
    small = 16
    rng = np.random.default_rng(21)
    ds = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [small, N], bounds=[(0.0, Lx), (0.0, Lz)])
    z = zgrid(N, Lz)
    wall = z * (Lz - z)
    st = {}
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        c = np.zeros((small // 2 + 1, N), complex)
        if var == oracle.RBC_B:
            c = np.asarray(oracle.forward_transform(ds, np.tile(Lz - z, (small, 1))))
        for m in (1, 3):
            prof = np.polynomial.chebyshev.chebfit(2 * z / Lz - 1, wall * np.sin((m + 1) * np.pi * z / Lz), 40)
            c[m, :41] += 1e-6 * small * prof * np.exp(1j * rng.uniform(0, 2 * np.pi))
        st[var] = c
    ref = oracle.RBC2D(small, N, Lx, Lz, kappa, nu, dt, ts)
    for var, c in st.items():
        ref.set(var, c)
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS[ts], dtype=dtype)
    scale = nx / small   # unnormalised rfft: coefficients scale with Nx
    for var, c in st.items():
        big = np.zeros((nx // 2 + 1, N), complex)
        big[:small // 2 + 1] = c * scale
        big[small // 2] = 0   # the small grid's Nyquist column is empty
        s.set(var, big)
    ref.step(3)
    s.step(3)
    tol = {np.float64: 1e-12, np.float32: 1e-5}[dtype]
    b0 = np.abs(ref.get(oracle.RBC_B)[0]).max()  # the O(1) conduction profile carried by kx = 0
    worst = 0.0
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        g = s.get(var) / scale
        r = ref.get(var)
        for m in (0, 1, 3):
            # modes 1, 3 against the perturbation size; at kx = 0 (where uz is roundoff) the
            # solve's rounding scales with the conduction profile of that mode
            den = max(np.abs(r[m]).max(), np.abs(r[1]).max(), b0 if m == 0 else 0.0)
            err = np.abs(g[m] - r[m]).max() / den
            worst = max(worst, err)
            assert err < tol, (var, m, err)
        assert np.abs(g[small // 2 + 1:]).max() < 1e-9 * np.abs(r[1]).max()
    print(f"  worst mode-wise rel err {worst:.3e}", flush=True)
    s.close()


def test_channel_rejects_unsupported(gpu):
    kappa, nu = params()
    with pytest.raises(_lib.TarangError) as e:
        gpu.chan(24, 16, 4.0, 1.0, kappa, nu, 1e-2)
    assert e.value.code == _lib.ERR_INVALID and "RealFourier size 24" in str(e.value)
    with pytest.raises(_lib.TarangError) as e:
        gpu.chan(32, 600, 4.0, 1.0, kappa, nu, 1e-2)
    assert e.value.code == _lib.ERR_INVALID
    with pytest.raises(_lib.TarangError) as e:
        gpu.chan(32, 16, 4.0, 1.0, kappa, nu, 1e-2, stepper=3)  # RK111: no explicit first stage
    assert e.value.code == _lib.ERR_INVALID
    with pytest.raises(_lib.TarangError):
        gpu.chan(2048, 16, 4.0, 1.0, kappa, nu, 1e-2, dtype=np.float64)
    with pytest.raises(_lib.TarangError) as e:  # split inverses belong to the FP32 state
        gpu.chan(32, 16, 4.0, 1.0, kappa, nu, 1e-2, dtype=np.float64, split=1)
    assert e.value.code == _lib.ERR_INVALID and "split_inverse" in str(e.value)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts", ["SBDF2", "SBDF1"])
def test_channel_sbdf_parity(gpu, oracle, ts, dtype):
    """Per-mode multistep (step_subproblem_multistep.jl:197-420): SBDF1 start, SBDF2 after,
    a dt change (w != 1) half way; p and the taus come from the solve."""
    nx, N, Lx, Lz = 32, 24, 4.0, 1.0
    print(f"channel {ts} {dtype.__name__}", flush=True)
    kappa, nu = params()
    st = random_state(oracle, nx, N, Lx, Lz, seed=13)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, 5e-3, ts)
    s = gpu.chan(nx, N, Lx, Lz, kappa, nu, 5e-3, stepper={"SBDF2": 2, "SBDF1": 4}[ts], dtype=dtype)
    for var, c in st.items():
        ref.set(var, c)
        s.set(var, c)
    tol = TOL[dtype]
    for nsteps, dt in ((1, 5e-3), (3, 5e-3), (2, 2.5e-3), (2, 4e-3)):
        ref.set_dt(dt)
        s.set_dt(dt)
        ref.step(nsteps)
        s.step(nsteps)
        for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ, oracle.RBC_P):
            assert relerr(s.get(var), ref.get(var)) < tol * (100 if var == oracle.RBC_P else 1), (nsteps, dt, var)
        # taus against the largest tau of the step (tau_p is roundoff around 0)
        scale = max(np.abs(ref.tau(t)).max() for t in range(7))
        for t in range(7):
            assert np.abs(s.tau(t) - ref.tau(t)).max() <= tol * 100 * scale, (t,)
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts", ["RK222", "SBDF2"])
def test_channel_refinement_sweeps(gpu, oracle, ts, dtype):
    """FP64-residual refinement of the per-mode solves (refine_sweeps 1 and 2) and, in FP32,
    split inverses (FP32 + 8-bit remainder, alone and with a sweep) against the oracle
    at the 2-D workload's parameters (Ra = 2e6, Pr = 1, Nz = 129): FP64 stays at 1e-12; in
    FP32 the error of the FP32-rounded inverses is removed, so the state error is at most
    that of the plain FP32 form and within 1e-5."""
    nx, N, Lx, Lz, dt = 32, 129, 4.0, 1.0, 1e-2
    kappa, nu = params(Ra=2e6)
    st = random_state(oracle, nx, N, Lx, Lz, seed=13)
    ref = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, ts)
    for var, c in st.items():
        ref.set(var, c)
    ref.step(3)
    want = {v: ref.get(v) for v in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ)}
    errs = {}
    forms = [(0, 0), (1, 0), (2, 0)] + ([(0, 1), (1, 1)] if dtype == np.float32 else [])
    for sweeps, split in forms:
        print(f"channel refinement {ts} sweeps={sweeps} split={split} {dtype.__name__}", flush=True)
        s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=TS[ts], dtype=dtype, refine=sweeps, split=split)
        for var, c in st.items():
            s.set(var, c)
        s.step(3)
        errs[(sweeps, split)] = max(relerr(s.get(v), w) for v, w in want.items())
        s.close()
    print(f"  max rel err by (sweeps, split): {errs}", flush=True)
    if dtype == np.float64:
        assert max(errs.values()) < 1e-12, errs
    else:
        for f in forms[1:]:
            assert errs[f] <= errs[(0, 0)] * 1.05 and errs[f] < 1e-5, (f, errs)


def test_channel_fp32_forms_long_run(gpu):
    """The FP32 forms against the FP64 form (itself pinned to the oracle at 1e-12) over 100
    RK222 steps of the 2-D workload's physics (Ra = 2e6, Pr = 1, dt = 1e-2, Nz = 256,
    conduction plus noise) on 32 x 256: split inverses (FP32 + 8-bit remainder) and
    refinement sweeps remove the error of the FP32-rounded inverses; b stays within 1e-6
    of FP64."""
    nx, N, Lx, Lz, dt, steps = 32, 256, 4.0, 1.0, 1e-2, 100
    kappa, nu = params(Ra=2e6)
    z = Lz * (1 - np.cos(np.pi * np.arange(N) / (N - 1))) / 2
    rng = np.random.default_rng(42)
    bg = (Lz - z)[None, :] + 1e-3 * rng.standard_normal((nx, N)) * (z * (Lz - z))[None, :]
    k = np.arange(N)
    T = np.cos(np.outer(np.pi * np.arange(N) / (N - 1), k)) * ((-1.0) ** k)[None, :]
    Bc = np.linalg.solve(T, np.fft.rfft(bg, axis=0).T).T
    out = {}
    for name, dtype, sweeps, split in (("f64", np.float64, 0, 0), ("plain", np.float32, 0, 0),
                                       ("split", np.float32, 0, 1), ("sweeps1", np.float32, 1, 0)):
        print(f"channel long run {name}", flush=True)
        s = gpu.chan(nx, N, Lx, Lz, kappa, nu, dt, stepper=0, dtype=dtype, refine=sweeps, split=split)
        s.set(_lib.CHAN_B, Bc)
        for v in (_lib.CHAN_P, _lib.CHAN_UX, _lib.CHAN_UZ):
            s.set(v, np.zeros_like(Bc))
        s.step(steps)
        out[name] = s.get(_lib.CHAN_B)
        s.close()
    errs = {n: relerr(out[n], out["f64"]) for n in ("plain", "split", "sweeps1")}
    print(f"  b rel err vs FP64 after {steps} steps: {errs}", flush=True)
    assert errs["split"] < 1e-6 and errs["sweeps1"] < 1e-6, errs
    assert errs["split"] <= errs["plain"] and errs["sweeps1"] <= errs["plain"], errs


def test_channel_rhs_full_size(gpu, oracle):
    """The right-hand side at the 2-D workload's size (1024 x 256, Lx = 4) on a full-spectrum
    random state, against the oracle's right-hand side (no per-mode matrices): every mode,
    the Nyquist line and the 1536-point padded products included.  FP64 1e-12, FP32 1e-5."""
    nx, N, Lx, Lz = 1024, 256, 4.0, 1.0
    print("channel rhs full size", flush=True)
    rng = np.random.default_rng(22)
    ref = oracle.RBC2D(nx, N, Lx, Lz, 0.0, 0.0, 0.0, rhs_only=True)
    st = {}
    for var in (oracle.RBC_B, oracle.RBC_UX, oracle.RBC_UZ):
        c = 1e-3 * (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape))
        st[var] = c
        ref.set(var, projected(c, nx))
    Fr = ref.rhs()
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        s = gpu.chan(nx, N, Lx, Lz, 1e-3, 1e-3, 1e-2, dtype=dtype)
        for var, c in st.items():
            s.set(var, c)
        errs = [relerr(a, b) for a, b in zip(s.rhs(), Fr)]
        print(f"  {dtype.__name__} rel errs {errs}", flush=True)
        assert max(errs) < tol, errs
        s.close()
