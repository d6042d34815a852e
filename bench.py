"""Benchmark: 2-D decaying turbulence, 1024x1024 RealFourier, RK443, dt=1e-3.

One step = one full RK443 timestep (4 dealiased RHS evaluations + stage updates)
of the vorticity/streamfunction system on the GPU, state resident in HBM.
Reports grid-point steps/s for the FP32 path (value) and the FP64 path, the
FP32-vs-FP64 error, the dominant kernel's roofline figure measured live with HIP
events, and the FP64 oracle timed on the host cores on a bounded sample.

`--workload channel` measures the 2-D Rayleigh-Benard channel instead
(1024x256 RealFourier x ChebyshevT, Ra=2e6, Pr=1, RK222, dt=0.01): one step =
3 stages, 3 dealiased RHS evaluations, 2 per-mode FP64 tau solves.

`--workload plugins` measures the stand-alone plug-ins instead, one JSON line
each: the Chebyshev grid derivative on the 256x256x128 grid (along z and x, FP64
and FP32 data) and the batched complex LU of the 2-D channel's stage matrices
(513 modes, n = 1031: factor and solve).

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--no-cpu]
                       [--workload turbulence|channel|channel3d|serving|plugins]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "tarang.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
N, NU, DT = 1024, 1e-4, 1e-3
WORKLOAD = "2-D decaying turbulence, vorticity-streamfunction, 1024x1024 RealFourier, 3/2 dealiasing, nu=1e-4, RK443, dt=1e-3"
KERNELS = ["k_ybwd (y-backward: multipliers, padding, inverse y FFT M=1536)",
           "k_xpass (inverse x FFTs, u.grad(q) product, forward real FFT M=1536)",
           "k_yfwd (forward y FFT M=1536, truncation, Nyquist folds)",
           "k_rk_stage (IMEX stage / final update)"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(name, kernel_label, dtype="float", variant=None):
    """(HBM bytes per launch, full template name) of a kernel from a committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/<name>), keyed by the kernels' full template
    names: the one entry for this kernel and dtype (and `variant`, a piece of its template
    argument list), or (None, None) when there is not exactly one."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    short = kernel_label.split()[0]
    if "kernels" not in d:  # summaries keyed by the short kernel name
        return d.get(short), d.get(short + "_counters", {}).get("kernel")
    cands = [k for k in d["kernels"] if k.startswith(short + "<" + dtype)]
    if variant is not None:
        cands = [k for k in cands if variant in k]
    if len(cands) != 1:
        return None, None
    return d["kernels"][cands[0]]["bytes"], cands[0]


def host_threads():
    """Threads for the CPU baseline: the cores this process may run on, at most 16 (a host
    shared by several GPUs gives each GPU a 16-core share)."""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    return max(1, min(visible, 16))


def host_info():
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = None
    return {"nproc": os.cpu_count(), "affinity_cores": visible, "threads_used": host_threads(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def calibration_copy(mib):
    """One plain device-to-device copy of `mib` MiB: the reference reading for HBM byte counters."""
    import torch

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"calibration copy: {mib} MiB read + {mib} MiB written")
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one plain device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--workload", choices=["all", "turbulence", "channel", "channel3d", "serving", "plugins"],
                    default="all")
    ap.add_argument("--requests", type=int, default=24, help="serving: number of requests")
    ap.add_argument("--cfl-horizon", action="store_true",
                    help="channel: also run the configuration to its t = 20 horizon with CFL-controlled dt "
                         "(FP64 and FP32 split-inverse forms through the API)")
    ap.add_argument("--no-group", action="store_true",
                    help="channel: skip the group-stepping comparison (used for profiler runs)")
    ap.add_argument("--cfl-wall", type=float, default=480.0,
                    help="channel --cfl-horizon: wall-clock limit per form in seconds")
    ap.add_argument("--horizon", type=int, default=None,
                    help="channel workloads: steps before the FP32-vs-FP64 comparison (default: the timed steps "
                         "in 2-D, the configuration's 500 in 3-D)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (no collectives); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_mib > 0:
        calibration_copy(a.copy_mib)
    if a.workload == "plugins":
        return bench_plugins(a)
    run = {"turbulence": bench_turbulence, "channel": bench_channel, "serving": bench_serving,
           "channel3d": bench_channel3d}
    if a.workload != "all":
        print(json.dumps(run[a.workload](a)), flush=True)
        return
    # every configuration in one run; the line is the 3-D Rayleigh-Benard configuration (the
    # metric's FP32+refinement vs FP64 comparison), the others are summarised inside it
    others = {}
    for name, fn, steps in (("turbulence", bench_turbulence, 200), ("channel", bench_channel, 100),
                            ("serving", bench_serving, None)):
        sub = argparse.Namespace(**vars(a))
        sub.no_cpu = True
        if steps is not None and a.steps == 1000:
            sub.steps = steps
        if name == "serving":
            sub.requests = min(a.requests, 16)
        log(f"all: {name}")
        r = fn(sub)
        others[r["config"]["workload"]] = {k: r.get(k) for k in (
            "value", "unit", "ms_per_step", "fp64", "speedup_fp32_vs_fp64", "fp32_form", "fp32_vs_fp64_rel_err",
            "fp32_forms", "roofline", "latency_s", "makespan_s", "rejected") if r.get(k) is not None}
    log("all: channel3d (headline)")
    line = bench_channel3d(a)
    line["other_configs"] = others
    print(json.dumps(line), flush=True)


def bench_turbulence(a):

    import numpy as np
    import torch

    from tarang_amd import _lib, configs

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda", 0)
    log(f"building initial condition {N}x{N}")
    q0 = configs.turbulence_2d(N, seed=42, kpeak=10.0, energy=0.5)
    C0 = np.fft.fft(np.fft.rfft(q0, axis=0), axis=1)  # reference layout (kx, ky), unnormalised
    E0, Z0 = configs.energy_enstrophy_from_coeff(C0, N)

    ctx = ctypes.c_void_p()
    _lib.call("tarang_ctx_create", 0, ctypes.byref(ctx))

    def make(dtype_code):
        d = _lib.FivpDesc()
        d.ndim = 2
        d.n[0] = d.n[1] = N
        d.L[0] = d.L[1] = 2 * math.pi
        d.dtype = dtype_code
        d.problem = 1
        d.timestepper = 1  # RK443
        d.nu = NU
        d.dt = DT
        h = ctypes.c_void_p()
        _lib.call("tarang_fivp_create", ctx, ctypes.byref(d), ctypes.byref(h))
        return h

    results = {}
    finals = {}
    solvers = {}
    tuning = {}
    for name, code, ct in (("f64", _lib.F64, torch.complex128), ("f32", _lib.F32, torch.complex64)):
        s = make(code)
        solvers[name] = s
        tv = (ctypes.c_int * 3)()
        _lib.call("tarang_fivp_get_tuning", s, tv)
        tuning[name] = {"ybwd": tv[0], "xpass": tv[1], "yfwd": tv[2]}
        Cd = torch.as_tensor(np.ascontiguousarray(C0.T), device=dev).to(ct).contiguous()
        torch.cuda.synchronize()
        _lib.call("tarang_fivp_set_state", s, ctypes.c_void_p(Cd.data_ptr()))
        log(f"{name}: warmup {a.warmup} steps")
        _lib.call("tarang_fivp_step", s, a.warmup)
        _lib.call("tarang_fivp_set_state", s, ctypes.c_void_p(Cd.data_ptr()))  # timed run starts from the IC
        _lib.call("tarang_ctx_sync", ctx)
        torch.cuda.synchronize()
        log(f"{name}: timing {a.steps} steps")
        t0 = time.perf_counter()
        _lib.call("tarang_fivp_step", s, a.steps)
        _lib.call("tarang_ctx_sync", ctx)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / a.steps
        results[name] = ms
        out = torch.zeros_like(Cd)
        _lib.call("tarang_fivp_get_state", s, ctypes.c_void_p(out.data_ptr()))
        _lib.call("tarang_ctx_sync", ctx)
        finals[name] = out.cpu().numpy().T.astype(np.complex128)
        log(f"{name}: {ms:.4f} ms/step, {N * N / (ms * 1e-3):.4e} grid-point steps/s")

    Ef = {k: configs.energy_enstrophy_from_coeff(v, N) for k, v in finals.items()}
    err_state = float(np.abs(finals["f32"] - finals["f64"]).max() / np.abs(finals["f64"]).max())
    err_E = abs(Ef["f32"][0] - Ef["f64"][0]) / Ef["f64"][0]
    err_Z = abs(Ef["f32"][1] - Ef["f64"][1]) / Ef["f64"][1]

    # live per-kernel durations (HIP events on the library stream) for the FP32 path
    log(f"profiling {a.profile_steps} eager steps with HIP events")
    avg = (ctypes.c_double * 4)()
    cnt = (ctypes.c_long * 4)()
    s32 = solvers["f32"]
    _lib.call("tarang_fivp_profile", s32, a.profile_steps, avg, cnt)
    yb, xp, yf = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _lib.call("tarang_fivp_kernel_bytes", s32, ctypes.byref(yb), ctypes.byref(xp), ctypes.byref(yf))
    nbytes = [yb.value, xp.value, yf.value, None]
    per_step = [avg[k] * cnt[k] / a.profile_steps for k in range(4)]
    dom = max(range(3), key=lambda k: per_step[k])
    achieved = nbytes[dom] / (avg[dom] * 1e-3) / 1e9
    kernels = {KERNELS[k]: {"avg_ms": avg[k], "launches_per_step": cnt[k] / a.profile_steps,
                            "algorithmic_bytes": nbytes[k],
                            "GBps": (nbytes[k] / (avg[k] * 1e-3) / 1e9) if nbytes[k] and avg[k] else None}
               for k in range(4)}
    # the timed FP32 variant's template arguments: y kernels (tile width, workgroup size) by
    # variant 0 (4, 256), 1 (2, 256), 2 (8, 512), 3 (4, 512); the x pass 256 or 512 threads
    tv = tuning.get("f32", {})
    yv = [(4, 256), (2, 256), (8, 512), (4, 512)]
    M = 3 * N // 2
    piece = [f"{N}, {M}, {yv[tv.get('ybwd', 0)][0]}, {yv[tv.get('ybwd', 0)][1]}",
             f"{N}, {M}, {(256, 512)[tv.get('xpass', 0)]}",
             f"{N}, {M}, {yv[tv.get('yfwd', 0)][0]}, {yv[tv.get('yfwd', 0)][1]}"][dom]
    traffic, traffic_kernel = measured_traffic("traffic_turbulence_1024_f32.json", KERNELS[dom], "float", piece)

    for s in solvers.values():
        _lib.call("tarang_fivp_destroy", s)
    _lib.call("tarang_ctx_destroy", ctx)

    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        threads = host_threads()
        O.set_threads(threads)
        d = O.domain([0, 0], [N, N])
        p = O.FourierIVP(d, O.PROB_TURB2D, "RK443", NU, DT)
        p.set_coeff(C0)
        log(f"cpu baseline: oracle, {a.cpu_steps} RK443 steps on {threads} threads")
        t0 = time.perf_counter()
        p.step(a.cpu_steps)
        el = time.perf_counter() - t0
        cpu = {"value": N * N * a.cpu_steps / el, "unit": "grid-point steps/s", "cores": threads, "host": host_info(), "kind": "port",
               "sample": f"{a.cpu_steps} RK443 steps of the same 1024^2 workload from the same initial condition, "
                         f"FP64 oracle (C, OpenMP over transform lines), {el:.2f} s"}

    v32 = N * N / (results["f32"] * 1e-3)
    v64 = N * N / (results["f64"] * 1e-3)
    line = {
        "metric": "grid-point steps/s, FP32+refine vs FP64 at <=1e-6 rel err; p50/p99 latency",
        "value": v32,
        "unit": "grid-point steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": results["f32"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-phase vorticity, E(k)~k^4 exp(-2(k/10)^2), E=0.5)",
        "config": {"workload": WORKLOAD, "grid": [N, N], "timestepper": "RK443", "dt": DT, "nu": NU,
                   "parallelism": "single GPU"},
        "fp64": {"value": v64, "ms_per_step": results["f64"]},
        "speedup_fp32_vs_fp64": v32 / v64,
        "fp32_vs_fp64_rel_err": {"state_maxnorm": err_state, "energy": err_E, "enstrophy": err_Z,
                                 "after_steps": a.steps},
        "diagnostics": {"E0": E0, "Z0": Z0, "E_final_f64": Ef["f64"][0], "Z_final_f64": Ef["f64"][1]},
        "roofline": {"bound": "hbm", "kernel": KERNELS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_kernel": traffic_kernel},
        "kernels_f32": kernels,
        "autotuned_variants": tuning,
        "cpu_baseline": cpu,
    }
    return line


CH = dict(nx=1024, nz=256, Lx=4.0, Lz=1.0, Ra=2e6, Pr=1.0, dt=0.01)
CH_WORKLOAD = ("2-D Rayleigh-Benard, RealFourier x ChebyshevT 1024x256, Lx=4, Lz=1, Ra=2e6, Pr=1, first-order tau "
               "form (two lifts per bounded variable), 3/2 dealiasing, RK222, dt=0.01")
CH_KERNELS = ["k_ch_zbwd (Chebyshev evaluation of 8 operand planes)",
              "k_ch_xpass (inverse x FFTs M=1536, products, forward x FFTs)",
              "k_ch_zfwd (Chebyshev forward of 3 products)",
              "k_ch_lx (structured L x per mode)",
              "k_ch_solve (per-mode explicit inverses, FP64 GEMV, parity-split systems)",
              "stage / copy / zero"]


def channel_ic(np, nx, nz, Lz, seed=42):
    """b0 = Lz - z + 1e-3 N(0,1) z (Lz - z), u = 0 (grid values, x fastest)."""
    z = Lz * (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
    rng = np.random.default_rng(seed)
    return (Lz - z)[None, :] + 1e-3 * rng.standard_normal((nx, nz)) * (z * (Lz - z))[None, :], z


def dt_change_cost(_lib, ctx, h, pre, dt, steady_ms):
    """Wall time of the two steps right after a dt change, against a steady step: the
    per-mode systems are refactored (once for RK; for variable-step SBDF2 once for the
    step with the new ratio and once when the ratio returns to 1) and the step graph is
    captured again."""
    out = {}
    for name, new_dt in (("dt x 0.9", dt * 0.9), ("back to dt", dt)):
        _lib.call(pre + "set_dt", h, new_dt)
        _lib.call("tarang_ctx_sync", ctx)
        ts = [time.perf_counter()]
        for _ in range(3):
            _lib.call(pre + "step", h, 1)
            _lib.call("tarang_ctx_sync", ctx)
            ts.append(time.perf_counter())
        out[name] = [round((b - a) * 1e3, 3) for a, b in zip(ts, ts[1:])]
    return {"ms_of_the_3_steps_after_a_change": out, "steady_ms_per_step": steady_ms}


# FP32 forms of the channel solvers: (name, refine_sweeps, split_inverse); the first is the
# headline form.  split_inverse stores every tau-system inverse entry as an FP32 value plus a
# 8-bit remainder (2^-32 relative); refine_sweeps adds FP64-residual refinement sweeps.
FP32_FORMS = (("f32", 0, 1), ("f32_plain", 0, 0), ("f32_sweeps1", 1, 0), ("f32_sweeps2", 2, 0))
FP32_HEADLINE = {"refine_sweeps": FP32_FORMS[0][1], "split_inverse": FP32_FORMS[0][2]}


def fp32_form_table(np, results, finals, nus):
    """Per FP32 form: time, speed against FP64 and its FP32-vs-FP64 errors."""
    table = {}
    for name, sw, sp in FP32_FORMS:
        table[name] = {
            "refine_sweeps": sw, "split_inverse": sp,
            "ms_per_step": results[name][0], "speedup_vs_fp64": results["f64"][0] / results[name][0],
            "nusselt_rel_err_vs_fp64": abs(nus[name] - nus["f64"]) / abs(nus["f64"]),
            "b_rel_err_vs_fp64": float(np.abs(finals[name][0] - finals["f64"][0]).max() /
                                       np.abs(finals["f64"][0]).max())}
    return table


def bench_channel(a):
    import numpy as np
    import torch

    from tarang_amd import _lib, diagnostics

    nx, nz, Lx, Lz, Ra, Pr, dt = (CH[k] for k in ("nx", "nz", "Lx", "Lz", "Ra", "Pr", "dt"))
    kappa, nu = (Ra * Pr) ** -0.5, (Ra / Pr) ** -0.5
    steps = a.steps if a.steps != 1000 else 200
    # the configuration's 2000 steps at the fixed dt = 0.01 leave the advective stability limit
    # once convection develops (both forms diverge before step 2000), so the FP32-vs-FP64
    # comparison defaults to the timed steps (t = 2)
    horizon = a.horizon if a.horizon is not None else steps
    dev = torch.device("cuda", 0)
    bg, z = channel_ic(np, nx, nz, Lz)
    # coefficients: rfft in x, Chebyshev interpolation in z (exact on the Gauss-Lobatto grid)
    k = np.arange(nz)
    T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
    Bc = np.linalg.solve(T, np.fft.rfft(bg, axis=0).T).T
    ctx = ctypes.c_void_p()
    _lib.call("tarang_ctx_create", 0, ctypes.byref(ctx))
    results, finals, solvers = {}, {}, {}
    forms = (("f64", _lib.F64, torch.complex128, 0, 0),) + tuple(
        (name, _lib.F32, torch.complex64, sw, sp) for name, sw, sp in FP32_FORMS)
    for name, code, ct, sweeps, split in forms:
        d = _lib.ChanDesc()
        d.nx, d.nz, d.Lx, d.Lz, d.kappa, d.nu, d.dt = nx, nz, Lx, Lz, kappa, nu, dt
        d.dtype, d.timestepper, d.b_bottom, d.b_top = code, 0, Lz, 0.0
        d.refine_sweeps, d.split_inverse = sweeps, split
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        _lib.call("tarang_chan_create", ctx, ctypes.byref(d), ctypes.byref(h))
        tf = time.perf_counter() - t0
        solvers[name] = h
        Bd = torch.as_tensor(np.ascontiguousarray(Bc.T), device=dev).to(ct).contiguous()
        Zd = torch.zeros_like(Bd)
        torch.cuda.synchronize()

        def reset():
            _lib.call("tarang_chan_set_field", h, _lib.CHAN_B, ctypes.c_void_p(Bd.data_ptr()))
            for v in (_lib.CHAN_P, _lib.CHAN_UX, _lib.CHAN_UZ):
                _lib.call("tarang_chan_set_field", h, v, ctypes.c_void_p(Zd.data_ptr()))
        reset()
        log(f"channel {name}: setup (assemble + factor) {tf:.3f} s; warmup {a.warmup} steps")
        _lib.call("tarang_chan_step", h, a.warmup)
        reset()
        _lib.call("tarang_ctx_sync", ctx)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.call("tarang_chan_step", h, steps)
        _lib.call("tarang_ctx_sync", ctx)
        t1 = time.perf_counter()
        results[name] = ((t1 - t0) * 1e3 / steps, tf)
        if horizon > steps:  # the FP32-vs-FP64 comparison runs to the configuration's horizon
            _lib.call("tarang_chan_step", h, horizon - steps)
        outs = []
        for v in (_lib.CHAN_B, _lib.CHAN_UZ):
            o = torch.zeros_like(Bd)
            _lib.call("tarang_chan_get_field", h, v, ctypes.c_void_p(o.data_ptr()))
            outs.append(o)
        _lib.call("tarang_ctx_sync", ctx)
        finals[name] = [o.cpu().numpy().T.astype(np.complex128) for o in outs]
        log(f"channel {name}: {results[name][0]:.4f} ms/step")
    nus = {k: diagnostics.nusselt(v[0], v[1], nx, Lz, kappa, Lz) for k, v in finals.items()}
    err_b = float(np.abs(finals["f32"][0] - finals["f64"][0]).max() / np.abs(finals["f64"][0]).max())
    by_form = fp32_form_table(np, results, finals, nus)
    err_w = float(np.abs(finals["f32"][1] - finals["f64"][1]).max() / max(np.abs(finals["f64"][1]).max(), 1e-300))
    # live per-kernel durations, FP32 path
    avg = (ctypes.c_double * 6)()
    cnt = (ctypes.c_long * 6)()
    _lib.call("tarang_chan_profile", solvers["f32"], a.profile_steps, avg, cnt)
    nb = (ctypes.c_double * 6)()
    _lib.call("tarang_chan_kernel_bytes", solvers["f32"], nb)
    per_step = [avg[i] * cnt[i] / a.profile_steps for i in range(6)]
    dom = max((0, 1, 2, 3, 5), key=lambda i: per_step[i])  # HBM class; the solve (4) is reported on its own
    tr_bytes, tr_name = measured_traffic("traffic_channel2d_1024x256_f32.json", CH_KERNELS[dom])
    solve_class = {"kernel": CH_KERNELS[4], "ms_per_step": per_step[4], "bytes_per_launch": nb[4],
                   "GBps": nb[4] / (avg[4] * 1e-3) / 1e9 if avg[4] else None,
                   "share_of_step": per_step[4] / sum(per_step) if sum(per_step) else None}
    achieved = nb[dom] / (avg[dom] * 1e-3) / 1e9
    kernels = {CH_KERNELS[i]: {"avg_ms": avg[i], "launches_per_step": cnt[i] / a.profile_steps,
                               "algorithmic_bytes": nb[i] or None,
                               "GBps": (nb[i] / (avg[i] * 1e-3) / 1e9) if nb[i] and avg[i] else None}
               for i in range(6)}
    dtc = dt_change_cost(_lib, ctx, solvers["f32"], "tarang_chan_", dt, results["f32"][0])
    for h in solvers.values():
        _lib.call("tarang_chan_destroy", h)
    _lib.call("tarang_ctx_destroy", ctx)
    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        threads = host_threads()
        O.set_threads(threads)
        sx, cs = 32, 2
        ref = O.RBC2D(sx, nz, Lx, Lz, kappa, nu, dt, "RK222")
        bgs, _ = channel_ic(np, sx, nz, Lz)
        ref.set(O.RBC_B, O.forward_transform(ref.d, bgs))
        ref.step(1)  # factorisation happens on the first step
        t0 = time.perf_counter()
        ref.step(cs)
        el = time.perf_counter() - t0
        cpu = {"value": sx * nz * cs / el, "unit": "grid-point steps/s", "cores": threads, "host": host_info(), "kind": "port",
               "sample": f"{cs} RK222 steps at Nx={sx}, Nz={nz} (17 of the 513 per-mode systems, same size each), "
                         f"FP64 oracle with dense complex per-mode LU, factorisation excluded, {el:.2f} s"}
    v32 = nx * nz / (results["f32"][0] * 1e-3)
    v64 = nx * nz / (results["f64"][0] * 1e-3)
    line = {
        "metric": "grid-point steps/s, FP32+refine vs FP64 at <=1e-6 rel err; p50/p99 latency",
        "value": v32, "unit": "grid-point steps/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": results["f32"][0], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (b0 = Lz - z + 1e-3 N(0,1) z (Lz - z), u = 0, seed 42)",
        "config": {"workload": CH_WORKLOAD, "grid": [nx, nz], "timestepper": "RK222", "dt": dt, "Ra": Ra, "Pr": Pr,
                   "parallelism": "single GPU"},
        "fp64": {"value": v64, "ms_per_step": results["f64"][0], "setup_s": results["f64"][1]},
        "setup_s_f32": results["f32"][1], "dt_change_f32": dtc,
        "speedup_fp32_vs_fp64": v32 / v64,
        "fp32_form": FP32_HEADLINE,
        "fp32_vs_fp64_rel_err": {"b_maxnorm": err_b, "uz_maxnorm": err_w,
                                 "nusselt": abs(nus["f32"] - nus["f64"]) / abs(nus["f64"]),
                                 "after_steps": max(horizon, steps)},
        "fp32_forms": by_form,
        "nusselt": nus,
        "roofline": {"bound": "hbm", "kernel": CH_KERNELS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr_bytes, "traffic_kernel": tr_name},
        "solve_class_f32": solve_class,
        "kernels_f32": kernels,
        "cpu_baseline": cpu,
    }
    if not a.no_group:
        line["group"] = channel_group_bench(a)
    if a.cfl_horizon:
        line["cfl_horizon"] = channel_cfl_horizon(a)
    return line


# a group of uneven 2-D channel problems (nx, nz, Ra): FP32 split inverses, RK222, dt = 1e-3
GROUP_MEMBERS = [(256, 64, 1e6)] * 4 + [(512, 128, 2e6)] * 4 + [(1024, 64, 5e5)] * 2 + [(128, 256, 1e7)] * 2


def channel_group_bench(a):
    """The GROUP_MEMBERS problems stepped together (tarang_chan_group_step: one per-mode solve
    launch per implicit stage for all of them, one graph per group step) against the same
    problems stepped one after another (one graph, with its own solve launches, per problem
    and step), on one context stream."""
    import numpy as np
    import torch

    from tarang_amd import _lib

    dev = torch.device("cuda", 0)
    ctx = ctypes.c_void_p()
    _lib.call("tarang_ctx_create", 0, ctypes.byref(ctx))
    steps = 50
    out = {}
    sets = {}
    for label in ("alone", "group"):
        hs, keep = [], []
        for i, (nx, nz, Ra) in enumerate(GROUP_MEMBERS):
            d = _lib.ChanDesc()
            d.nx, d.nz, d.Lx, d.Lz = nx, nz, 4.0, 1.0
            d.kappa, d.nu, d.dt = Ra ** -0.5, Ra ** -0.5, 1e-3
            d.dtype, d.timestepper, d.b_bottom, d.b_top = _lib.F32, 0, 1.0, 0.0
            d.refine_sweeps, d.split_inverse = 0, 1
            h = ctypes.c_void_p()
            _lib.call("tarang_chan_create", ctx, ctypes.byref(d), ctypes.byref(h))
            bg, _ = channel_ic(np, nx, nz, 1.0, seed=100 + i)
            k = np.arange(nz)
            T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
            Bc = np.linalg.solve(T, np.fft.rfft(bg, axis=0).T).T
            Bd = torch.as_tensor(np.ascontiguousarray(Bc.T), device=dev).to(torch.complex64).contiguous()
            torch.cuda.synchronize()
            _lib.call("tarang_chan_set_field", h, _lib.CHAN_B, ctypes.c_void_p(Bd.data_ptr()))
            keep.append(Bd)
            hs.append(h)
        sets[label] = (hs, keep)
    g = ctypes.c_void_p()
    hs = sets["group"][0]
    _lib.call("tarang_chan_group_create", (ctypes.c_void_p * len(hs))(*[h.value for h in hs]), len(hs),
              ctypes.byref(g))
    points = sum(nx * nz for nx, nz, _ in GROUP_MEMBERS)
    for label in ("alone", "group"):
        hs = sets[label][0]

        def run(n):
            if label == "group":
                _lib.call("tarang_chan_group_step", g, n)
            else:
                for _ in range(n):
                    for h in hs:
                        _lib.call("tarang_chan_step", h, 1)
        run(5)
        _lib.call("tarang_ctx_sync", ctx)
        t0 = time.perf_counter()
        run(steps)
        _lib.call("tarang_ctx_sync", ctx)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out[label] = {"ms_per_step": ms, "grid_point_steps_per_s": points / (ms * 1e-3)}
        log(f"channel group of {len(GROUP_MEMBERS)} problems, {label}: {ms:.3f} ms per step of the set")
    # the two sets took the same steps from the same states: every member's state agrees
    same = True
    for i, (nx, nz, _) in enumerate(GROUP_MEMBERS):
        o = [torch.zeros((nz, nx // 2 + 1), dtype=torch.complex64, device=dev) for _ in range(2)]
        for j, label in enumerate(("alone", "group")):
            _lib.call("tarang_chan_get_field", sets[label][0][i], _lib.CHAN_B, ctypes.c_void_p(o[j].data_ptr()))
        _lib.call("tarang_ctx_sync", ctx)
        same = same and bool(torch.equal(o[0], o[1]))
    _lib.call("tarang_chan_group_destroy", g)
    for label in ("alone", "group"):
        for h in sets[label][0]:
            _lib.call("tarang_chan_destroy", h)
    _lib.call("tarang_ctx_destroy", ctx)
    return {"members": [{"nx": nx, "nz": nz, "Ra": Ra} for nx, nz, Ra in GROUP_MEMBERS],
            "form": "FP32 split inverses, RK222, dt = 1e-3", "steps": steps, "alone": out["alone"],
            "group": out["group"], "speedup": out["alone"]["ms_per_step"] / out["group"]["ms_per_step"],
            "same_bits": same}


def channel_cfl_horizon(a):
    """The 2-D configuration run to its horizon t = 20 (2000 steps of the fixed dt = 0.01 at
    most) with the CFL controller choosing dt (max_dt = 0.01, safety 0.5, cadence 10), FP64
    and the FP32 split-inverse form, through the API mirror.  Convection at Ra = 2e6 is
    chaotic, so the two trajectories decorrelate: they are compared by the time-averaged
    Nusselt number over t in [10, 20], not pointwise."""
    import numpy as np

    import tarang_amd as T
    from tarang_amd import _lib, flow_tools

    nx, nz, Lx, Lz, Ra, Pr = (CH[k] for k in ("nx", "nz", "Lx", "Lz", "Ra", "Pr"))
    kappa, nu = (Ra * Pr) ** -0.5, (Ra / Pr) ** -0.5
    out = {}
    for name, dtype in (("f64", np.float64), ("f32", np.float32)):
        coords = T.CartesianCoordinates("x", "z")
        dist = T.Distributor(coords, dtype=dtype, device=T.GPU())
        xb = T.RealFourier(coords["x"], size=nx, bounds=(0.0, Lx), dealias=3 / 2)
        zb = T.ChebyshevT(coords["z"], size=nz, bounds=(0.0, Lz), dealias=3 / 2)
        dom = T.Domain(dist, (xb, zb))
        p, b, u = T.ScalarField(dom, "p"), T.ScalarField(dom, "b"), T.VectorField(dom, "u")
        taus = [T.ScalarField(dist, "tau_p", (), dtype), T.ScalarField(dist, "tau_b1", (xb,), dtype),
                T.ScalarField(dist, "tau_b2", (xb,), dtype), T.VectorField(dist, coords, "tau_u1", (xb,), dtype),
                T.VectorField(dist, coords, "tau_u2", (xb,), dtype)]
        ex, ez = T.unit_vector_fields(coords, dist)
        lift = lambda A: T.lift(A, T.derivative_basis(zb, 1), -1)  # noqa: E731
        pb = T.IVP([p, b, u] + taus)
        T.add_parameters_(pb, kappa=kappa, nu=nu, Lz=Lz, ez=ez, τ_lift=lift,
                          grad_u=T.grad(u) + ez * lift(taus[3]), grad_b=T.grad(b) + ez * lift(taus[1]))
        T.add_equation_(pb, "trace(grad_u) + tau_p = 0")
        T.add_equation_(pb, "∂t(b) - kappa*div(grad_b) + τ_lift(tau_b2) = -u⋅∇(b)")
        T.add_equation_(pb, "∂t(u) - nu*div(grad_u) + ∇(p) - b*ez + τ_lift(tau_u2) = -u⋅∇(u)")
        for bc in ("b(z=0) = Lz", "u(z=0) = 0", "b(z=Lz) = 0", "u(z=Lz) = 0", "integ(p) = 0"):
            T.add_bc_(pb, bc)
        solver = T.InitialValueSolver(pb, T.RK222(), dt=CH["dt"])
        bg, _ = channel_ic(np, nx, nz, Lz)
        b["g"] = bg
        cfl = flow_tools.CFL(solver, initial_dt=CH["dt"], cadence=10, safety=0.5, max_dt=CH["dt"])
        flow_tools.add_velocity_(cfl, u)
        nus, dts = [], set()

        t0 = time.perf_counter()
        last = [t0]

        def sample(sv):
            v = ctypes.c_double()
            _lib.call("tarang_chan_nusselt", sv.handle, ctypes.byref(v))
            nus.append((sv.sim_time, v.value))
            dts.add(sv.dt)
            now = time.perf_counter()
            if now - last[0] >= 15.0:
                last[0] = now
                log(f"channel cfl horizon {name}: step {sv.iteration}, t = {sv.sim_time:.4f}, dt = {sv.dt:.4g}, "
                    f"Nu = {v.value:.6g}, {now - t0:.0f} s")
        # the wall-clock guard ends a run that the CFL dt would make longer than the bench allows
        T.run_(solver, stop_time=20.0, stop_wall_time=a.cfl_wall, cfl=cfl, callbacks=[(50, sample)])
        el = time.perf_counter() - t0
        late = [v for t, v in nus if t >= 10.0]
        out[name] = {"steps": solver.iteration, "sim_time": solver.sim_time, "wall_s": el,
                     "reached_horizon": bool(solver.sim_time >= 20.0 - 1e-9),
                     "dt_min": min(dts) if dts else None, "dt_max": max(dts) if dts else None,
                     "distinct_dt": len(dts), "finite": bool(np.all(np.isfinite(np.asarray([v for _, v in nus])))),
                     "nusselt_final": nus[-1][1] if nus else None,
                     "nusselt_mean_t10_20": float(np.mean(late)) if late else None,
                     "nusselt_std_t10_20": float(np.std(late)) if late else None}
        log(f"channel cfl horizon {name}: {solver.iteration} steps to t = {solver.sim_time:.3f} in {el:.1f} s, "
            f"<Nu>(10..20) = {out[name]['nusselt_mean_t10_20']}")
        solver.close()
    return out


CH3 = dict(nx=256, ny=256, nz=128, Lx=4.0, Ly=4.0, Lz=1.0, Ra=1e7, Pr=0.7, dt=2e-3)
CH3_WORKLOAD = ("3-D Rayleigh-Benard, RealFourier^2 x ChebyshevT 256x256x128, Lx=Ly=4, Lz=1, Ra=1e7, Pr=0.7, "
                "no-slip walls, b fixed at both walls, 3/2 dealiasing, SBDF2, dt=2e-3")
CH3_KERNELS = ["k3_zbwd (Chebyshev evaluation of 15 operand planes)", "k3_ybwd (padded inverse y FFT, MY=384)",
               "k3_xpass (inverse x FFTs MX=384, 12 products, forward x FFTs)",
               "k3_yfwd (forward y FFT, truncation, Nyquist folds)", "k_ch_zfwd (Chebyshev forward of 4 products)",
               "k3_sbdf (multistep right-hand side)",
               "k3_solve (explicit inverses per k_h^2 key, 8-mode chunks on FP64 matrix cores)",
               "k3_lx (structured L x, RK only)"]


def bench_channel3d(a):
    import numpy as np
    import torch

    from tarang_amd import _lib, diagnostics

    c = CH3
    nx, ny, nz, Lz = c["nx"], c["ny"], c["nz"], c["Lz"]
    kappa, nu = (c["Ra"] * c["Pr"]) ** -0.5, (c["Ra"] / c["Pr"]) ** -0.5
    steps = a.steps if a.steps != 1000 else 100
    horizon = a.horizon if a.horizon is not None else 500  # SBDF2, dt = 2e-3, 500 steps
    dev = torch.device("cuda", 0)
    z = Lz * (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
    rng = np.random.default_rng(42)
    bg = (Lz - z)[None, None, :] + 1e-3 * rng.standard_normal((nx, ny, nz)) * (z * (Lz - z))[None, None, :]
    k = np.arange(nz)
    T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
    F = np.fft.fft(np.fft.rfft(bg, axis=0), axis=1)
    Bc = np.einsum("kz,mnz->mnk", np.linalg.inv(T), F)
    ctx = ctypes.c_void_p()
    _lib.call("tarang_ctx_create", 0, ctypes.byref(ctx))
    results, finals, solvers, keys, setups, tunings = {}, {}, {}, {}, {}, {}
    # FP64, then the FP32 forms (the headline form first)
    forms = (("f64", _lib.F64, torch.complex128, 0, 0),) + tuple(
        (name, _lib.F32, torch.complex64, sw, sp) for name, sw, sp in FP32_FORMS)
    for name, code, ct, sweeps, split in forms:
        d = _lib.Chan3Desc()
        d.nx, d.ny, d.nz, d.Lx, d.Ly, d.Lz = nx, ny, nz, c["Lx"], c["Ly"], Lz
        d.kappa, d.nu, d.dt, d.dtype, d.timestepper, d.b_bottom, d.b_top = kappa, nu, c["dt"], code, 2, Lz, 0.0
        d.refine_sweeps, d.split_inverse = sweeps, split
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        _lib.call("tarang_chan3_create", ctx, ctypes.byref(d), ctypes.byref(h))
        tf = time.perf_counter() - t0
        solvers[name] = h
        kk = ctypes.c_int()
        _lib.call("tarang_chan3_num_keys", h, ctypes.byref(kk))
        keys[name] = kk.value
        tv = (ctypes.c_int * 2)()
        _lib.call("tarang_chan3_get_tuning", h, tv)
        tunings[name] = {"solve_variant": tv[0], "zbwd_16_mode_tile": tv[1]}
        Bd = torch.as_tensor(np.ascontiguousarray(Bc.T), device=dev).to(ct).contiguous()
        Zd = torch.zeros_like(Bd)
        torch.cuda.synchronize()

        def reset():
            _lib.call("tarang_chan3_set_field", h, _lib.CHAN3_B, ctypes.c_void_p(Bd.data_ptr()))
            for v in (_lib.CHAN3_P, _lib.CHAN3_UX, _lib.CHAN3_UY, _lib.CHAN3_UZ):
                _lib.call("tarang_chan3_set_field", h, v, ctypes.c_void_p(Zd.data_ptr()))
        reset()
        # the warmup covers the SBDF1 start and the first SBDF2 step, each with its own
        # factorisation; the timed steps are steady SBDF2 steps (one graph replay each)
        warm = max(a.warmup, 3)
        log(f"channel3d {name}: setup {tf:.3f} s ({kk.value} |k_h|^2 keys); warmup {warm} steps")
        t0 = time.perf_counter()
        _lib.call("tarang_chan3_step", h, warm)
        _lib.call("tarang_ctx_sync", ctx)
        tw = time.perf_counter() - t0
        setups[name] = tw
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.call("tarang_chan3_step", h, steps)
        _lib.call("tarang_ctx_sync", ctx)
        t1 = time.perf_counter()
        results[name] = ((t1 - t0) * 1e3 / steps, tf)
        if horizon > warm + steps:  # the FP32-vs-FP64 comparison runs to the configuration's horizon
            _lib.call("tarang_chan3_step", h, horizon - warm - steps)
        outs = []
        for v in (_lib.CHAN3_B, _lib.CHAN3_UZ):
            o = torch.zeros_like(Bd)
            _lib.call("tarang_chan3_get_field", h, v, ctypes.c_void_p(o.data_ptr()))
            outs.append(o)
        _lib.call("tarang_ctx_sync", ctx)
        finals[name] = [o.cpu().numpy().T.astype(np.complex128) for o in outs]
        log(f"channel3d {name}: {results[name][0]:.4f} ms/step")
    nus = {kx: diagnostics.nusselt3(v[0], v[1], nx, ny, Lz, kappa, Lz) for kx, v in finals.items()}
    err_b = float(np.abs(finals["f32"][0] - finals["f64"][0]).max() / np.abs(finals["f64"][0]).max())
    by_form = fp32_form_table(np, results, finals, nus)
    avg = (ctypes.c_double * 8)()
    cnt = (ctypes.c_long * 8)()
    _lib.call("tarang_chan3_profile", solvers["f32"], a.profile_steps, avg, cnt)
    nb = (ctypes.c_double * 8)()
    _lib.call("tarang_chan3_kernel_bytes", solvers["f32"], nb)
    per_step = [avg[i] * cnt[i] / a.profile_steps for i in range(8)]
    # the roofline names the dominant HBM-class kernel (transforms, products, stage sweeps);
    # the solve class (band factors and sweeps) is reported on its own
    dom = max(range(6), key=lambda i: per_step[i])
    achieved = nb[dom] / (avg[dom] * 1e-3) / 1e9
    tr_bytes, tr_name = measured_traffic("traffic_channel3d_256x256x128_f32.json", CH3_KERNELS[dom])
    solve_class = {"kernel": CH3_KERNELS[6], "ms_per_step": per_step[6], "bytes_per_launch": nb[6],
                   "GBps": nb[6] / (avg[6] * 1e-3) / 1e9 if avg[6] else None,
                   "share_of_step": per_step[6] / sum(per_step) if sum(per_step) else None}
    kernels = {CH3_KERNELS[i]: {"avg_ms": avg[i], "launches_per_step": cnt[i] / a.profile_steps,
                                "algorithmic_bytes": nb[i], "GBps": nb[i] / (avg[i] * 1e-3) / 1e9 if avg[i] else None}
               for i in range(8)}
    dtc = dt_change_cost(_lib, ctx, solvers["f32"], "tarang_chan3_", c["dt"], results["f32"][0])
    for h in solvers.values():
        _lib.call("tarang_chan3_destroy", h)
    _lib.call("tarang_ctx_destroy", ctx)
    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        threads = host_threads()
        O.set_threads(threads)
        sx, cs = 16, 2
        ref = O.RBC3D(sx, sx, nz, c["Lx"], c["Ly"], Lz, kappa, nu, c["dt"], "SBDF2")
        zz = Lz * (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
        ref.set(O.RBC3_B, O.forward_transform(ref.d, np.tile((Lz - zz)[None, None, :], (sx, sx, 1))))
        ref.step(2)  # SBDF1 start and the SBDF2 factorisation
        t0 = time.perf_counter()
        ref.step(cs)
        el = time.perf_counter() - t0
        cpu = {"value": sx * sx * nz * cs / el, "unit": "grid-point steps/s", "cores": threads, "host": host_info(), "kind": "port",
               "sample": f"{cs} SBDF2 steps at 16x16x128 (144 of the 33,024 per-mode systems, same size), FP64 "
                         f"oracle with dense complex per-mode LU, factorisation excluded, {el:.2f} s"}
    pts = nx * ny * nz
    v32, v64 = pts / (results["f32"][0] * 1e-3), pts / (results["f64"][0] * 1e-3)
    line = {
        "metric": "grid-point steps/s, FP32+refine vs FP64 at <=1e-6 rel err; p50/p99 latency",
        "value": v32, "unit": "grid-point steps/s", "n_gpus": 1, "steps": steps, "warmup": max(a.warmup, 3),
        "ms_per_step": results["f32"][0], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (b0 = Lz - z + 1e-3 N(0,1) z (Lz - z), u = 0, seed 42)",
        "config": {"workload": CH3_WORKLOAD, "grid": [nx, ny, nz], "timestepper": "SBDF2", "dt": c["dt"],
                   "parallelism": "single GPU"},
        "fp64": {"value": v64, "ms_per_step": results["f64"][0], "setup_s": results["f64"][1],
                 "warmup_s": setups["f64"]},
        "setup_s_f32": results["f32"][1], "warmup_s_f32": setups["f32"], "kh2_keys": keys["f32"],
        "autotuned_variants": tunings,
        "dt_change_f32": dtc,
        "speedup_fp32_vs_fp64": v32 / v64,
        "fp32_form": FP32_HEADLINE,
        "fp32_vs_fp64_rel_err": {"b_maxnorm": err_b, "nusselt": abs(nus["f32"] - nus["f64"]) / abs(nus["f64"]),
                                 "after_steps": max(horizon, steps + max(a.warmup, 3))},
        "fp32_forms": by_form,
        "nusselt": nus,
        "roofline": {"bound": "hbm", "kernel": CH3_KERNELS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr_bytes, "traffic_kernel": tr_name},
        "solve_class_f32": solve_class,
        "kernels_f32": kernels,
        "cpu_baseline": cpu,
    }
    return line


SERVE_WORKLOAD = ("serving stream: Poisson arrivals at 80% of the measured saturation capacity; 40% 2-D turbulence "
                  "N in {128,256,512,1024}^2 (RK443, dt=1e-3), 40% 2-D Rayleigh-Benard Nx x Nz in {128x32, 256x64, "
                  "512x128, 1024x256}, Ra log-uniform in [1e5, 1e8], Pr=1 (RK222, dt=1e-3), 20% 1-D Burgers N in "
                  "{64..4096}, nu=1e-2 (RK222, dt=1e-3); FP32 state, FP64 tau factorisation, inverses stored as FP32 "
                  "values plus 8-bit remainders; steps log-uniform in [200, 2000]")


def serving_stream(np, torch, _lib, configs, n, seed, dev):
    """Seeded request list: (kind, desc, nsteps, ic tensors, output shapes, grid points)."""
    rng = np.random.default_rng(seed)
    reqs = []
    cdt = torch.complex64

    def to_dev(a):
        return torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), device=dev).to(cdt).contiguous()
    for i in range(n):
        u = rng.uniform()
        steps = int(np.exp(rng.uniform(np.log(200), np.log(2000))))
        if u < 0.4:
            N = int(rng.choice([128, 256, 512, 1024]))
            d = _lib.FivpDesc()
            d.ndim, d.n[0], d.n[1], d.L[0], d.L[1] = 2, N, N, 2 * math.pi, 2 * math.pi
            d.dtype, d.problem, d.timestepper, d.nu, d.dt = _lib.F32, 1, 1, 1e-4, 1e-3
            q = configs.turbulence_2d(N, seed=int(rng.integers(1 << 30)), kpeak=10.0, energy=0.5)
            C = np.fft.fft(np.fft.rfft(q, axis=0), axis=1)
            reqs.append((_lib.SERVE_PERIODIC, d, steps, [to_dev(C)], [(N // 2 + 1, N)], N * N))
        elif u < 0.8:
            nx, nz = [(128, 32), (256, 64), (512, 128), (1024, 256)][int(rng.integers(4))]
            Ra = float(np.exp(rng.uniform(np.log(1e5), np.log(1e8))))
            d = _lib.ChanDesc()
            d.nx, d.nz, d.Lx, d.Lz, d.kappa, d.nu, d.dt = nx, nz, 4.0, 1.0, Ra ** -0.5, Ra ** -0.5, 1e-3
            d.dtype, d.timestepper, d.b_bottom, d.b_top = _lib.F32, 0, 1.0, 0.0
            d.refine_sweeps = FP32_HEADLINE["refine_sweeps"]  # FP32 state, refined tau inverses
            d.split_inverse = FP32_HEADLINE["split_inverse"]
            bg, _ = channel_ic(np, nx, nz, 1.0, seed=int(rng.integers(1 << 30)))
            k = np.arange(nz)
            T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
            Bc = np.linalg.solve(T, np.fft.rfft(bg, axis=0).T).T
            reqs.append((_lib.SERVE_CHANNEL, d, steps, [None, to_dev(Bc), None, None], [(nx // 2 + 1, nz)] * 4,
                         nx * nz))
        else:
            N = int(rng.choice([64, 128, 256, 512, 1024, 2048, 4096]))
            d = _lib.FivpDesc()
            d.ndim, d.n[0], d.n[1], d.L[0], d.L[1] = 1, N, 1, 2 * math.pi, 1.0
            d.dtype, d.problem, d.timestepper, d.nu, d.dt = _lib.F32, 2, 0, 1e-2, 1e-3
            x = 2 * math.pi * np.arange(N) / N
            u0 = np.sin(x) + 0.1 * sum(rng.standard_normal() * np.sin(j * x + rng.uniform(0, 6.3)) for j in (2, 3, 4))
            reqs.append((_lib.SERVE_PERIODIC, d, steps, [to_dev(np.fft.rfft(u0))], [(N // 2 + 1,)], N))
    return reqs


def bench_serving(a):
    import numpy as np
    import torch

    from tarang_amd import _lib, configs
    from tarang_amd.serving import Server, latency_stats

    dev = torch.device("cuda", 0)
    n = a.requests
    reqs = serving_stream(np, torch, _lib, configs, n, 7, dev)
    budget = 200 << 30

    def run(arrivals):
        srv = Server(0, lanes=4, mem_budget=budget, chunk_steps=25)
        ids = []
        for (kind, d, steps, ic, shapes, pts), t in zip(reqs, arrivals):
            o = [torch.zeros(tuple(reversed(sh)), dtype=torch.complex64, device=dev) for sh in shapes]
            ids.append(srv.submit(kind, d, steps, t, ic, o))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        srv.run()
        el = time.perf_counter() - t0
        recs = [srv.record(i) for i in ids]
        srv.close()
        return recs, el
    log(f"serving: saturation pass, {n} requests arriving at t=0")
    recs0, el0 = run([0.0] * n)
    cap = n / el0
    rate = 0.8 * cap
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n)) - 0.0
    log(f"serving: capacity {cap:.2f} req/s, Poisson arrivals at {rate:.2f} req/s")
    recs, el = run(list(arrivals))
    st = latency_stats(recs)
    total = sum(r.grid_point_steps for r in recs if r.status == 2)
    line = {
        "metric": "grid-point steps/s, FP32+refine vs FP64 at <=1e-6 rel err; p50/p99 latency",
        "value": st["grid_point_steps_per_s"], "unit": "grid-point steps/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded initial conditions per request)",
        "config": {"workload": SERVE_WORKLOAD, "requests": n, "lanes": 4, "chunk_steps": 25,
                   "parallelism": "single GPU, continuous batching"},
        "latency_s": {"p50": st["p50_s"], "p99": st["p99_s"], "mean": st["mean_s"]},
        "saturation": {"makespan_s": el0, "capacity_req_per_s": cap,
                       "grid_point_steps_per_s": sum(r.grid_point_steps for r in recs0) / el0},
        "offered_rate_req_per_s": rate, "makespan_s": el, "total_grid_point_steps": total,
        "rejected": st.get("rejected", 0),
    }
    return line


FP64_MFMA_PEAK_TFLOPS = 78.6


def bench_plugins(a):
    """Chebyshev grid derivative and batched LU, timed with HIP events on the
    library's stream (the context is bound to torch's current stream)."""
    import numpy as np
    import torch

    from tarang_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(dev)  # a real stream (the library reads NULL as "its own stream")
    torch.cuda.set_stream(stream)
    ctx = ctypes.c_void_p()
    _lib.call("tarang_ctx_create", 0, ctypes.byref(ctx))
    _lib.call("tarang_ctx_set_stream", ctx, ctypes.c_void_p(stream.cuda_stream))
    reps = a.steps if a.steps != 1000 else 50

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / n

    # ---- Chebyshev grid derivative, 256 x 256 x 128 ----
    nx, ny, nz = 256, 256, 128
    shape = (ctypes.c_int * 3)(nx, ny, nz)
    g = torch.Generator(device=dev).manual_seed(7)
    for tdt, code, esz in ((torch.float64, _lib.F64, 8), (torch.float32, _lib.F32, 4)):
        src = torch.rand((nz, ny, nx), dtype=torch.float64, device=dev, generator=g).to(tdt)  # column-major (nx, ny, nz)
        dst = torch.empty_like(src)
        for axis, n in ((2, nz), (0, nx)):
            def call():
                _lib.call("tarang_cheb_deriv", ctx, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                          3, shape, axis, 1, 2.0, code, 0)
            for _ in range(max(a.warmup, 3)):
                call()
            ms = timed(call, reps)
            pts = nx * ny * nz
            flops = float(n) * pts  # parity split: two (n/2)^2 products per line of n points
            tflops = flops / (ms * 1e-3) / 1e12
            gbs = 2.0 * esz * pts / (ms * 1e-3) / 1e9
            mfma_bound = flops / (2.0 * esz * pts) > FP64_MFMA_PEAK_TFLOPS * 1e3 / HBM_PEAK_GBS
            cpu = None
            if not a.no_cpu:
                from oracle import oracle as O
                O.set_threads(min(os.cpu_count() or 1, 16))
                sample = np.random.default_rng(0).uniform(-1, 1, (nx, 16, nz) if axis == 2 else (nx, 16, nz))
                t0 = time.perf_counter()
                O.cheb_deriv(sample, axis, 1, 2.0)
                el = time.perf_counter() - t0
                cpu = {"value": sample.size / el, "unit": "grid points/s", "cores": O.get_threads(), "kind": "port",
                       "sample": f"first derivative of a {nx}x16x{nz} block along axis {axis} (DCT-I chain, FP64), "
                                 f"{el:.3f} s"}
            print(json.dumps({
                "metric": "Chebyshev grid derivative throughput", "value": pts / (ms * 1e-3),
                "unit": "grid points/s", "n_gpus": 1, "steps": reps, "warmup": max(a.warmup, 3), "ms_per_step": ms,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f64" if code == _lib.F64 else "f32 data, f64 arithmetic", "data": "synthetic uniform(0,1)",
                "config": {"workload": f"tarang_cheb_deriv, order 1, {nx}x{ny}x{nz} along axis {axis} (n={n})"},
                "roofline": ({"bound": "mfma", "kernel": "k_axis_apply", "achieved": tflops,
                              "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": tflops / FP64_MFMA_PEAK_TFLOPS, "hbm_GBps": gbs, "traffic": None}
                             if mfma_bound else
                             {"bound": "hbm", "kernel": "k_axis_apply", "achieved": gbs, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "fp64_TFLOPs": tflops, "traffic": None}),
                "cpu_baseline": cpu}), flush=True)
        del src, dst
    torch.cuda.empty_cache()

    # ---- transform plug-in: RealFourier^2 x ChebyshevT 256 x 256 x 128, forward and backward ----
    for tdt, ct, code, esz in ((torch.float64, torch.complex128, _lib.F64, 8),
                               (torch.float32, torch.complex64, _lib.F32, 4)):
        d = _lib.DomainDesc()
        d.ndim = 3
        for i, (kind, n_, hi) in enumerate(((_lib.REALFOURIER, nx, 4.0), (_lib.REALFOURIER, ny, 4.0),
                                            (_lib.CHEBYSHEVT, nz, 1.0))):
            d.kind[i], d.n[i], d.lo[i], d.hi[i], d.dealias[i] = kind, n_, 0.0, hi, 1.5
        d.dtype = code
        plan = ctypes.c_void_p()
        _lib.call("tarang_plan_create", ctx, ctypes.byref(d), ctypes.byref(plan))
        grid = torch.rand((nz, ny, nx), dtype=torch.float64, device=dev, generator=g).to(tdt)
        coeff = torch.empty((nz, ny, nx // 2 + 1), dtype=ct, device=dev)
        back = torch.empty_like(grid)
        fwd = lambda: _lib.call("tarang_forward_transform", plan, ctypes.c_void_p(grid.data_ptr()),  # noqa: E731
                                ctypes.c_void_p(coeff.data_ptr()))
        bwd = lambda: _lib.call("tarang_backward_transform", plan, ctypes.c_void_p(coeff.data_ptr()),  # noqa: E731
                                ctypes.c_void_p(back.data_ptr()))
        for _ in range(max(a.warmup, 3)):
            fwd()
            bwd()
        torch.cuda.synchronize()
        rt = float(((back - grid).abs().max() / grid.abs().max()).item())
        for name, fn in (("forward", fwd), ("backward", bwd)):
            ms = timed(fn, reps)
            # algorithmic bytes: the grid array once and the coefficient array once
            nbytes = nx * ny * nz * esz + (nx // 2 + 1) * ny * nz * 2 * esz
            gbs = nbytes / (ms * 1e-3) / 1e9
            print(json.dumps({
                "metric": "spectral transform throughput", "value": nx * ny * nz / (ms * 1e-3),
                "unit": "grid points/s", "n_gpus": 1, "steps": reps, "warmup": max(a.warmup, 3), "ms_per_step": ms,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f64" if code == _lib.F64 else "f32", "data": "synthetic uniform(0,1)",
                "config": {"workload": f"tarang_{name}_transform, RealFourier^2 x ChebyshevT {nx}x{ny}x{nz} "
                                       f"(rfft x, FFT y, Chebyshev z on the matrix cores)"},
                "round_trip_rel_err": rt,
                "roofline": {"bound": "hbm", "kernel": f"{name} chain (rows, y FFT, k_cheb_axis)", "achieved": gbs,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None},
                "cpu_baseline": None}), flush=True)
        _lib.call("tarang_plan_destroy", plan)
        del grid, coeff, back
    torch.cuda.empty_cache()

    # ---- ChebyshevT axis: dense matrix (method 1) against one FFT of length 2(n-1) per column
    # (method 2), RealFourier^2 x ChebyshevT 256 x 256 x n at the n whose 2(n-1) the LDS FFT takes
    for nzc in (17, 33, 65, 129, 257, 513, 769):
        for tdt, ct, code, esz in ((torch.float64, torch.complex128, _lib.F64, 8),
                                   (torch.float32, torch.complex64, _lib.F32, 4)):
            d = _lib.DomainDesc()
            d.ndim = 3
            for i, (kind, n_, hi) in enumerate(((_lib.REALFOURIER, nx, 4.0), (_lib.REALFOURIER, ny, 4.0),
                                                (_lib.CHEBYSHEVT, nzc, 1.0))):
                d.kind[i], d.n[i], d.lo[i], d.hi[i], d.dealias[i] = kind, n_, 0.0, hi, 1.5
            d.dtype = code
            plan = ctypes.c_void_p()
            _lib.call("tarang_plan_create", ctx, ctypes.byref(d), ctypes.byref(plan))
            grid = torch.rand((nzc, ny, nx), dtype=torch.float64, device=dev, generator=g).to(tdt)
            coeff = torch.empty((nzc, ny, nx // 2 + 1), dtype=ct, device=dev)
            back = torch.empty_like(grid)
            res = {}
            for method, mname in ((1, "dense"), (2, "fft"), (0, "automatic")):
                _lib.call("tarang_plan_set_cheb_method", plan, method)
                fwd = lambda: _lib.call("tarang_forward_transform", plan, ctypes.c_void_p(grid.data_ptr()),  # noqa
                                        ctypes.c_void_p(coeff.data_ptr()))
                bwd = lambda: _lib.call("tarang_backward_transform", plan, ctypes.c_void_p(coeff.data_ptr()),  # noqa
                                        ctypes.c_void_p(back.data_ptr()))
                for _ in range(max(a.warmup, 3)):
                    fwd()
                    bwd()
                torch.cuda.synchronize()
                # alternate the two directions' timing rounds so that drift falls on both
                tf, tb = [], []
                for _ in range(3):
                    tf.append(timed(fwd, reps))
                    tb.append(timed(bwd, reps))
                res[mname] = {"forward_ms": sorted(tf)[1], "backward_ms": sorted(tb)[1],
                              "round_trip_rel_err": float(((back - grid).abs().max() / grid.abs().max()).item())}
            nbytes = nx * ny * nzc * esz + (nx // 2 + 1) * ny * nzc * 2 * esz
            print(json.dumps({
                "metric": "ChebyshevT transform forms",
                "value": nx * ny * nzc / (res["automatic"]["forward_ms"] * 1e-3),
                "unit": "grid points/s (forward, automatic form)", "n_gpus": 1, "steps": reps,
                "warmup": max(a.warmup, 3), "ms_per_step": res["automatic"]["forward_ms"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f64" if code == _lib.F64 else "f32",
                "data": "synthetic uniform(0,1)",
                "config": {"workload": f"tarang_forward/backward_transform, RealFourier^2 x ChebyshevT "
                                       f"{nx}x{ny}x{nzc}: dense Chebyshev matrix vs FFT of length {2 * (nzc - 1)}"},
                "forms": res, "algorithmic_bytes_per_transform": nbytes,
                "roofline": None, "cpu_baseline": None}), flush=True)
            _lib.call("tarang_plan_destroy", plan)
            del grid, coeff, back
    torch.cuda.empty_cache()

    # ---- batched complex LU, 2-D channel stage matrices: 513 modes, n = 1031 ----
    n, nm = 1031, 513
    log(f"batched LU: {nm} modes of n={n} ({nm * n * n * 16 / 1e9:.2f} GB per copy)")
    A0 = torch.randn((nm, n, n), dtype=torch.complex128, device=dev, generator=g)  # storage of (n, n, nm)
    A0.diagonal(dim1=1, dim2=2).add_(2.0 * math.sqrt(n))
    A = torch.empty_like(A0)
    piv = torch.empty((nm, n), dtype=torch.int32, device=dev)
    info = np.zeros(nm, np.int32)
    B = torch.randn((nm, n), dtype=torch.complex128, device=dev, generator=g)
    X = torch.empty_like(B)

    def factor():
        _lib.call("tarang_batched_lu_factor", ctx, ctypes.c_void_p(A.data_ptr()), n, nm,
                  ctypes.c_void_p(piv.data_ptr()), info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))

    def solve():
        _lib.call("tarang_batched_lu_solve", ctx, ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(piv.data_ptr()),
                  n, nm, ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(X.data_ptr()))

    fms = []
    for _ in range(3):  # assemble (copy) then factor, never factor twice
        A.copy_(A0)
        fms.append(timed(factor, 1))
    log(f"batched LU factor: {fms} ms")
    for _ in range(3):
        solve()
    sms = timed(solve, max(reps // 5, 5))
    # residual of a few modes: max |A x - b| / (|A| |x|)
    idx = torch.arange(0, nm, max(nm // 8, 1), device=dev)
    Ax = torch.matmul(A0[idx].transpose(1, 2).contiguous(), X[idx].unsqueeze(2)).squeeze(2)
    res = float(((Ax - B[idx]).abs().max() / (A0[idx].abs().max() * X[idx].abs().max())).item())
    flop_factor = nm * 8.0 / 3.0 * n ** 3
    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        O.set_threads(min(os.cpu_count() or 1, 16))
        ms_cpu = min(16, nm)
        Ah = A0[:ms_cpu].cpu().numpy().transpose(2, 1, 0)  # logical (n, n, modes), column-major
        t0 = time.perf_counter()
        O.batched_lu_factor(np.asfortranarray(Ah))
        el = time.perf_counter() - t0
        cpu = {"value": ms_cpu / el, "unit": "modes factored/s", "cores": O.get_threads(), "kind": "port",
               "sample": f"{ms_cpu} of the {nm} modes (n={n}), unblocked zgetf2 restatement, {el:.2f} s"}
    fbest = min(fms)
    print(json.dumps({
        "metric": "batched complex LU factor + solve", "value": nm / (fbest * 1e-3), "unit": "modes factored/s",
        "n_gpus": 1, "steps": 3, "warmup": 0, "ms_per_step": fbest, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic N(0,1) complex + 2 sqrt(n) I",
        "config": {"workload": f"tarang_batched_lu_factor/solve, {nm} modes, n={n} (2-D channel 1024x256 stage "
                               f"matrices' size)"},
        "factor_ms": fms, "factor_TFLOPs": flop_factor / (fbest * 1e-3) / 1e12,
        "solve_ms": sms, "solve_GBps": nm * n * n * 16 / (sms * 1e-3) / 1e9, "residual": res,
        "roofline": {"bound": "hbm", "kernel": "k_zgetrs", "achieved": nm * n * n * 16 / (sms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": nm * n * n * 16 / (sms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None},
        "cpu_baseline": cpu}), flush=True)
    _lib.call("tarang_ctx_destroy", ctx)


if __name__ == "__main__":
    main()
