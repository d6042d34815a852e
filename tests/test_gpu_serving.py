"""Serving loop (tarang_serve_*): continuous batching of mixed requests on one GPU.

Each request's final state must be identical (bit for bit: the same kernels run on the
same inputs) whether it runs alone or in the mix, with overlapping lanes or with a
memory budget that forces requests to wait.
"""
import ctypes

import numpy as np
import pytest
import torch

from tarang_amd import _lib, configs
from tarang_amd.serving import Server, latency_stats
from tests import gpu_helpers as H

pytestmark = pytest.mark.gpu


def periodic_desc(n, problem, stepper, nu, dt, dtype=_lib.F32):
    d = _lib.FivpDesc()
    d.ndim = len(n)
    d.n[0] = n[0]
    d.n[1] = n[1] if len(n) > 1 else 1
    d.L[0] = 2 * np.pi
    d.L[1] = 2 * np.pi if len(n) > 1 else 1.0
    d.dtype, d.problem, d.timestepper, d.nu, d.dt = dtype, problem, stepper, nu, dt
    return d


def chan_desc(nx, nz, Ra, stepper, dt, dtype=_lib.F32):
    d = _lib.ChanDesc()
    d.nx, d.nz, d.Lx, d.Lz = nx, nz, 4.0, 1.0
    d.kappa, d.nu = Ra ** -0.5, Ra ** -0.5
    d.dt, d.dtype, d.timestepper, d.b_bottom, d.b_top = dt, dtype, stepper, 1.0, 0.0
    return d


def chan3_desc(nx, ny, nz, Ra, dt, dtype=_lib.F32):
    d = _lib.Chan3Desc()
    d.nx, d.ny, d.nz, d.Lx, d.Ly, d.Lz = nx, ny, nz, 4.0, 4.0, 1.0
    d.kappa, d.nu = Ra ** -0.5, Ra ** -0.5
    d.dt, d.dtype, d.timestepper, d.b_bottom, d.b_top = dt, dtype, 2, 1.0, 0.0
    return d


def requests(dev):
    """(kind, desc, nsteps, arrival, ic tensors, output shapes) of a small mixed stream."""
    out = []
    cdt = torch.complex64
    q = configs.turbulence_2d(64, seed=3, kpeak=4.0)
    C = np.fft.fft(np.fft.rfft(q, axis=0), axis=1)
    out.append((_lib.SERVE_PERIODIC, periodic_desc([64, 64], 1, 1, 1e-3, 2e-3), 40, 0.0,
                [H.to_dev(C, cdt, dev)], [(33, 64)]))
    x = 2 * np.pi * np.arange(256) / 256
    u = np.fft.rfft(np.sin(x) + 0.1 * np.cos(3 * x))
    out.append((_lib.SERVE_PERIODIC, periodic_desc([256], 2, 0, 1e-2, 1e-3), 100, 0.002,
                [H.to_dev(u, cdt, dev)], [(129,)]))
    rng = np.random.default_rng(5)
    for (nx, nz, Ra, ts, arr) in ((32, 16, 1e5, 0, 0.004), (64, 24, 1e6, 2, 0.006)):
        z = (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
        bg = (1 - z)[None, :] + 1e-3 * rng.standard_normal((nx, nz)) * (z * (1 - z))[None, :]
        k = np.arange(nz)
        T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
        Bc = np.linalg.solve(T, np.fft.rfft(bg, axis=0).T).T
        out.append((_lib.SERVE_CHANNEL, chan_desc(nx, nz, Ra, ts, 5e-3), 20, arr,
                    [None, H.to_dev(Bc, cdt, dev), None, None], [(nx // 2 + 1, nz)] * 4))
    nx = ny = 16
    nz = 9
    z = (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
    bg = (1 - z)[None, None, :] + 1e-3 * rng.standard_normal((nx, ny, nz)) * (z * (1 - z))[None, None, :]
    k = np.arange(nz)
    T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
    F = np.fft.fft(np.fft.rfft(bg, axis=0), axis=1)
    Bc = np.einsum("kz,mnz->mnk", np.linalg.inv(T), F)
    out.append((_lib.SERVE_CHANNEL3, chan3_desc(nx, ny, nz, 1e5, 2e-3), 10, 0.008,
                [None, H.to_dev(Bc, cdt, dev), None, None, None], [(nx // 2 + 1, ny, nz)] * 5))
    return out


def run_server(dev, reqs, lanes, budget, chunk=7):
    srv = Server(0, lanes=lanes, mem_budget=budget, chunk_steps=chunk)
    outs = []
    ids = []
    for kind, desc, nsteps, arr, ic, shapes in reqs:
        o = [torch.zeros(tuple(reversed(s)), dtype=torch.complex64, device=dev) for s in shapes]
        outs.append(o)
        ids.append(srv.submit(kind, desc, nsteps, arr, ic, o))
    torch.cuda.synchronize()
    srv.run()
    recs = [srv.record(i) for i in ids]
    res = [[t.cpu().numpy() for t in o] for o in outs]
    srv.close()
    return res, recs


def test_serving_mix_equals_alone(gpu):
    dev = gpu.torch_device
    reqs = requests(dev)
    mix, recs = run_server(dev, reqs, lanes=3, budget=8 << 30)
    assert all(r.status == 2 for r in recs), [(r.status, r.error) for r in recs]
    # more than one request was resident at once (the work overlapped)
    spans = sorted((r.admitted, r.done) for r in recs)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for i, rq in enumerate(reqs):
        alone, rec = run_server(dev, [rq[:3] + (0.0,) + rq[4:]], lanes=1, budget=8 << 30)
        assert rec[0].status == 2
        for a, b in zip(mix[i], alone[0]):
            assert np.array_equal(a, b), i
    st = latency_stats(recs)
    assert st["n"] == len(reqs) and st["p99_s"] >= st["p50_s"] > 0 and st["grid_point_steps_per_s"] > 0


def test_serving_budget_serialises_and_rejects(gpu):
    dev = gpu.torch_device
    reqs = requests(dev)
    fps = []
    srv = Server(0, lanes=2, mem_budget=1 << 40)
    for kind, desc, nsteps, arr, ic, shapes in reqs:
        o = [torch.zeros(tuple(reversed(s)), dtype=torch.complex64, device=dev) for s in shapes]
        fps.append(srv.record(srv.submit(kind, desc, nsteps, arr, ic, o)).footprint)
    srv.close()
    # the largest request cannot share the GPU with any other one
    mix, recs = run_server(dev, reqs, lanes=2, budget=max(fps) + min(fps) - 1)
    assert all(r.status == 2 for r in recs), [(r.status, r.error) for r in recs]
    order = sorted(recs, key=lambda r: r.admitted)
    used = []
    for r in order:  # at every admission the resident footprints fit the budget
        used = [u for u in used if u[0] > r.admitted] + [(r.done, r.footprint)]
        assert sum(u[1] for u in used) <= max(fps) + min(fps) - 1
    full, _ = run_server(dev, reqs, lanes=2, budget=8 << 30)
    for a, b in zip(mix, full):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    # a request larger than the whole budget is rejected, the rest still complete
    res, recs = run_server(dev, reqs, lanes=2, budget=max(fps) - 1)
    assert any(r.status == -1 for r in recs) and all(r.status in (2, -1) for r in recs)


def test_serving_failures_stay_with_their_request(gpu):
    """A request whose per-mode system is singular (kappa = 0: the b system's lift column
    vanishes, caught by the factorisation's info check) and one whose state turns
    non-finite (a NaN in its initial condition, caught by the check after its first
    chunk) fail with a message and free their memory; every other request finishes with
    the bits it has alone."""
    dev = gpu.torch_device
    reqs = requests(dev)
    bad_chan = chan_desc(32, 16, 1e5, 0, 5e-3)
    bad_chan.kappa = 0.0
    nz = 16
    Bc = np.zeros((17, nz), complex)
    Bc[0, 0] = 1.0
    u = np.fft.rfft(np.sin(2 * np.pi * np.arange(128) / 128))
    u[3] = np.nan
    bad = [(_lib.SERVE_CHANNEL, bad_chan, 20, 0.001, [None, H.to_dev(Bc, torch.complex64, dev), None, None],
            [(17, nz)] * 4),
           (_lib.SERVE_PERIODIC, periodic_desc([128], 2, 0, 1e-2, 1e-3), 30, 0.003,
            [H.to_dev(u, torch.complex64, dev)], [(65,)])]
    mix, recs = run_server(dev, reqs[:2] + bad + reqs[2:], lanes=3, budget=8 << 30)
    good = [r for i, r in enumerate(recs) if i not in (2, 3)]
    assert all(r.status == 2 for r in good), [(r.status, r.error) for r in recs]
    assert recs[2].status == -2 and "singular" in recs[2].error, recs[2]
    assert recs[3].status == -2 and "non-finite" in recs[3].error and recs[3].steps_done == 7, recs[3]
    st = latency_stats(recs)
    assert st["failed"] == 2 and st["n"] == len(reqs)
    keep = mix[:2] + mix[4:]
    for i, rq in enumerate(reqs):
        alone, rec = run_server(dev, [rq[:3] + (0.0,) + rq[4:]], lanes=1, budget=8 << 30)
        assert rec[0].status == 2
        for a, b in zip(keep[i], alone[0]):
            assert np.array_equal(a, b), i


def test_serving_submit_refuses_unsupported_descriptors(gpu):
    dev = gpu.torch_device
    srv = Server(0, lanes=2, mem_budget=1 << 30)
    o = [torch.zeros((12, 7), dtype=torch.complex64, device=dev)]
    ic = [torch.zeros((12, 7), dtype=torch.complex64, device=dev)]
    with pytest.raises(_lib.TarangError, match="no fused 2-D kernels"):
        srv.submit(_lib.SERVE_PERIODIC, periodic_desc([12, 12], 1, 1, 1e-3, 1e-3), 5, 0.0, ic, o)
    bad = chan_desc(32, 600, 1e5, 0, 1e-3)
    with pytest.raises(_lib.TarangError, match="ChebyshevT size"):
        srv.submit(_lib.SERVE_CHANNEL, bad, 5, 0.0, [None, ic[0], None, None], o * 4)
    srv.close()


@pytest.mark.parametrize("which", ["periodic", "channel", "channel3"])
def test_serving_footprint_bounds_the_allocation(gpu, which):
    """The admission footprint is an upper bound of the device memory a solver holds once
    built (it also counts the FP64 factor scratch of set-up), and not a gross overestimate."""
    import gc
    from tarang_amd.serving import footprint
    if which == "periodic":
        kind, desc = _lib.SERVE_PERIODIC, periodic_desc([512, 512], 1, 1, 1e-4, 1e-3)
        make = lambda: gpu.fivp([512, 512], [2 * np.pi] * 2, 1, 1, 1e-4, 1e-3, dtype=np.float32)  # noqa: E731
    elif which == "channel":
        kind, desc = _lib.SERVE_CHANNEL, chan_desc(512, 128, 1e6, 0, 1e-3)
        make = lambda: gpu.chan(512, 128, 4.0, 1.0, 1e-3, 1e-3, 1e-3, stepper=0, dtype=np.float32)  # noqa: E731
    else:
        kind, desc = _lib.SERVE_CHANNEL3, chan3_desc(64, 64, 33, 1e5, 1e-3)
        make = lambda: gpu.chan3(64, 64, 33, 4.0, 4.0, 1.0, 1e-3, 1e-3, 1e-3, dtype=np.float32)  # noqa: E731
    fp = footprint(kind, desc)
    gc.collect()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    s = make()
    gpu.sync()
    free1, _ = torch.cuda.mem_get_info()
    held = free0 - free1
    s.close()
    print(f"{which}: footprint {fp} B, held after set-up {held} B", flush=True)
    assert 0 < held <= fp and fp <= 4 * held + (32 << 20), (held, fp)


def config_size_requests(dev):
    """Requests at the serving stream's own sizes: 2-D turbulence 128^2 and 1024^2, 2-D
    Rayleigh-Benard 128 x 32 (Ra 1e5), 512 x 128 (Ra 3e6) and 1024 x 256 (Ra 1e8), 1-D
    Burgers N = 64 and 4096; FP32, the channel solves with one FP64-residual refinement sweep.
    Fewer steps than the stream (20-60) keep the test short."""
    cdt = torch.complex64
    rng = np.random.default_rng(31)
    out = []
    for N, steps, arr in ((128, 60, 0.0), (1024, 20, 0.001)):
        q = configs.turbulence_2d(N, seed=int(rng.integers(1 << 30)), kpeak=10.0, energy=0.5)
        C = np.fft.fft(np.fft.rfft(q, axis=0), axis=1)
        out.append((_lib.SERVE_PERIODIC, periodic_desc([N, N], 1, 1, 1e-4, 1e-3), steps, arr,
                    [H.to_dev(C, cdt, dev)], [(N // 2 + 1, N)]))
    for (nx, nz, Ra, steps, arr) in ((128, 32, 1e5, 60, 0.002), (512, 128, 3e6, 30, 0.003), (1024, 256, 1e8, 20, 0.004)):
        d = chan_desc(nx, nz, Ra, 0, 1e-3)
        d.split_inverse = 1  # the stream's FP32 form; one request also refines
        d.refine_sweeps = 1 if nx == 512 else 0
        z = (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
        bg = (1 - z)[None, :] + 1e-3 * rng.standard_normal((nx, nz)) * (z * (1 - z))[None, :]
        k = np.arange(nz)
        T = np.cos(np.outer(np.pi * np.arange(nz) / (nz - 1), k)) * ((-1.0) ** k)[None, :]
        Bc = np.linalg.solve(T, np.fft.rfft(bg, axis=0).T).T
        out.append((_lib.SERVE_CHANNEL, d, steps, arr, [None, H.to_dev(Bc, cdt, dev), None, None],
                    [(nx // 2 + 1, nz)] * 4))
    for N, arr in ((64, 0.005), (4096, 0.006)):
        x = 2 * np.pi * np.arange(N) / N
        u0 = np.sin(x) + 0.1 * sum(rng.standard_normal() * np.sin(j * x + rng.uniform(0, 6.3)) for j in (2, 3, 4))
        out.append((_lib.SERVE_PERIODIC, periodic_desc([N], 2, 0, 1e-2, 1e-3), 60, arr,
                    [H.to_dev(np.fft.rfft(u0), cdt, dev)], [(N // 2 + 1,)]))
    return out


def test_serving_config_sizes_mix_equals_alone(gpu):
    dev = gpu.torch_device
    reqs = config_size_requests(dev)
    print(f"serving {len(reqs)} requests at the stream's sizes", flush=True)
    mix, recs = run_server(dev, reqs, lanes=4, budget=64 << 30, chunk=10)
    assert all(r.status == 2 for r in recs), [r.error for r in recs]
    spans = sorted((r.admitted, r.done) for r in recs)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for i, rq in enumerate(reqs):
        alone, rec = run_server(dev, [rq[:3] + (0.0,) + rq[4:]], lanes=1, budget=64 << 30, chunk=10)
        assert rec[0].status == 2
        for a, b in zip(mix[i], alone[0]):
            assert np.array_equal(a, b), i
            assert np.isfinite(a).all(), i
    st = latency_stats(recs)
    print(f"  p50 {st['p50_s']:.3f} s, p99 {st['p99_s']:.3f} s, {st['grid_point_steps_per_s']:.3e} gp-steps/s",
          flush=True)
