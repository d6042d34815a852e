"""Serving loop: continuous batching of IVP solve requests on one GPU.

A thin host wrapper over the native request loop of libtarang_hip.so
(tarang_serve_*).  The reference has no serving layer; each request is one of the
equation families the GPU path supports (periodic diffusion / vorticity / Burgers,
2-D and 3-D Rayleigh-Benard), stated by its library descriptor.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib


@dataclass
class Record:
    arrival: float
    admitted: float
    done: float
    grid_point_steps: float
    footprint: int
    lane: int            # the slot the request occupied
    status: int          # 2 done, -1 rejected (footprint above the budget), -2 failed
    steps_done: int = 0
    error: str = ""      # why the request failed or was rejected

    @property
    def latency(self) -> float:
        """time to completion from arrival"""
        return self.done - self.arrival


def _request(kind: int, desc, nsteps: int = 0, arrival: float = 0.0):
    r = _lib.ServeRequest()
    r.kind = kind
    if kind == _lib.SERVE_PERIODIC:
        r.fivp = desc
    elif kind == _lib.SERVE_CHANNEL:
        r.chan = desc
    else:
        r.chan3 = desc
    r.nsteps = int(nsteps)
    r.arrival = float(arrival)
    return r


def footprint(kind: int, desc) -> int:
    """Device bytes a request of this kind and descriptor holds while resident."""
    b = ctypes.c_size_t()
    _lib.call("tarang_serve_footprint", ctypes.byref(_request(kind, desc)), ctypes.byref(b))
    return b.value


class Server:
    """Continuous-batching request loop (tarang_serve_create/submit/run/record).

    `lanes` is the number of slots: requests resident at once, each on its own stream."""

    def __init__(self, device: int = 0, lanes: int = 4, mem_budget: int = 64 << 30, chunk_steps: int = 10):
        h = ctypes.c_void_p()
        _lib.call("tarang_serve_create", int(device), int(lanes), ctypes.c_size_t(int(mem_budget)),
                  int(chunk_steps), ctypes.byref(h))
        self.h = h
        self._keep: List[object] = []   # device buffers referenced by submitted requests

    def submit(self, kind: int, desc, nsteps: int, arrival: float, ic: Sequence, out: Sequence) -> int:
        """ic / out: torch device tensors (or None) in the reference layout; kept alive until run() ends."""
        r = _request(kind, desc, nsteps, arrival)
        for i, t in enumerate(ic):
            r.ic[i] = t.data_ptr() if t is not None else None
        for i, t in enumerate(out):
            r.out[i] = t.data_ptr() if t is not None else None
        self._keep.extend([t for t in list(ic) + list(out) if t is not None])
        rid = ctypes.c_int()
        _lib.call("tarang_serve_submit", self.h, ctypes.byref(r), ctypes.byref(rid))
        return rid.value

    def run(self):
        _lib.call("tarang_serve_run", self.h)

    def record(self, rid: int) -> Record:
        t = (ctypes.c_double * 3)()
        gps = ctypes.c_double()
        fp = ctypes.c_size_t()
        lane = ctypes.c_int()
        st = ctypes.c_int()
        _lib.call("tarang_serve_record", self.h, int(rid), t, ctypes.byref(gps), ctypes.byref(fp),
                  ctypes.byref(lane), ctypes.byref(st))
        nd = ctypes.c_int()
        msg = ctypes.create_string_buffer(1024)
        _lib.call("tarang_serve_record_error", self.h, int(rid), ctypes.byref(nd), msg, 1024)
        return Record(t[0], t[1], t[2], gps.value, fp.value, lane.value, st.value, nd.value, msg.value.decode())

    def close(self):
        if getattr(self, "h", None):
            _lib.call("tarang_serve_destroy", self.h)
            self.h = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def latency_stats(records: Sequence[Record]) -> dict:
    """p50 / p99 time to completion and aggregate grid-point steps/s over the finished requests."""
    done = [r for r in records if r.status == 2]
    if not done:
        return {"n": 0}
    lat = np.array([r.latency for r in done])
    span = max(r.done for r in done) - min(r.arrival for r in done)
    return {"n": len(done), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": span,
            "grid_point_steps_per_s": float(sum(r.grid_point_steps for r in done) / span) if span > 0 else None,
            "rejected": sum(1 for r in records if r.status == -1),
            "failed": sum(1 for r in records if r.status == -2)}
