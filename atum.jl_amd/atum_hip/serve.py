"""Serving loop: simulation requests arrive over time and share the device.

`Server` drives the library's serving scheduler (atm_server_*): requests are
built (operators cached per mesh, so repeated meshes cost one upload) and
submitted from this thread as they arrive, while the library's worker thread
admits them (while the device holds fewer than `max_active` requests and
their HBM footprint fits the budget), advances every resident request by a
common chunk of steps per tick — the explicit (LSRK) requests in one batched
launch per kernel group and stage, the implicit-explicit (ARK23) requests on
`imex_streams` side streams beside it — and retires the finished ones,
copying their final state to the host.  Building the next requests on the host
therefore overlaps the device work.  A request's result is bitwise the one it
would have alone (same kernels, same order).
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib
from .odesolvers import ARK23, LSRK54, BandedOperator
from .tuning import TuneCache, autotune

c_int32, c_int64, c_double, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p

REQ_LSRK, REQ_ARK23 = 0, 1
QUEUED, ACTIVE, DONE = 0, 1, 2


class _Desc(ctypes.Structure):
    _fields_ = [("max_active", c_int32), ("chunk", c_int32), ("imex_streams", c_int32), ("reserved", c_int32),
                ("hbm_budget", c_double)]


class _Req(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("scheme", c_int32), ("split_rhs", c_int32), ("paperversion", c_int32),
                ("dg", c_void_p), ("lin", c_void_p), ("band", c_void_p), ("q0", c_void_p), ("result", c_void_p),
                ("dt", c_double), ("steps", c_int64)]


_STAT_INT = ("submitted", "completed", "failed", "ticks", "max_active")
_STAT_DBL = ("dof_updates", "busy_s", "gpu_span_s", "admit_s", "enqueue_s", "wait_s", "retire_s")
_STAT_POOL_INT = ("pool_hits", "pool_misses")
_STAT_MEM = ("peak_bytes", "peak_operator_bytes", "pool_bytes")


class _Stats(ctypes.Structure):
    _fields_ = ([(n, c_int64) for n in _STAT_INT] + [(n, c_double) for n in _STAT_DBL]
                + [(n, c_int64) for n in _STAT_POOL_INT] + [(n, c_double) for n in _STAT_MEM])


_P = ctypes.POINTER
_lib._sig("atm_server_create", ctypes.c_int, [c_void_p, _P(_Desc), _P(c_void_p)])
_lib._sig("atm_server_destroy", ctypes.c_int, [c_void_p])
_lib._sig("atm_server_submit", ctypes.c_int, [c_void_p, _P(_Req), _P(c_int64)])
_lib._sig("atm_server_poll", ctypes.c_int, [c_void_p, c_int64])
_lib._sig("atm_server_wait", ctypes.c_int, [c_void_p, c_int64])
_lib._sig("atm_server_next_done", ctypes.c_int, [c_void_p, c_double, _P(c_int64)])
_lib._sig("atm_server_drain", ctypes.c_int, [c_void_p])
_lib._sig("atm_server_clock", c_double, [c_void_p])
_lib._sig("atm_server_times", ctypes.c_int, [c_void_p, c_int64, _P(c_double)])
_lib._sig("atm_server_error", ctypes.c_char_p, [c_void_p, c_int64])
_lib._sig("atm_server_release", ctypes.c_int, [c_void_p, c_int64])
_lib._sig("atm_server_get_stats", ctypes.c_int, [c_void_p, _P(_Stats)])


@dataclass
class Request:
    """One simulation request: `build(cache)` returns (dg, q0, dt) for explicit runs or
    (ex, lin, q0, dt) for IMEX runs (kind == "imex"); `steps` time steps are taken."""
    rid: int
    kind: str
    key: tuple
    steps: int
    arrival: float
    build: Callable
    footprint: int = 0
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[np.ndarray] = None
    error: Optional[str] = None
    dof: int = 0
    _id: int = -1
    _keep: object = None  # host buffers and operators the library reads until the request ends


@dataclass
class ServeStats:
    completed: int = 0
    failed: int = 0
    latencies: list = field(default_factory=list)
    dof_updates: float = 0.0
    wall: float = 0.0
    ticks: int = 0
    max_active: int = 0
    build_s: float = 0.0     # host (caller thread): building and submitting requests
    admit_s: float = 0.0     # worker: allocating and uploading admitted requests
    enqueue_s: float = 0.0   # worker: launching the ticks
    wait_s: float = 0.0      # worker: waiting for the device
    retire_s: float = 0.0    # worker: downloading results, freeing
    gpu_span_s: float = 0.0  # device: from the first to the last kernel of each tick (events)
    clock0: float = 0.0      # scheduler clock at the start of the run (arrival times count from it)

    def summary(self):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"completed": self.completed, "failed": self.failed, "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "dof_updates_per_s": self.dof_updates / self.wall if self.wall > 0 else 0.0,
                "wall_s": self.wall, "ticks": self.ticks, "max_active": self.max_active,
                "build_s": self.build_s, "admit_s": self.admit_s, "enqueue_s": self.enqueue_s,
                "wait_s": self.wait_s, "retire_s": self.retire_s, "gpu_span_s": self.gpu_span_s}


class Server:
    def __init__(self, ctx=None, hbm_budget=0.9 * 288e9, max_active=512, chunk=10, keep_results=True,
                 imex_streams=3, tune: TuneCache | None = None):
        """tune: a TuneCache: every new operator takes its elements per workgroup from
        it, timed once per kernel configuration and size class on first use (while the
        device serves other requests, so the timing is taken under load)."""
        self.ctx = ctx or _lib.context()
        self.hbm_budget = float(hbm_budget)
        self.tune = tune
        self._tuned = set()
        self.keep_results = keep_results
        self.imex_streams = int(imex_streams)
        self.cache = {}
        self._live = {}
        desc = _Desc(int(max_active), int(chunk), self.imex_streams, 0, float(hbm_budget))
        h = c_void_p()
        self.ctx.check(_lib.lib.atm_server_create(self.ctx.h, ctypes.byref(desc), ctypes.byref(h)),
                       "atm_server_create")
        self.h = h

    def submit(self, r: Request) -> int:
        """Build the request's operators (cached per mesh) and hand it to the scheduler."""
        built = r.build(self.cache)
        req = _Req()
        if r.kind == "imex":
            ex, lin, q0, dt = built
            # requests on the same mesh with the same time step share one factorisation
            bkey = ("banded", r.key, dt)
            if bkey not in self.cache:
                self.cache[bkey] = BandedOperator(lin, -dt * (1 - 1 / math.sqrt(2)))
            band = self.cache[bkey]
            req.kind, req.dg, req.lin, req.band = REQ_ARK23, (ex.h if ex is not None else None), lin.h, band.h
            req.split_rhs, req.paperversion = 0, 0
            shape_dg = lin
        else:
            dg, q0, dt = built
            band = None
            req.kind, req.scheme, req.dg = REQ_LSRK, 0, dg.h
            shape_dg = dg
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        if self.tune is not None:
            for op in ((built[0], built[1]) if r.kind == "imex" else (built[0],)):
                if op is not None and id(op) not in self._tuned:
                    autotune(op, q0, reps=10, cache=self.tune)
                    self._tuned.add(id(op))
        ns, nn = int(_lib.lib.atm_dg_nstates(shape_dg.h)), int(_lib.lib.atm_dg_nnodes(shape_dg.h))
        if q0.size != ns * nn:
            raise ValueError(f"request {r.rid}: initial state has {q0.size} values, the operator {ns * nn}")
        r.result = np.empty((ns, nn)) if self.keep_results else None
        req.q0 = q0.ctypes.data
        req.result = r.result.ctypes.data if r.result is not None else None
        req.dt, req.steps = float(dt), int(r.steps)
        r._keep = (built, band, q0)
        r.dof = int(q0.size)
        i = c_int64()
        self.ctx.check(_lib.lib.atm_server_submit(self.h, ctypes.byref(req), ctypes.byref(i)), "atm_server_submit")
        r._id = int(i.value)
        self._live[r._id] = r
        return r._id

    def _finish(self, i: int) -> Request:
        r = self._live.pop(i)
        st = _lib.lib.atm_server_poll(self.h, i)
        t = (c_double * 3)()
        _lib.lib.atm_server_times(self.h, i, t)
        r.admitted, r.finished = t[1], t[2]
        if st < 0:
            r.error = (_lib.lib.atm_server_error(self.h, i) or b"").decode()
            r.result = None
        _lib.lib.atm_server_release(self.h, i)
        r._keep = None
        return r

    def next_done(self, timeout: float = -1.0) -> Optional[Request]:
        """The next finished (or failed) request, or None after `timeout` seconds (< 0: wait)."""
        i = c_int64()
        if _lib.lib.atm_server_next_done(self.h, float(timeout), ctypes.byref(i)) != 1:
            return None
        return self._finish(int(i.value))

    def clock(self) -> float:
        """Seconds on the scheduler's clock (request times use it)."""
        return _lib.lib.atm_server_clock(self.h)

    def stats(self) -> dict:
        st = _Stats()
        _lib.lib.atm_server_get_stats(self.h, ctypes.byref(st))
        return {n: getattr(st, n) for n in _STAT_INT + _STAT_DBL + _STAT_POOL_INT + _STAT_MEM}

    def footprint_stats(self) -> dict:
        """Device memory as the scheduler counted it against its budget: the peak of
        resident buffers plus operators, the operators' share, and the idle pool."""
        st = self.stats()
        return {"budget_bytes": self.hbm_budget, "peak_bytes": st["peak_bytes"],
                "peak_operator_bytes": st["peak_operator_bytes"], "pool_bytes": st["pool_bytes"],
                "pool_hits": st["pool_hits"], "pool_misses": st["pool_misses"]}

    def run(self, requests):
        """Serve `requests`: a request is submitted when the wall clock since the start of
        the run passes its arrival time; returns when all have finished.  Raises
        AtumError if any request failed (after the others have finished)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        before = self.stats()
        stats = ServeStats()
        t0 = time.perf_counter()
        c0 = self.clock()
        stats.clock0 = c0
        pi, waiting, failed = 0, 0, []
        while pi < len(pending) or waiting:
            tb = time.perf_counter()
            while pi < len(pending) and pending[pi].arrival <= tb - t0:
                self.submit(pending[pi])
                pi += 1
                waiting += 1
            stats.build_s += time.perf_counter() - tb
            timeout = 0.05
            if pi < len(pending):
                timeout = max(0.0, min(timeout, pending[pi].arrival - (time.perf_counter() - t0)))
            r = self.next_done(timeout)
            while r is not None:
                waiting -= 1
                if r.error is not None:
                    failed.append(r)
                else:
                    stats.latencies.append(r.finished - (c0 + r.arrival))
                r = self.next_done(0.0)
        stats.wall = time.perf_counter() - t0
        after = self.stats()
        stats.completed = after["completed"] - before["completed"]
        stats.failed = after["failed"] - before["failed"]
        stats.ticks = after["ticks"] - before["ticks"]
        stats.max_active = after["max_active"]
        stats.dof_updates = after["dof_updates"] - before["dof_updates"]
        for k in ("admit_s", "enqueue_s", "wait_s", "retire_s", "gpu_span_s"):
            setattr(stats, k, after[k] - before[k])
        if failed:
            raise _lib.AtumError(f"{len(failed)} request(s) failed; request {failed[0].rid}: {failed[0].error}")
        return stats

    def close(self):
        """Finish every submitted request and stop the scheduler."""
        if getattr(self, "h", None):
            _lib.lib.atm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_alone(r: Request, cache=None):
    """The request's result on its own (same kernels, no batching), for comparison."""
    cache = {} if cache is None else cache
    built = r.build(cache)
    if r.kind == "imex":
        ex, lin, q0, dt = built
        q = lin.fieldarray(q0)
        ARK23(ex, lin, q, dt, split_rhs=False).run(q, r.steps)
        lin.ctx.sync()
        return q.numpy()
    dg, q0, dt = built
    q = dg.fieldarray(q0)
    LSRK54(dg, q, dt).run(q, r.steps)
    dg.ctx.sync()
    return q.numpy()
