"""Serving loop: simulation requests arrive over time and share the device.

Continuous batching: every tick admits the requests that have arrived (while
the device still holds fewer than `max_active` requests and their HBM footprint
fits the budget), advances every active request by a common chunk of steps —
all explicit (LSRK) requests in one batched launch per kernel group and stage
on the context's stream, implicit-explicit (ARK23) requests spread over
`imex_streams` further streams that run beside it (each request's steps in
order on one stream; the streams fork from and join back into the context's
stream every tick) — and retires
the requests that finished, freeing their slots and memory for the next tick.
A request's state stays resident in HBM from admission to completion, and its
result is bitwise the one it would have alone (same kernels, same order).
Operators (geometry on the device) are cached per mesh, so repeated meshes
cost one upload.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib
from .batch import Batch
from .fields import Field
from .odesolvers import ARK23, LSRK54, BandedOperator


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
    dof: int = 0
    # device-side state while active
    _dg: object = None
    _q: Optional[Field] = None
    _dq: Optional[Field] = None
    _scratch: Optional[Field] = None
    _dt: float = 0.0
    _ark: object = None
    _left: int = 0


@dataclass
class ServeStats:
    completed: int = 0
    latencies: list = field(default_factory=list)
    dof_updates: float = 0.0
    wall: float = 0.0
    ticks: int = 0
    max_active: int = 0
    admit_s: float = 0.0     # host: building / uploading admitted requests
    enqueue_s: float = 0.0   # host: launching a tick's work
    wait_s: float = 0.0      # host: waiting for the device at the end of a tick
    retire_s: float = 0.0    # host: downloading results, freeing
    gpu_span_s: float = 0.0  # device: from the first to the last kernel of each tick (events)

    def summary(self):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"completed": self.completed, "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "dof_updates_per_s": self.dof_updates / self.wall if self.wall > 0 else 0.0,
                "wall_s": self.wall, "ticks": self.ticks, "max_active": self.max_active,
                "admit_s": self.admit_s, "enqueue_s": self.enqueue_s, "wait_s": self.wait_s,
                "retire_s": self.retire_s, "gpu_span_s": self.gpu_span_s}


class Server:
    def __init__(self, ctx=None, hbm_budget=0.9 * 288e9, max_active=512, chunk=10, keep_results=True,
                 imex_streams=3):
        self.ctx = ctx or _lib.context()
        self.imex_streams = int(imex_streams)
        self._streams = []
        self.hbm_budget = float(hbm_budget)
        self.max_active = int(max_active)
        self.chunk = int(chunk)
        self.keep_results = keep_results
        self.cache = {}
        self.in_use = 0

    def _admit(self, r: Request, now):
        built = r.build(self.cache)
        if r.kind == "imex":
            ex, lin, q0, dt = built
            r._dg = lin
            r._q = lin.fieldarray(q0)
            # requests on the same mesh with the same time step share one factorisation
            bkey = ("banded", r.key, dt)
            if bkey not in self.cache:
                self.cache[bkey] = BandedOperator(lin, -dt * (1 - 1 / math.sqrt(2)))
            r._ark = ARK23(ex, lin, r._q, dt, split_rhs=False, banded=self.cache[bkey])
        else:
            dg, q0, dt = built
            r._dg = dg
            r._q = dg.fieldarray(q0)
            r._dq = r._q.similar().fill_(0.0)
            r._scratch = r._q.similar()
        r._dt = dt
        r.dof = int(q0.size)
        r._left = r.steps
        r.admitted = now
        self.in_use += r.footprint

    def _retire(self, r: Request, now, stats: ServeStats):
        self.ctx.sync()
        if self.keep_results:
            r.result = r._q.numpy()
        r.finished = now
        stats.completed += 1
        stats.latencies.append(r.finished - r.arrival)
        self.in_use -= r.footprint
        r._q = r._dq = r._scratch = r._ark = r._dg = None

    def _run_implicit(self, implicit, step):
        """ARK23 chunks of the implicit requests, round-robin over the side streams
        (or on the context's stream when imex_streams is 0)."""
        if self.imex_streams <= 0:
            for r in implicit:
                r._ark.run(r._q, step)
            return
        while len(self._streams) < self.imex_streams:
            self._streams.append(self.ctx.new_stream())
        main = self.ctx.stream
        used = self._streams[:min(len(implicit), self.imex_streams)]
        for s in used:
            self.ctx.stream_wait(s, main)
        try:
            for i, r in enumerate(implicit):
                self.ctx.use_stream(used[i % len(used)])
                r._ark.run(r._q, step)
        finally:
            self.ctx.use_stream(main)
        for s in used:
            self.ctx.stream_wait(main, s)

    def close(self):
        """Release the side streams (after their work has finished)."""
        for s in self._streams:
            self.ctx.destroy_stream(s)
        self._streams = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, requests, on_tick=None):
        """Serve `requests`; a request becomes available when the wall clock since the
        start of the run passes its arrival time."""
        pending = sorted(requests, key=lambda r: r.arrival)
        active = []
        stats = ServeStats()
        batch, members = None, None
        t0 = time.perf_counter()
        pi = 0
        while pi < len(pending) or active:
            now = time.perf_counter() - t0
            ta = time.perf_counter()
            while (pi < len(pending) and pending[pi].arrival <= now and len(active) < self.max_active
                   and self.in_use + pending[pi].footprint <= self.hbm_budget):
                self._admit(pending[pi], time.perf_counter() - t0)
                active.append(pending[pi])
                pi += 1
            stats.admit_s += time.perf_counter() - ta
            if not active:
                if pi < len(pending):
                    time.sleep(max(0.0, min(0.01, pending[pi].arrival - now)))
                continue
            stats.max_active = max(stats.max_active, len(active))
            explicit = [r for r in active if r.kind != "imex"]
            implicit = [r for r in active if r.kind == "imex"]
            step = min(self.chunk, min(r._left for r in active))
            te = time.perf_counter()
            self.ctx.record(14)
            if explicit:
                ids = [r.rid for r in explicit]
                if ids != members:
                    batch = Batch([r._dg for r in explicit], self.ctx)
                    members = ids
                batch.lsrk_run([r._q for r in explicit], [r._dq for r in explicit],
                               [r._scratch for r in explicit], [r._dt for r in explicit], step, LSRK54)
            if implicit:
                self._run_implicit(implicit, step)
            self.ctx.record(15)
            tw = time.perf_counter()
            stats.enqueue_s += tw - te
            self.ctx.sync()
            stats.wait_s += time.perf_counter() - tw
            stats.gpu_span_s += self.ctx.elapsed_ms(14, 15) / 1e3
            stats.ticks += 1
            for r in active:
                r._left -= step
                stats.dof_updates += r.dof * step * (3 if r.kind == "imex" else 5)
            now = time.perf_counter() - t0
            done = [r for r in active if r._left == 0]
            tr = time.perf_counter()
            for r in done:
                self._retire(r, now, stats)
            stats.retire_s += time.perf_counter() - tr
            active = [r for r in active if r._left > 0]
            if on_tick:
                on_tick(stats, active)
        stats.wall = time.perf_counter() - t0
        return stats


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
