"""Autotuner of launch shapes: elements per workgroup of the element kernels and
the solve kernel of the banded vertical solve (with, for the matrix-core solve, its
waves per workgroup).

The number of elements per workgroup (epw) of an operator's element kernel
changes how the device is filled, never the results (each element is computed
by its own threads in a fixed order).  `autotune` times the shapes that fit
(`atm_dg_autotune`: tendency launches timed with HIP events) and keeps the
fastest.  The banded solve has five kernels of different tile shapes (16-line
block form on the matrix cores or the vector units, 64-line register-blocked
substitution, LDS and plain substitution) whose availability and speed depend
on the degree; `autotune_banded` times the ones that can run a factorisation
(`atm_banded_autotune`) and keeps the fastest.  A `TuneCache` remembers each
choice per kernel configuration and size class, in memory or in a JSON file, so
later operators of the same kind skip the timing.
"""
from __future__ import annotations

import ctypes
import json
import math
import os

from . import _lib
from .fields import Field


def _size_class(K: int) -> int:
    """Element counts within a factor of two share a tuning."""
    return 0 if K <= 0 else int(round(math.log2(K)))


class TuneCache:
    """Chosen epw per (kernel configuration, size class); persisted to `path` if given."""

    def __init__(self, path: str | None = None):
        self.path = path
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    @staticmethod
    def key(dg) -> str:
        return json.dumps(list(dg.kernel_key) + [_size_class(dg.grid.K)])

    def get(self, dg):
        return self.table.get(self.key(dg))

    def put(self, dg, epw: int):
        self.store(self.key(dg), int(epw))

    def store(self, key: str, value):
        self.table[key] = value
        if self.path:
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)


def autotune(dg, q, reps: int = 20, cache: TuneCache | None = None):
    """Choose the operator's elements per workgroup by timing; returns (epw, {epw: us per launch}).
    `q` is a representative state (Field or host array of the operator's shape); the
    timings are empty when the cache already holds the choice."""
    if cache is not None:
        hit = cache.get(dg)
        if hit is not None and hit in dg.epw_candidates():
            dg.epw = hit
            return hit, {}
    qf = q if isinstance(q, Field) else dg.fieldarray(q)
    dq = dg.fieldarray()
    cap = 8
    ep = (ctypes.c_int32 * cap)()
    us = (ctypes.c_float * cap)()
    best = _lib.lib.atm_dg_autotune(dg.h, qf.ptr, dq.ptr, int(reps), ep, us, cap)
    dg.ctx.check(best if best < 0 else 0, "atm_dg_autotune")
    n = len(dg.epw_candidates())
    times = {int(ep[i]): float(us[i]) for i in range(min(n, cap))}
    if cache is not None:
        cache.put(dg, best)
    return int(best), times


def autotune_batch(dgs, q0s, dts, steps: int = 2, rounds: int = 1):
    """Choose elements per workgroup for batched launches: the members sharing a kernel
    configuration form one launch group, so each group gets one shape, found by
    coordinate descent over the groups (each candidate timed on a full batched run of
    `steps` LSRK54 steps from copies of the initial states, the other groups at their
    current shapes).  Returns {kernel configuration: (epw, {epw: ms per run})}."""
    from .batch import Batch, LSRK54

    if not dgs:
        return {}
    ctx = dgs[0].ctx
    groups = {}
    for d in dgs:
        groups.setdefault(d.kernel_key, []).append(d)
    qs = [d.fieldarray(q) for d, q in zip(dgs, q0s)]
    dqs = [d.fieldarray() for d in dgs]
    scr = [d.fieldarray() for d in dgs]

    def run_ms():
        b = Batch(dgs, ctx)
        for q, d, q0 in zip(qs, dgs, q0s):
            q.upload(q0)
        b.lsrk_run(qs, dqs, scr, dts, steps, LSRK54)  # capture and warm up
        ctx.sync()
        ctx.record(12)
        b.lsrk_run(qs, dqs, scr, dts, steps, LSRK54)
        ctx.record(13)
        return ctx.elapsed_ms(12, 13)

    out = {}
    for _ in range(rounds):
        for key, members in groups.items():
            times = {}
            for e in members[0].epw_candidates():
                if not all(e in m.epw_candidates() for m in members):
                    continue
                for m in members:
                    m.epw = e
                times[e] = run_ms()
            best = min(times, key=times.get)
            for m in members:
                m.epw = best
            out[key] = (best, times)
    return out


def banded_key(band) -> str:
    """Tuning key of a banded factorisation: law, degree, elements per column and the
    size class of its line count."""
    lin = band.linrhs
    _, KV, ncols = lin.grid.columns()
    return json.dumps(["banded", lin.law.law, lin.grid.dim, lin.grid.N, int(KV),
                       _size_class(int(ncols) * (lin.grid.Np // lin.grid.Nq))])


def autotune_banded(band, reps: int = 10, cache: TuneCache | None = None):
    """Choose the banded solve kernel by timing solves of a right-hand side of ones;
    returns (kernel name, {name: us per solve}) (timings empty on a cache hit)."""
    key = banded_key(band)
    if cache is not None and key in cache.table:
        entry = cache.table[key]
        try:
            kernel, waves = (entry["kernel"], entry.get("waves", 0)) if isinstance(entry, dict) else (entry, 0)
            band.kernel = kernel
            if kernel == "mfma":
                band.waves = waves
            return band.kernel, {}
        except (_lib.AtumError, KeyError, TypeError):
            pass
    lin = band.linrhs
    rhs = lin.fieldarray()
    rhs.fill_(1.0)
    best, times = band.autotune(lin.fieldarray(), rhs, reps=reps)
    if cache is not None:
        cache.store(key, {"kernel": best, "waves": band.waves if best == "mfma" else 0})
    return best, times
