"""Snapshot stream: state output from a running time loop.

`SnapshotWriter` is an `after_step` hook for `solve_` (the reference's
`solve!(q, timeend, solver; after_step)`, src/odesolvers.jl:9-26): every
`every` steps it stages the state on the device, has the library copy it to
pinned host memory on a second stream (`atm_snapshot`) while the compute
stream continues, and a writer thread puts it on disk once the copy has landed.
A ring of `ring` buffers bounds the memory and the copies in flight.

File format (little-endian): b"ATMSNAP1", uint64 header length, a JSON header
(shape, dtype, step, time and caller metadata), then the array bytes in C order.
With fmt="vtk" the writer thread writes each snapshot as a .vtu instead (the
fields interpolated to equally spaced points, as the reference's experiments
write them, `vtk.py`) and `close()` writes a .pvd collection of the run.
"""
from __future__ import annotations

import ctypes
import json
import queue
import struct
import threading

import numpy as np

from . import _lib
from . import vtk as _vtk
from .fields import Field

MAGIC = b"ATMSNAP1"

_lib._sig("atm_host_alloc", _lib.c_int, [_lib.c_void_p, ctypes.c_size_t, ctypes.POINTER(_lib.c_void_p)])
_lib._sig("atm_host_free", _lib.c_int, [_lib.c_void_p, _lib.c_void_p])
_lib._sig("atm_snapshot", _lib.c_int, [_lib.c_void_p, _lib.c_int, _lib.c_void_p, _lib.c_void_p, _lib.c_void_p,
                                       ctypes.c_size_t])
_lib._sig("atm_snapshot_query", _lib.c_int, [_lib.c_void_p, _lib.c_int])
_lib._sig("atm_snapshot_wait", _lib.c_int, [_lib.c_void_p, _lib.c_int])


def write_snapshot(path, array, meta=None):
    """Write one snapshot file (see the module docstring for the format)."""
    a = np.ascontiguousarray(array)
    head = dict(meta or {})
    head.update(shape=list(a.shape), dtype=a.dtype.str)
    hb = json.dumps(head).encode()
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(hb)))
        f.write(hb)
        f.write(memoryview(a).cast("B"))


def read_snapshot(path):
    """(array, header) of a snapshot file."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a snapshot file")
        (n,) = struct.unpack("<Q", f.read(8))
        head = json.loads(f.read(n).decode())
        a = np.frombuffer(f.read(), dtype=np.dtype(head["dtype"]))
    return a.reshape(head["shape"]), head


def load_checkpoint(path, dg):
    """Resume from a snapshot: (device field of `dg`, step, time).  The solvers of
    this package carry no state between steps beyond q and the time (LSRK's first
    stage overwrites dq, ARK23 and RLSRK steps start from q alone), so a run
    continued from a snapshot with `t0 = time` repeats the uninterrupted run bit
    for bit."""
    a, head = read_snapshot(path)
    q = dg.fieldarray(a)
    return q, int(head.get("step", 0)), float(head.get("time", 0.0))


class PinnedBuffer:
    """Page-locked host memory from the library (hipHostMalloc), viewed as numpy."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _lib.c_void_p()
        ctx.check(_lib.lib.atm_host_alloc(ctx.h, self.nbytes, ctypes.byref(p)), "atm_host_alloc")
        self.ptr = p
        buf = (ctypes.c_char * self.nbytes).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=self.dtype).reshape(self.shape)

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                self.array = None
                _lib.lib.atm_host_free(self.ctx.h, self.ptr)
                self.ptr = None
        except Exception:
            pass


class SnapshotWriter:
    """after_step hook writing the state every `every` steps to
    pattern.format(step=..., time=...) without stalling the device."""

    def __init__(self, q: Field, pattern: str, every: int = 1, ring: int = 2, meta=None, first_slot=0,
                 solver=None, fmt: str = "atm", grid=None, names=None, reference=None, pvd: str | None = None):
        """solver: if given, the header's time is solver.time when the hook runs (the
        time of the state written); otherwise the time the hook receives.
        fmt="vtk": write .vtu files; needs `grid` (node coordinates `points`, N, dim) and
        takes field `names` (default q1, q2, ...), an optional `reference` state
        subtracted before writing (as the reference's experiments write q - q_ref) and
        the `pvd` collection path written by close()."""
        if not 1 <= ring <= 16 or first_slot < 0 or first_slot + ring > 16:
            raise ValueError("ring slots must lie in [0, 16)")
        if fmt not in ("atm", "vtk"):
            raise ValueError("fmt must be 'atm' or 'vtk'")
        if fmt == "vtk" and grid is None:
            raise ValueError("VTK output needs the grid")
        self.fmt, self.grid, self.pvd = fmt, grid, pvd
        self.names = list(names) if names is not None else [f"q{s + 1}" for s in range(q.shape[0])]
        if len(self.names) != q.shape[0]:
            raise ValueError("one name per state component")
        self.reference = None if reference is None else np.asarray(reference, dtype=np.float64)
        self.pvd_entries = []
        self.ctx = q.ctx
        self.pattern, self.every = pattern, int(every)
        self.meta = dict(meta or {})
        self.solver = solver
        self.staging = [q.similar() for _ in range(ring)]
        self.host = [PinnedBuffer(self.ctx, q.shape) for _ in range(ring)]
        self.slots = list(range(first_slot, first_slot + ring))
        self.free = [threading.Event() for _ in range(ring)]
        for ev in self.free:
            ev.set()
        self.files = []
        self.errors = []
        self._k = 0
        self._tasks = queue.Queue()
        self._thread = threading.Thread(target=self._writer, daemon=True)
        self._thread.start()

    def _writer(self):
        while True:
            task = self._tasks.get()
            if task is None:
                return
            i, path, head = task
            try:
                self.ctx.check(_lib.lib.atm_snapshot_wait(self.ctx.h, self.slots[i]), "atm_snapshot_wait")
                if self.fmt == "vtk":
                    a = self.host[i].array
                    if self.reference is not None:
                        a = a - self.reference
                    _vtk.write_vtu(path, self.grid.points, dict(zip(self.names, a)), self.grid.N, self.grid.dim)
                    self.pvd_entries.append((head["time"], path))
                else:
                    write_snapshot(path, self.host[i].array, head)
                self.files.append(path)
            except Exception as exc:  # reported by close()
                self.errors.append(exc)
            finally:
                self.free[i].set()

    def __call__(self, step, time, q: Field):
        if step % self.every:
            return
        i = self._k % len(self.slots)
        self._k += 1
        self.free[i].wait()  # the writer has finished with this buffer
        self.free[i].clear()
        self.ctx.check(_lib.lib.atm_snapshot(self.ctx.h, self.slots[i], self.host[i].ptr, self.staging[i].ptr,
                                             q.ptr, q.nbytes), "atm_snapshot")
        t = self.solver.time if self.solver is not None else time
        head = dict(self.meta, step=int(step), time=float(t))
        self._tasks.put((i, self.pattern.format(step=step, time=time), head))

    def close(self):
        """Wait for every snapshot to be on disk; raise the first write error."""
        self._tasks.put(None)
        self._thread.join()
        if self.errors:
            raise self.errors[0]
        if self.fmt == "vtk" and self.pvd:
            _vtk.write_pvd(self.pvd, self.pvd_entries)
        return list(self.files)
