"""Batched instances: many DGSEM operators advance in one launch per stage.

Members that share a kernel form a group whose elements are laid out back to
back (a flattened element table with per-member offsets); each group is one
launch per stage, and the two-step stage sequence of all groups is captured
in one hipGraph (atm_batch_* in include/atum_hip.h).
"""
from __future__ import annotations

import ctypes

from . import _lib
from .fields import Field
from .odesolvers import LSRK54, LSRK144

_lib._sig("atm_batch_create", _lib.c_int, [_lib.c_void_p, _lib.c_int32, ctypes.POINTER(_lib.c_void_p),
                                           ctypes.POINTER(_lib.c_void_p)])
_lib._sig("atm_batch_destroy", _lib.c_int, [_lib.c_void_p])
_lib._sig("atm_batch_ngroups", _lib.c_int, [_lib.c_void_p])
_lib._sig("atm_batch_lsrk_run", _lib.c_int, [_lib.c_void_p, _lib.c_int, ctypes.POINTER(_lib.c_void_p),
                                             ctypes.POINTER(_lib.c_void_p), ctypes.POINTER(_lib.c_void_p),
                                             ctypes.POINTER(ctypes.c_double), _lib.c_int64, _lib.c_int])


def _ptrs(fields):
    arr = (_lib.c_void_p * len(fields))()
    for i, f in enumerate(fields):
        arr[i] = f.ptr.value
    return arr


class Batch:
    """A set of DGSEM operators advanced together (atm_batch)."""

    def __init__(self, dgs, ctx=None):
        self.ctx = ctx or (dgs[0].ctx if dgs else _lib.context())
        self.dgs = list(dgs)
        hs = (_lib.c_void_p * len(self.dgs))(*[d.h.value for d in self.dgs])
        h = _lib.c_void_p()
        self.ctx.check(_lib.lib.atm_batch_create(self.ctx.h, len(self.dgs), hs, ctypes.byref(h)),
                       "atm_batch_create")
        self.h = h

    @property
    def ngroups(self):
        return _lib.lib.atm_batch_ngroups(self.h)

    def lsrk_run(self, qs, dqs, scratches, dts, nsteps, scheme=LSRK54, use_graph=True):
        dt = (ctypes.c_double * len(dts))(*[float(x) for x in dts])
        self.ctx.check(_lib.lib.atm_batch_lsrk_run(self.h, scheme.scheme, _ptrs(qs), _ptrs(dqs),
                                                   _ptrs(scratches), dt, int(nsteps), int(use_graph)),
                       "atm_batch_lsrk_run")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib.atm_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LSRKEnsemble:
    """LSRK time stepping of many instances, each with its own dt, in batched launches."""

    def __init__(self, dgs, qs, dts, scheme=LSRK54, use_graph=True):
        if not (len(dgs) == len(qs) == len(dts)):
            raise ValueError("one state and one dt per operator")
        self.batch = Batch(dgs)
        self.qs = list(qs)
        self.dqs = [q.similar().fill_(0.0) for q in qs]
        self.scratch = [q.similar() for q in qs]
        self.dts = [float(x) for x in dts]
        self.scheme = scheme
        self.use_graph = use_graph
        self.steps = 0

    def run(self, nsteps):
        self.batch.lsrk_run(self.qs, self.dqs, self.scratch, self.dts, nsteps, self.scheme,
                            self.use_graph)
        self.steps += nsteps


__all__ = ["Batch", "LSRKEnsemble", "LSRK54", "LSRK144", "Field"]
