"""Device-resident field arrays (Bennu `fieldarray` replacement).

A field with Ns components lives in one device allocation of Ns*Np*K doubles,
component-major, each component a column-major (Np, K) array — the reference's
SoA data model.  Host copies are numpy arrays of shape (Ns, Np*K).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Field:
    def __init__(self, ns: int, ntot: int, ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.context()
        self.shape = (int(ns), int(ntot))
        self.nbytes = 8 * self.shape[0] * self.shape[1]
        p = ctypes.c_void_p()
        self.ctx.check(_lib.lib.atm_alloc(self.ctx.h, self.nbytes, ctypes.byref(p)), "atm_alloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, ctx=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.ndim == 1:
            a = a[None, :]
        f = cls(a.shape[0], a.shape[1], ctx)
        f.upload(a)
        return f

    def similar(self):
        return Field(self.shape[0], self.shape[1], self.ctx)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        self.ctx.check(_lib.lib.atm_upload(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes), "atm_upload")

    def numpy(self):
        out = np.empty(self.shape)
        self.ctx.check(_lib.lib.atm_download(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes),
                       "atm_download")
        return out

    def copy_(self, other: "Field"):
        assert other.shape == self.shape
        self.ctx.check(_lib.lib.atm_copy(self.ctx.h, self.ptr, other.ptr, self.nbytes), "atm_copy")
        return self

    def fill_(self, value: float):
        self.upload(np.full(self.shape, float(value)))
        return self

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and self.ptr.value:
                _lib.lib.atm_free(self.ctx.h, self.ptr)
                self.ptr = None
        except Exception:
            pass


def fieldarray(law_or_field, grid=None, ctx=None):
    """`fieldarray(law, grid)` or `fieldarray(q)` (a zeroed field of the same shape)."""
    if isinstance(law_or_field, Field):
        f = law_or_field.similar()
    else:
        f = Field(law_or_field.nstates, grid.Ntot, ctx)
    f.fill_(0.0)
    return f
