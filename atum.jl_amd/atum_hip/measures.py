"""Measures of DGSEM fields on the device (src/measures.jl:6-33).

`weightednorm`, `entropyintegral` and `entropyproduct` reduce over all nodes
in the library (`atm_weightednorm`, `atm_entropy_integral`,
`atm_entropy_product`); a numpy argument is uploaded first.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .fields import Field


def _field(dg, a):
    if isinstance(a, Field):
        return a
    return Field.from_numpy(np.asarray(a, dtype=np.float64).reshape(dg.Ns, -1), dg.ctx)


def weightednorm(dg, q, p=2, componentwise=False):
    """MJ-weighted p-norm over all components (src/measures.jl:6-21); p may be np.inf."""
    f = _field(dg, q)
    out = np.zeros(f.shape[0] if componentwise else 1)
    dg.ctx.check(_lib.lib.atm_weightednorm(dg.h, f.ptr, float(p), int(bool(componentwise)),
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                 "atm_weightednorm")
    return out if componentwise else float(out[0])


def entropyintegral(dg, q):
    """sum(MJ .* entropy.(law, q, aux)) (src/measures.jl:23-27)."""
    f = _field(dg, q)
    out = ctypes.c_double(0.0)
    dg.ctx.check(_lib.lib.atm_entropy_integral(dg.h, f.ptr, ctypes.byref(out)), "atm_entropy_integral")
    return out.value


def entropyproduct(dg, p, q):
    """sum(MJ .* dot.(entropyvariables.(law, p, aux), q)) (src/measures.jl:28-33)."""
    fp, fq = _field(dg, p), _field(dg, q)
    out = ctypes.c_double(0.0)
    dg.ctx.check(_lib.lib.atm_entropy_product(dg.h, fp.ptr, fq.ptr, ctypes.byref(out)),
                 "atm_entropy_product")
    return out.value
