"""ctypes binding of libatumhip.so (the C ABI declared in include/atum_hip.h).

The library is the product: there is no CPU fallback.  Importing this module
when the library has not been built raises immediately.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ATUM_HIP_LIB",
                          os.path.join(os.path.dirname(_HERE), "build", "libatumhip.so"))

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libatumhip.so not found at {LIB_PATH}; build it first "
                      "(python atum.jl_amd/build_lib.py or __graft_entry__.build())")

lib = ctypes.CDLL(LIB_PATH)

c_int, c_int32, c_int64, c_double, c_void_p = (ctypes.c_int, ctypes.c_int32, ctypes.c_int64,
                                                ctypes.c_double, ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)

ATM_OK = 0

LAWS = {"advection": 0, "euler": 1, "euler_gravity": 2, "linear_euler_gravity": 3, "shallow_water": 4}
FLUX_IDS = {"central": 0, "rusanov": 1, "ec": 2, "kg": 3, "roe": 4, "matrix": 5}
BC_IDS = {"copy": 0, "reflect": 1, "zero_inflow": 2, "prescribed": 3}
SOURCE_IDS = {None: 0, "coriolis": 1}
PRECISION_IDS = {"reference": 0, "fast": 1}
FD_KERNEL_IDS = {"per_dir": 0, "per_dir_symmetric": 1, "naive": 2}


class DgDesc(ctypes.Structure):
    _fields_ = [("law", c_int32), ("dim", c_int32), ("N", c_int32), ("K", c_int32),
                ("volume_form", c_int32), ("volume_flux", c_int32), ("surface_flux", c_int32),
                ("dirmask", c_int32), ("bc", c_int32), ("source", c_int32),
                ("pde_level_balance", c_int32), ("naux", c_int32), ("n_bc_states", c_int32),
                ("precision", c_int32), ("fd_kernel", c_int32), ("reserved", c_int32), ("gamma", c_double), ("grav", c_double),
                ("omega", c_double), ("u", c_double * 3), ("bc_states", _dp), ("D", _dp),
                ("g", _dp), ("MJ", _dp), ("MJI", _dp), ("faceMJ", _dp), ("normal", _dp),
                ("aux", _dp), ("faceix_m", _i64p), ("faceix_p", _i64p), ("bctag", _i32p)]


class BrickDesc(ctypes.Structure):
    _fields_ = [("Kd", c_int32 * 3), ("periodic", c_int32 * 3), ("stacked", c_int32),
                ("coords", _dp * 3), ("xi", _dp), ("w", _dp)]


# atm_dg_geometry arrays: id, dtype, rows (1, dim, dim*dim), per (node, face, tag)
GEOMETRY = {"points": (0, np.float64, "dim", "node"), "g": (1, np.float64, "dim2", "node"),
            "MJ": (2, np.float64, "1", "node"), "MJI": (3, np.float64, "1", "node"),
            "faceMJ": (4, np.float64, "1", "face"), "normal": (5, np.float64, "dim", "face"),
            "faceix_p": (6, np.int64, "1", "face"), "bctag": (7, np.int32, "1", "tag")}


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("atm_ctx_create", c_int, [c_int, ctypes.POINTER(c_void_p)])
_sig("atm_ctx_destroy", c_int, [c_void_p])
_sig("atm_ctx_set_stream", c_int, [c_void_p, c_void_p])
_sig("atm_ctx_get_stream", c_void_p, [c_void_p])
_sig("atm_stream_create", c_int, [c_void_p, ctypes.POINTER(c_void_p)])
_sig("atm_stream_destroy", c_int, [c_void_p, c_void_p])
_sig("atm_stream_sync", c_int, [c_void_p, c_void_p])
_sig("atm_stream_wait", c_int, [c_void_p, c_void_p, c_void_p])
_sig("atm_last_error", ctypes.c_char_p, [c_void_p])
_sig("atm_sync", c_int, [c_void_p])
_sig("atm_version", ctypes.c_char_p, [])
_sig("atm_alloc", c_int, [c_void_p, ctypes.c_size_t, ctypes.POINTER(c_void_p)])
_sig("atm_free", c_int, [c_void_p, c_void_p])
_sig("atm_upload", c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t])
_sig("atm_download", c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t])
_sig("atm_copy", c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t])
_sig("atm_dg_create", c_int, [c_void_p, ctypes.POINTER(DgDesc), ctypes.POINTER(c_void_p)])
_sig("atm_dg_create_brick", c_int, [c_void_p, ctypes.POINTER(DgDesc), ctypes.POINTER(BrickDesc),
                                    ctypes.POINTER(c_void_p)])
_sig("atm_dg_geometry", c_int, [c_void_p, c_int, c_void_p])
_sig("atm_dg_destroy", c_int, [c_void_p])
_sig("atm_dg_nstates", c_int, [c_void_p])
_sig("atm_dg_nnodes", c_int64, [c_void_p])
_sig("atm_dg_device_bytes", c_int64, [c_void_p])
_sig("atm_dg_get_epw", c_int, [c_void_p])
_sig("atm_dg_set_schedule", c_int, [c_void_p, ctypes.POINTER(ctypes.c_int32), c_int64])
_sig("atm_dg_set_epw", c_int, [c_void_p, c_int])
_sig("atm_dg_epw_candidates", c_int, [c_void_p, ctypes.POINTER(ctypes.c_int32), c_int])
_sig("atm_dg_autotune", c_int, [c_void_p, c_void_p, c_void_p, c_int, ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_float), c_int])
_sig("atm_dg_tendency", c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int])
_sig("atm_lsrk_stage", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double])
_sig("atm_lsrk_run", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_double, c_double,
                             c_int64, c_int])
_sig("atm_lsrk_run_noalloc", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double, c_double, c_int64])
_sig("atm_banded_create", c_int, [c_void_p, c_int32, _i32p, c_int32, c_double,
                                  ctypes.POINTER(c_void_p)])
_sig("atm_banded_destroy", c_int, [c_void_p])
_sig("atm_banded_solve", c_int, [c_void_p, c_void_p, c_void_p])
_sig("atm_banded_shape", c_int, [c_void_p, _i64p, _i32p, _i32p])
_sig("atm_banded_download", c_int, [c_void_p, _dp])
_sig("atm_banded_nunique", c_int64, [c_void_p])
_sig("atm_banded_device_bytes", c_int64, [c_void_p])
_sig("atm_banded_set_kernel", c_int, [c_void_p, c_int])
_sig("atm_banded_get_kernel", c_int, [c_void_p])
_sig("atm_banded_set_waves", c_int, [c_void_p, c_int])
_sig("atm_banded_get_waves", c_int, [c_void_p])
_sig("atm_banded_autotune", c_int, [c_void_p, c_void_p, c_void_p, c_int, _i32p, ctypes.POINTER(ctypes.c_float), c_int])
BAND_KERNELS = {"auto": -1, "mfma": 0, "block": 1, "reg": 2, "lds": 3, "plain": 4}
_sig("atm_ark23_run", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double,
                              c_int64, c_int, c_int])
_sig("atm_event_record", c_int, [c_void_p, c_int])
_sig("atm_event_elapsed_ms", c_int, [c_void_p, c_int, c_int, ctypes.POINTER(ctypes.c_float)])
_sig("atm_axpby", c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_int64])
_sig("atm_entropy_integral", c_int, [c_void_p, c_void_p, _dp])
_sig("atm_entropy_product", c_int, [c_void_p, c_void_p, c_void_p, _dp])
_sig("atm_weightednorm", c_int, [c_void_p, c_void_p, c_double, c_int, _dp])
_sig("atm_rlsrk_run", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double, c_double, c_int64, _dp])

EXPORTED = [n for n in dir(lib) if n.startswith("atm_")]


class AtumError(RuntimeError):
    pass


class Context:
    """One device and one HIP stream (atm_ctx)."""

    def __init__(self, device: int = 0):
        h = c_void_p()
        rc = lib.atm_ctx_create(device, ctypes.byref(h))
        if rc != ATM_OK:
            raise AtumError(f"atm_ctx_create(device={device}) failed with status {rc}")
        self.h = h
        self.device = device

    def check(self, rc, what="call"):
        if rc != ATM_OK:
            msg = lib.atm_last_error(self.h)
            raise AtumError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    def sync(self):
        self.check(lib.atm_sync(self.h), "atm_sync")

    @property
    def stream(self):
        return lib.atm_ctx_get_stream(self.h)

    def new_stream(self):
        """A further HIP stream on this device (default priority); see `use_stream`."""
        h = c_void_p()
        self.check(lib.atm_stream_create(self.h, ctypes.byref(h)), "atm_stream_create")
        return h

    def destroy_stream(self, stream):
        self.check(lib.atm_stream_destroy(self.h, stream), "atm_stream_destroy")

    def stream_wait(self, waiter=None, waitee=None):
        """Order `waiter` after the work already enqueued on `waitee` (None: the
        context's current stream)."""
        self.check(lib.atm_stream_wait(self.h, waiter, waitee), "atm_stream_wait")

    def use_stream(self, stream=None):
        """Launch on `stream` from now on (None: the context's own stream)."""
        self.check(lib.atm_ctx_set_stream(self.h, stream), "atm_ctx_set_stream")

    def record(self, slot):
        self.check(lib.atm_event_record(self.h, slot), "atm_event_record")

    def elapsed_ms(self, a, b):
        ms = ctypes.c_float(0)
        self.check(lib.atm_event_elapsed_ms(self.h, a, b, ctypes.byref(ms)), "atm_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib.atm_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass


_default = None


def context() -> Context:
    """The process-wide default context on device 0."""
    global _default
    if _default is None:
        _default = Context(0)
    return _default
