"""Time integrators: `solve!`, LSRK54 / LSRK144 and ARK23 (src/odesolvers.jl).

The stage arithmetic runs on the device: an LSRK stage is one fused kernel
(`atm_lsrk_stage`), whole runs of steps are launched by `atm_lsrk_run` (with
hipGraph replay), ARK23 steps by `atm_ark23_run` with the banded vertical
solver.  `solve_` keeps the reference's host loop semantics: the final step is
clipped to land on `timeend` when `adjust_final` (src/odesolvers.jl:9-26).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .fields import Field

# src/odesolvers.jl:148-174, as Float64(num)/Float64(den)
LSRK54_A = (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
            -3550918686646 / 2091501179385, -1275806237668 / 842570457699)
LSRK54_B = (1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
            1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
            2277821191437 / 14882151754819)
LSRK54_C = (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
            2006345519317 / 3224310063776, 2802321613138 / 2924317926251)


class LSRK:
    """Low-storage RK (src/odesolvers.jl:28-57) over a DGSEM operator."""

    scheme = 0
    nstages = 5

    def __init__(self, rhs, q: Field, dt: float, t0: float = 0.0, use_graph: bool = True):
        self.rhs = rhs
        self.dt = float(dt)
        self.time = float(t0)
        self.dq = q.similar().fill_(0.0)
        self._scratch = q.similar()
        self.use_graph = use_graph

    def run(self, q: Field, nsteps: int, dt: float | None = None):
        """Advance `nsteps` steps of size dt (default self.dt) on the device."""
        dt = self.dt if dt is None else float(dt)
        if nsteps <= 0:
            return
        ctx = self.rhs.ctx
        ctx.check(_lib.lib.atm_lsrk_run(self.rhs.h, self.scheme, q.ptr, self.dq.ptr,
                                        self._scratch.ptr, dt, self.time, int(nsteps),
                                        int(self.use_graph)), "atm_lsrk_run")
        for _ in range(nsteps):
            self.time += dt

    def dostep(self, q: Field, after_stage=None):
        if after_stage is None:
            self.run(q, 1)
            return
        # stage by stage, so that the hook sees every stage (src/odesolvers.jl:54)
        A, B, C = self.coefficients()
        for a, b, c in zip(A, B, C):
            self.rhs.lsrk_stage(q, self._scratch, self.dq, a, b * self.dt)
            q.copy_(self._scratch)
            after_stage(self.time + c * self.dt, q)
        self.time += self.dt


class LSRK54(LSRK):
    scheme = 0
    nstages = 5

    @staticmethod
    def coefficients():
        return LSRK54_A, LSRK54_B, LSRK54_C


class LSRK144(LSRK):
    scheme = 1
    nstages = 14

    @staticmethod
    def coefficients():
        a = (0, -0.7188012108672410, -0.7785331173421570, -0.0053282796654044,
             -0.8552979934029281, -3.9564138245774565, -1.5780575380587385,
             -2.0837094552574054, -0.7483334182761610, -0.7032861106563359,
             0.0013917096117681, -0.0932075369637460, -0.9514200470875948,
             -7.1151571693922548)
        b = (0.0367762454319673, 0.3136296607553959, 0.1531848691869027, 0.0030097086818182,
             0.3326293790646110, 0.2440251405350864, 0.3718879239592277, 0.6204126221582444,
             0.1524043173028741, 0.0760894927419266, 0.0077604214040978, 0.0024647284755382,
             0.0780348340049386, 5.5059777270269628)
        c = (0, 0.0367762454319673, 0.1249685262725025, 0.2446177702277698, 0.2476149531070420,
             0.2969311120382472, 0.3978149645802642, 0.5270854589440328, 0.6981269994175695,
             0.8190890835352128, 0.8527059887098624, 0.8604711817462826, 0.8627060376969976,
             0.8734213127600976)
        return a, b, c


class RLSRK:
    """Relaxation LSRK (src/odesolvers.jl:69-136): after the LSRK stages the step
    is scaled, q = q0 + gamma*(q - q0), with gamma the root near 1 of
    eta(q0 + gamma*k) - eta0 - gamma*deta (deta the entropy production of the
    stages), so that the entropy evolves as the semi-discretisation dictates.
    Stages, entropy reductions and the Newton iteration all run on the device
    (`atm_rlsrk_run`)."""

    scheme = 0
    nstages = 5

    def __init__(self, rhs, q: Field, dt: float, t0: float = 0.0):
        self.rhs = rhs
        self.dt = float(dt)
        self.time = float(t0)
        self.gamma = 1.0
        self.work = Field(3 * q.shape[0], q.shape[1], rhs.ctx).fill_(0.0)  # dq, q0, k

    def run(self, q: Field, nsteps: int, dt: float | None = None):
        dt = self.dt if dt is None else float(dt)
        if nsteps <= 0:
            return
        g = ctypes.c_double(self.gamma)
        self.rhs.ctx.check(_lib.lib.atm_rlsrk_run(self.rhs.h, self.scheme, q.ptr, self.work.ptr, dt,
                                                  self.time, int(nsteps), ctypes.byref(g)),
                           "atm_rlsrk_run")
        self.gamma = g.value
        for _ in range(nsteps):
            self.time += dt

    def dostep(self, q: Field, after_stage=None):
        if after_stage is not None:
            raise NotImplementedError("the device RLSRK step has no per-stage hook; use after_step")
        self.run(q, 1)


class RLSRK54(RLSRK):
    scheme = 0
    nstages = 5


class RLSRK144(RLSRK):
    scheme = 1
    nstages = 14


class BandedOperator:
    """Per-line LU factors of A = I + scale*L for a vertical linear DGSEM operator L
    (Bennu's batchedbandedmatrix + batchedbandedlu!, src/odesolvers.jl:264-268),
    built on the device by `atm_banded_create`; lines with bit-identical bands share
    one factor.  Several solvers with the same operator and scale may share it."""

    def __init__(self, linrhs, scale: float):
        self.linrhs = linrhs
        self.scale = float(scale)
        colelem, KV, ncols = linrhs.grid.columns()
        self._colelem = np.ascontiguousarray(colelem, dtype=np.int32)
        h = ctypes.c_void_p()
        linrhs.ctx.check(_lib.lib.atm_banded_create(
            linrhs.h, KV, self._colelem.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ncols,
            self.scale, ctypes.byref(h)), "atm_banded_create")
        self.h = h

    @property
    def nunique(self):
        return int(_lib.lib.atm_banded_nunique(self.h))

    def solve(self, x: Field, rhs: Field):
        """x = A \\ rhs (x may be rhs)."""
        self.linrhs.ctx.check(_lib.lib.atm_banded_solve(self.h, x.ptr, rhs.ptr), "atm_banded_solve")

    @property
    def kernel(self) -> str:
        """The solve kernel a solve uses ("mfma", "block", "reg", "lds" or "plain")."""
        k = _lib.lib.atm_banded_get_kernel(self.h)
        return {v: n for n, v in _lib.BAND_KERNELS.items()}[k]

    @kernel.setter
    def kernel(self, name: str):
        self.linrhs.ctx.check(_lib.lib.atm_banded_set_kernel(self.h, _lib.BAND_KERNELS[name]),
                              "atm_banded_set_kernel")

    @property
    def waves(self) -> int:
        """Waves per workgroup of the MFMA solve (4: 64 lines share each block's operators in
        LDS; 1: 16 lines); 0 when the MFMA solve cannot run these factors."""
        return int(_lib.lib.atm_banded_get_waves(self.h))

    @waves.setter
    def waves(self, w: int):
        """4, 1, or 0 for the largest that fits; the solution is the same bits for either."""
        self.linrhs.ctx.check(_lib.lib.atm_banded_set_waves(self.h, int(w)), "atm_banded_set_waves")

    def autotune(self, x: Field, rhs: Field, reps: int = 10):
        """Time every solve kernel that can run these factors (the MFMA solve in both of its
        workgroup shapes), keep the fastest; returns (name, {name: microseconds per solve}).
        x is overwritten."""
        ks = (ctypes.c_int32 * 8)()
        us = (ctypes.c_float * 8)()
        best = _lib.lib.atm_banded_autotune(self.h, x.ptr, rhs.ptr, int(reps), ks, us, 8)
        self.linrhs.ctx.check(best if best < 0 else 0, "atm_banded_autotune")
        names = {v: n for n, v in _lib.BAND_KERNELS.items()}
        times = {names[ks[i]]: float(us[i]) for i in range(5) if names.get(ks[i]) and us[i] > 0}
        return names[best], times

    def host(self):
        """The factors of every line, (nlines, n, 2bw+1), band storage."""
        n = ctypes.c_int32(0)
        bw = ctypes.c_int32(0)
        nl = ctypes.c_int64(0)
        _lib.lib.atm_banded_shape(self.h, ctypes.byref(nl), ctypes.byref(n), ctypes.byref(bw))
        out = np.empty((nl.value, n.value, 2 * bw.value + 1))
        self.linrhs.ctx.check(_lib.lib.atm_banded_download(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "atm_banded_download")
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib.atm_banded_destroy(self.h)
                self.h = None
        except Exception:
            pass


class ARK23:
    """3-stage IMEX ARK of Giraldo, Kelly and Constantinescu (src/odesolvers.jl:231-388).

    `rhs` (explicit, may be None) and `linrhs` (the vertical linear operator)
    are DGSEM operators; the linear part is solved per vertical node line with
    a banded LU of I - dt*(1 - 1/sqrt 2)*L built once at construction (or
    `banded`, a BandedOperator of that operator and scale, shared with other
    solvers).
    """

    def __init__(self, rhs, linrhs, q: Field, dt: float, t0: float = 0.0, paperversion=False,
                 split_rhs=True, banded: BandedOperator | None = None):
        if linrhs is None:
            raise NotImplementedError("ARK23 without a linear operator is plain explicit RK")
        self.rhs, self.linrhs = rhs, linrhs
        self.dt = self.dt_fac = float(dt)
        self.time = float(t0)
        self.paperversion, self.split_rhs = bool(paperversion), bool(split_rhs)
        scale = -self.dt_fac * (1 - 1 / math.sqrt(2))
        if banded is None:
            banded = BandedOperator(linrhs, scale)
        elif banded.linrhs is not linrhs or banded.scale != scale:
            raise ValueError("the banded operator was built for another operator or time step")
        self.banded = banded
        self.work = Field(9 * q.shape[0], q.shape[1], linrhs.ctx)

    @property
    def band(self):
        """The library handle of the banded factors."""
        return self.banded.h

    def run(self, q: Field, nsteps: int, dt: float | None = None):
        dt = self.dt if dt is None else float(dt)
        ctx = self.linrhs.ctx
        ctx.check(_lib.lib.atm_ark23_run(self.rhs.h if self.rhs is not None else None,
                                         self.linrhs.h, self.band, q.ptr, self.work.ptr, dt,
                                         self.time, int(nsteps), int(self.split_rhs),
                                         int(self.paperversion)), "atm_ark23_run")
        for _ in range(nsteps):
            self.time += dt

    def dostep(self, q: Field, after_stage=None):
        self.run(q, 1)

    def band_host(self):
        return self.banded.host()


def solve_(q: Field, timeend: float, solver, after_step=None, after_stage=None,
           adjust_final=True):
    """`solve!(q, timeend, solver; after_step, after_stage, adjust_final)`.

    Without hooks, the run of equal steps is launched in one call; the step
    count and the clipped final step follow the reference's floating-point
    time accumulation exactly.
    """
    step = 0
    if after_step is None and after_stage is None:
        # count the equal steps before the final one
        t, n = solver.time, 0
        while not (t + solver.dt >= timeend):
            t += solver.dt
            n += 1
        if n:
            solver.run(q, n)
        step = n
        step += 1
        if adjust_final:
            solver.dt = timeend - solver.time
        solver.dostep(q)
        return step
    while True:
        step += 1
        time = solver.time
        finalstep = False
        if time + solver.dt >= timeend:
            if adjust_final:
                solver.dt = timeend - time
            finalstep = True
        solver.dostep(q, after_stage)
        if after_step is not None:
            after_step(step, time, q)
        if finalstep:
            break
    return step
