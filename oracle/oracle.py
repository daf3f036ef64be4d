"""CPU ORACLE — test infrastructure only.

ctypes front end of ``oracle/build/liboracle.so``, the FP64 CPU restatement of
Atum.jl's DGSEM tendency and time-step path (see ``atum_oracle.cpp`` for the
reference file:line each part follows).  Only the tests, the smoke check and
the bench's CPU-baseline leg import this module; it is what results are compared with, never the
thing measured or shipped.

Arrays follow the reference data model: one FP64 array per state component,
each ``(Np, K)`` column major, stacked here as a C-contiguous ``(Ns, Np*K)``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ATUM_ORACLE_SO names another build of the same sources (the mutation checks
# of the flux tests load deliberately broken copies this way)
_SO = os.environ.get("ATUM_ORACLE_SO") or os.path.join(_HERE, "build", "liboracle.so")

LAWS = {"advection": 0, "euler": 1, "euler_gravity": 2,
        "linear_euler_gravity": 3, "shallow_water": 4}
FLUXES = {"central": 0, "rusanov": 1, "ec": 2, "kg": 3, "roe": 4, "matrix": 5}
BCS = {"copy": 0, "reflect": 1, "zero_inflow": 2, "prescribed": 3}
SOURCES = {None: 0, "coriolis": 1}
FD_KERNELS = {"per_dir": 0, "per_dir_symmetric": 1, "naive": 2}

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


class _Geom(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("Nq", ctypes.c_int32), ("K", ctypes.c_int32),
                ("naux", ctypes.c_int32), ("D", _dp), ("g", _dp), ("MJ", _dp), ("MJI", _dp),
                ("faceMJ", _dp), ("normal", _dp), ("aux", _dp), ("faceix_m", _i64p),
                ("faceix_p", _i64p), ("bctag", _i32p)]


class _DG(ctypes.Structure):
    _fields_ = [("geom", _Geom), ("law", ctypes.c_int32), ("volume_form", ctypes.c_int32),
                ("fd_kernel", ctypes.c_int32), ("vflux", ctypes.c_int32),
                ("sflux", ctypes.c_int32), ("dirmask", ctypes.c_int32), ("bc", ctypes.c_int32),
                ("source", ctypes.c_int32), ("pde_level_balance", ctypes.c_int32),
                ("n_bc_states", ctypes.c_int32), ("gamma", ctypes.c_double),
                ("grav", ctypes.c_double), ("omega", ctypes.c_double),
                ("u", ctypes.c_double * 3), ("bc_states", _dp)]


_lib = None


def lib():
    """Load the oracle library, building it with its Makefile when absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = ctypes.CDLL(_SO)
        L.orc_brickgrid.restype = ctypes.c_int
        L.orc_tendency.restype = ctypes.c_int
        L.orc_check.restype = ctypes.c_int
        L.orc_nstates.restype = ctypes.c_int
        L.orc_weightednorm.restype = ctypes.c_double
        L.orc_weightednorm.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp]
        L.orc_entropyintegral.restype = ctypes.c_double
        L.orc_entropyproduct.restype = ctypes.c_double
        L.orc_set_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def set_threads(n: int) -> int:
    return lib().orc_set_threads(int(n))


def lgl(N: int):
    """LGL nodes, weights and derivative matrix D (D[r, c] = l_c'(xi_r))."""
    Nq = N + 1
    xi = np.zeros(Nq)
    w = np.zeros(Nq)
    D = np.zeros(Nq * Nq)
    lib().orc_lgl(ctypes.c_int(N), _p(xi), _p(w), _p(D))
    return xi, w, D.reshape(Nq, Nq, order="F")


class Grid:
    """The oracle's own brick grid (reconstruction of Bennu's ``brickgrid``)."""

    def __init__(self, N, coords, periodic, stacked=False):
        self.dim = dim = len(coords)
        self.N = N
        self.Nq = Nq = N + 1
        self.Kd = [len(c) - 1 for c in coords]
        self.K = K = int(np.prod(self.Kd))
        self.Np = Np = Nq ** dim
        self.Nfp = Nfp = 2 * dim * Nq ** (dim - 1)
        self.Ntot = Ntot = Np * K
        self.stacked = bool(stacked)
        self.periodic = tuple(bool(p) for p in periodic)
        self.coords = [np.asarray(c, dtype=np.float64) for c in coords]
        cat = np.ascontiguousarray(np.concatenate(self.coords))
        Kd = np.array(self.Kd, dtype=np.int32)
        per = np.array(self.periodic, dtype=np.int32)
        self.points = np.zeros((dim, Ntot))
        self.g = np.zeros((dim * dim, Ntot))
        self.MJ = np.zeros(Ntot)
        self.MJI = np.zeros(Ntot)
        self.faceMJ = np.zeros(Nfp * K)
        self.normal = np.zeros((dim, Nfp * K))
        self.faceix_m = np.zeros(Nfp * K, dtype=np.int64)
        self.faceix_p = np.zeros(Nfp * K, dtype=np.int64)
        self.bctag = np.zeros(2 * dim * K, dtype=np.int32)
        self.D = np.zeros(Nq * Nq)
        rc = lib().orc_brickgrid(
            ctypes.c_int(dim), ctypes.c_int(N), _p(Kd, _i32p), _p(cat), _p(per, _i32p),
            ctypes.c_int(int(stacked)), _p(self.points), _p(self.g), _p(self.MJ), _p(self.MJI),
            _p(self.faceMJ), _p(self.normal), _p(self.faceix_m, _i64p), _p(self.faceix_p, _i64p),
            _p(self.bctag, _i32p), _p(self.D))
        assert rc == 0
        self.Dmat = self.D.reshape(Nq, Nq, order="F")

    def colelem(self):
        """Element ids of each vertical column (last dimension), bottom to top."""
        KV = self.Kd[-1]
        if self.stacked:
            return np.arange(self.K, dtype=np.int32).reshape(-1, KV).ravel(), KV, self.K // KV
        H = self.K // KV
        return (np.arange(self.K, dtype=np.int32).reshape(KV, H).T.ravel()), KV, H

    def min_node_distance(self, dims=None):
        xi, _, _ = lgl(self.N)
        gap = np.min(np.diff(xi))
        dims = range(self.dim) if dims is None else dims
        return min(np.min(np.diff(self.coords[d])) / 2 * gap for d in dims)


class DG:
    """A DGSEM operator evaluated by the oracle (src/dgsem.jl:38-98)."""

    def __init__(self, law, grid, *, surface_flux="rusanov", volume_form="weak", volume_flux=None,
                 fd_kernel="per_dir", directions=None, gamma=1.4, grav=9.81, u=(1.0, 1.0, 1.0),
                 bc="copy", bc_states=None, source=None, omega=0.0, aux=None,
                 pde_level_balance=False):
        self.grid = grid
        self.law = law
        dim = grid.dim
        directions = tuple(range(1, dim + 1)) if directions is None else tuple(directions)
        self.directions = directions
        self.aux = (np.zeros((0, grid.Ntot)) if aux is None
                    else np.ascontiguousarray(aux, dtype=np.float64))
        self.bc_states = (np.zeros(1) if bc_states is None
                          else np.ascontiguousarray(bc_states, dtype=np.float64).ravel())
        nbc = 0 if bc_states is None else len(bc_states)
        s = _DG()
        G = s.geom
        G.dim, G.Nq, G.K, G.naux = dim, grid.Nq, grid.K, self.aux.shape[0]
        G.D, G.g, G.MJ, G.MJI = _p(grid.D), _p(grid.g), _p(grid.MJ), _p(grid.MJI)
        G.faceMJ, G.normal = _p(grid.faceMJ), _p(grid.normal)
        G.aux = _p(self.aux)
        G.faceix_m, G.faceix_p = _p(grid.faceix_m, _i64p), _p(grid.faceix_p, _i64p)
        G.bctag = _p(grid.bctag, _i32p)
        s.law = LAWS[law]
        s.volume_form = 0 if volume_form == "weak" else 1
        s.fd_kernel = FD_KERNELS[fd_kernel]
        s.vflux = FLUXES[volume_flux] if volume_flux else 0
        s.sflux = FLUXES[surface_flux]
        s.dirmask = sum(1 << (d - 1) for d in directions)
        s.bc = BCS[bc]
        s.source = SOURCES[source]
        s.pde_level_balance = int(pde_level_balance)
        s.n_bc_states = nbc
        s.gamma, s.grav, s.omega = float(gamma), float(grav), float(omega)
        for d in range(3):
            s.u[d] = float(u[d]) if d < len(u) else 0.0
        s.bc_states = _p(self.bc_states)
        self._s = s
        rc = lib().orc_check(ctypes.byref(s))
        if rc != 0:
            raise ValueError(f"oracle: unsupported DG configuration (code {rc})")
        self.Ns = lib().orc_nstates(ctypes.byref(s))

    @property
    def MJ(self):
        return self.grid.MJ

    def new_state(self):
        return np.zeros((self.Ns, self.grid.Ntot))

    def tendency(self, q, dq=None, increment=False):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if dq is None:
            dq = np.zeros_like(q)
        rc = lib().orc_tendency(ctypes.byref(self._s), _p(q), _p(dq), ctypes.c_int(int(increment)))
        assert rc == 0
        return dq

    def lsrk_stage(self, q, dq, a, bdt):
        rc = lib().orc_lsrk_stage(ctypes.byref(self._s), _p(q), _p(dq), ctypes.c_double(a),
                                  ctypes.c_double(bdt))
        assert rc == 0

    def solve_lsrk(self, q, dt, timeend, scheme="LSRK54", t0=0.0, adjust_final=True, max_steps=0):
        n = ctypes.c_int64(0)
        t = ctypes.c_double(0)
        rc = lib().orc_solve_lsrk(ctypes.byref(self._s), _p(q), ctypes.c_double(dt),
                                  ctypes.c_double(t0), ctypes.c_double(timeend),
                                  ctypes.c_int(0 if scheme == "LSRK54" else 1),
                                  ctypes.c_int(int(adjust_final)), ctypes.c_int64(max_steps),
                                  ctypes.byref(n), ctypes.byref(t))
        assert rc == 0
        return n.value, t.value

    def lsrk_steps(self, q, dt, nsteps, scheme="LSRK54", dq=None):
        if dq is None:
            dq = np.zeros_like(q)
        rc = lib().orc_lsrk_steps(ctypes.byref(self._s), _p(q), _p(dq), ctypes.c_double(dt),
                                  ctypes.c_int64(nsteps),
                                  ctypes.c_int(0 if scheme == "LSRK54" else 1))
        assert rc == 0
        return dq

    def solve_rlsrk(self, q, dt, timeend, scheme="LSRK54", t0=0.0, adjust_final=True):
        n = ctypes.c_int64(0)
        gam = ctypes.c_double(0)
        rc = lib().orc_solve_rlsrk(ctypes.byref(self._s), _p(q), ctypes.c_double(dt),
                                   ctypes.c_double(t0), ctypes.c_double(timeend),
                                   ctypes.c_int(0 if scheme == "LSRK54" else 1),
                                   ctypes.c_int(int(adjust_final)), ctypes.byref(n),
                                   ctypes.byref(gam))
        assert rc == 0
        return n.value, gam.value

    def weightednorm(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        return lib().orc_weightednorm(q.shape[0], q.shape[1], _p(self.grid.MJ), _p(q))

    def entropyintegral(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().orc_entropyintegral.argtypes = [ctypes.c_void_p, _dp]
        return lib().orc_entropyintegral(ctypes.byref(self._s), _p(q))

    def entropyproduct(self, p, q):
        p = np.ascontiguousarray(p, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().orc_entropyproduct.argtypes = [ctypes.c_void_p, _dp, _dp]
        return lib().orc_entropyproduct(ctypes.byref(self._s), _p(p), _p(q))

    # --- banded vertical operator ------------------------------------------
    def band_shape(self):
        colelem, KV, ncols = self.grid.colelem()
        n = ctypes.c_int32(0)
        bw = ctypes.c_int32(0)
        nl = ctypes.c_int64(0)
        lib().orc_band_shape(ctypes.byref(self._s), ctypes.c_int(KV), ctypes.c_int(ncols),
                             ctypes.byref(n), ctypes.byref(bw), ctypes.byref(nl))
        return n.value, bw.value, nl.value

    def banded_assemble(self):
        """Band of the linear operator, shape (nlines, n, 2*bw+1), [line, row, bw+col-row]."""
        colelem, KV, ncols = self.grid.colelem()
        n, bw, nl = self.band_shape()
        band = np.zeros((nl, n, 2 * bw + 1))
        rc = lib().orc_banded_assemble(ctypes.byref(self._s), ctypes.c_int(KV),
                                       _p(colelem, _i32p), ctypes.c_int(ncols), _p(band))
        assert rc == 0
        return band

    def banded_factor(self, band, scale):
        colelem, KV, ncols = self.grid.colelem()
        lib().orc_banded_factor(ctypes.byref(self._s), ctypes.c_int(KV), ctypes.c_int(ncols),
                                _p(band), ctypes.c_double(scale))
        return band

    def banded_solve(self, band, b):
        colelem, KV, ncols = self.grid.colelem()
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().orc_banded_solve(ctypes.byref(self._s), ctypes.c_int(KV), _p(colelem, _i32p),
                               ctypes.c_int(ncols), _p(band), _p(x), _p(b))
        return x


def solve_ark23(ex, lin, q, dt, timeend, *, split_rhs=True, paperversion=False, adjust_final=True,
                t0=0.0, max_steps=0):
    """solve! with ARK23 (src/odesolvers.jl:231-388); ``ex`` may be None."""
    colelem, KV, ncols = lin.grid.colelem()
    n = ctypes.c_int64(0)
    t = ctypes.c_double(0)
    rc = lib().orc_solve_ark23(
        ctypes.byref(ex._s) if ex is not None else None, ctypes.byref(lin._s), _p(q),
        ctypes.c_double(dt), ctypes.c_double(t0), ctypes.c_double(timeend),
        ctypes.c_int(int(split_rhs)), ctypes.c_int(int(paperversion)),
        ctypes.c_int(int(adjust_final)), ctypes.c_int(KV), _p(colelem, _i32p),
        ctypes.c_int(ncols), ctypes.c_int64(max_steps), ctypes.byref(n), ctypes.byref(t))
    assert rc == 0
    return n.value, t.value


_POINT_WHICH = {"flux": 0, "twopoint": 1, "surface": 2, "entropyvariables": 3}


def point_flux(law, dim, which, q1, a1=None, q2=None, a2=None, *, flux="central", n=None, gamma=1.4,
               grav=9.81):
    """One flux of ``law`` at one point (pair), as the oracle's tendency evaluates it.

    which: "flux" F(q1) -> (dim, Ns); "twopoint" -> (dim, Ns); "surface" (normal n)
    -> (Ns,); "entropyvariables" of q1 -> (Ns,).
    """
    s = _DG()
    s.geom.dim, s.geom.Nq, s.geom.K = dim, 2, 0
    q1 = np.ascontiguousarray(q1, dtype=np.float64)
    Ns = q1.shape[0]
    a1 = np.ascontiguousarray(np.zeros(1) if a1 is None else a1, dtype=np.float64)
    q2 = np.ascontiguousarray(q1 if q2 is None else q2, dtype=np.float64)
    a2 = np.ascontiguousarray(a1 if a2 is None else a2, dtype=np.float64)
    n = np.ascontiguousarray(np.zeros(dim) if n is None else n, dtype=np.float64)
    s.geom.naux = a1.shape[0]
    s.law = LAWS[law]
    s.gamma, s.grav = float(gamma), float(grav)
    w = _POINT_WHICH[which]
    out = np.zeros(dim * Ns if w in (0, 1) else Ns)
    L = lib()
    L.orc_pointflux.restype = ctypes.c_int
    rc = L.orc_pointflux(ctypes.byref(s), ctypes.c_int(w), ctypes.c_int(FLUXES[flux]), _p(n), _p(q1), _p(a1),
                         _p(q2), _p(a2), _p(out))
    if rc != 0:
        raise ValueError(f"oracle: {flux} {which} unsupported for {law} (code {rc})")
    return out.reshape(dim, Ns) if w in (0, 1) else out
