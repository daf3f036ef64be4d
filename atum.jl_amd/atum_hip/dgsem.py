"""DGSEM operator: the drop-in for `DGSEM(...)` and its `rhs!` protocol.

`DGSEM(law=..., grid=..., surface_numericalflux=..., volume_form=...,
directions=..., auxstate=...)` mirrors src/dgsem.jl:38-56; calling the object
as `dg(dq, q, time, increment=True)` mirrors src/dgsem.jl:59-98 and runs the
fused volume + surface kernel of libatumhip.so on the device fields.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .fields import Field
from .laws import AbstractNumericalFlux, EntropyConservativeFlux


class WeakForm:
    """src/dgsem.jl:7."""

    def __repr__(self):
        return "WeakForm()"

    def __eq__(self, other):
        return isinstance(other, WeakForm)

    def __hash__(self):
        return hash("WeakForm")


class FluxDifferencingForm:
    """src/dgsem.jl:8-14.  The three reference kernel variants (:per_dir,
    :per_dir_symmetric, :naive; src/dgsem.jl:135-212) compute the same sum in
    different orders; in precision="reference" the device kernel follows the
    order `kernel_type` names (their last bits differ), and "naive" needs all
    directions, as in the reference.  precision="fast" evaluates its own
    pair-symmetric order and accepts only the default kernel_type."""

    def __init__(self, volume_numericalflux: AbstractNumericalFlux = EntropyConservativeFlux(),
                 kernel_type: str = "per_dir"):
        if kernel_type not in ("per_dir", "per_dir_symmetric", "naive"):
            raise ValueError(f"Unknown kernel type {kernel_type}")
        self.volume_numericalflux = volume_numericalflux
        self.kernel_type = kernel_type

    def __repr__(self):
        return f"FluxDifferencingForm({self.volume_numericalflux!r}, {self.kernel_type!r})"

    def __eq__(self, other):
        return (isinstance(other, FluxDifferencingForm)
                and other.volume_numericalflux == self.volume_numericalflux
                and other.kernel_type == self.kernel_type)

    def __hash__(self):
        return hash(("FD", self.volume_numericalflux, self.kernel_type))


class DGSEM:
    def __init__(self, *, law, grid, surface_numericalflux, volume_form=None, directions=None,
                 auxstate=None, precision="reference", geometry="host", ctx=None, element_order="id"):
        """precision: "reference" keeps the reference's summation order with every
        product and sum rounded; "fast" evaluates symmetric flux pairs once with FMA
        (same operator, round-off-level differences).
        geometry: "host" copies the grid's geometry arrays to the device; "device"
        has the library build them from the brick's vertex coordinates
        (atm_dg_create_brick), bit-identical, without forming them on the host.
        element_order: "id" (default) has the element kernels take the elements in
        element-id order, "morton" in Morton order of the brick (atm_dg_set_schedule);
        results are the same either way (Morton measured slower: the indirection adds
        a dependent load before each element's first load)."""
        if element_order not in ("morton", "id"):
            raise ValueError("element_order must be 'morton' or 'id'")
        self.ctx = ctx or _lib.context()
        self.law = law
        self.grid = grid
        self.volume_form = volume_form if volume_form is not None else WeakForm()
        self.surface_numericalflux = surface_numericalflux
        dim = grid.dim
        if law.dim != dim:
            raise ValueError("law and grid dimensions differ")
        self.directions = tuple(directions) if directions is not None else tuple(range(1, dim + 1))
        if not self.directions or any(d < 1 or d > dim for d in self.directions):
            raise ValueError("directions must be a non-empty subset of 1:dim")
        if geometry not in ("host", "device"):
            raise ValueError("geometry must be 'host' or 'device'")
        self.geometry = geometry
        self._geom = {}
        if auxstate is None:
            auxstate = law.auxiliary(grid.points)
        self.auxstate = None if auxstate is None else np.ascontiguousarray(auxstate, dtype=np.float64)
        naux = 0 if self.auxstate is None else self.auxstate.shape[0]
        prob = law.problem
        d = _lib.DgDesc()
        d.law = _lib.LAWS[law.law]
        d.dim, d.N, d.K = dim, grid.N, grid.K
        if isinstance(self.volume_form, WeakForm):
            d.volume_form, d.volume_flux = 0, 0
        else:
            d.volume_form = 1
            d.volume_flux = _lib.FLUX_IDS[self.volume_form.volume_numericalflux.name]
            d.fd_kernel = _lib.FD_KERNEL_IDS[self.volume_form.kernel_type]
        d.surface_flux = _lib.FLUX_IDS[surface_numericalflux.name]
        d.dirmask = sum(1 << (r - 1) for r in self.directions)
        d.bc = _lib.BC_IDS[prob.boundary]
        d.source = _lib.SOURCE_IDS[prob.source]
        d.pde_level_balance = int(getattr(law, "pde_level_balance", False))
        d.naux = naux
        d.precision = _lib.PRECISION_IDS[precision]
        self.precision = precision
        # what selects the element kernel (with K, the key of the workgroup-shape tuning)
        self.kernel_key = (law.law, dim, grid.N, d.volume_form, d.volume_flux, d.surface_flux, d.dirmask,
                           precision, d.fd_kernel)
        d.gamma = float(getattr(law, "gamma", 1.4))
        d.grav = float(getattr(law, "grav", 9.81))
        d.omega = float(prob.omega)
        u = getattr(law, "u", (0.0,) * dim)
        for i in range(3):
            d.u[i] = float(u[i]) if i < len(u) else 0.0
        keep = []

        def ptr(a, t=ctypes.c_double):
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(t))

        if prob.boundary == "prescribed":
            st = np.ascontiguousarray(prob.bc_states, dtype=np.float64)
            d.n_bc_states = st.shape[0]
            d.bc_states = ptr(st)
        d.D = ptr(np.ascontiguousarray(grid.cell.D).ravel(order="F"))
        if naux:
            d.aux = ptr(self.auxstate)
        h = ctypes.c_void_p()
        if geometry == "device":
            b = _lib.BrickDesc()
            for i in range(dim):
                b.Kd[i] = grid.Kd[i]
                b.periodic[i] = int(grid.periodic[i])
                b.coords[i] = ptr(np.asarray(grid.coords[i], dtype=np.float64))
            b.stacked = int(grid.ordering == "stacked")
            b.xi, b.w = ptr(grid.cell.points_1d), ptr(grid.cell.weights_1d)
            self.ctx.check(_lib.lib.atm_dg_create_brick(self.ctx.h, ctypes.byref(d), ctypes.byref(b),
                                                        ctypes.byref(h)), "atm_dg_create_brick")
        else:
            d.g, d.MJ, d.MJI = ptr(grid.g), ptr(grid.MJ), ptr(grid.MJI)
            d.faceMJ, d.normal = ptr(grid.faceMJ), ptr(grid.normal)
            d.faceix_m = ptr(grid.faceix_m, ctypes.c_int64)
            d.faceix_p = ptr(grid.faceix_p, ctypes.c_int64)
            d.bctag = ptr(grid.bctag, ctypes.c_int32)
            self.ctx.check(_lib.lib.atm_dg_create(self.ctx.h, ctypes.byref(d), ctypes.byref(h)),
                           "atm_dg_create")
        self.h = h
        self.Ns = _lib.lib.atm_dg_nstates(h)
        self.element_order = element_order
        if element_order == "morton" and grid.K > 1:
            self.set_schedule(grid.morton_order())

    def set_schedule(self, order=None):
        """Processing order of the elements (a permutation of the element ids; None:
        element-id order).  Results do not depend on it."""
        if order is None:
            self.ctx.check(_lib.lib.atm_dg_set_schedule(self.h, None, 0), "atm_dg_set_schedule")
            return
        o = np.ascontiguousarray(order, dtype=np.int32)
        self.ctx.check(_lib.lib.atm_dg_set_schedule(self.h, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                    int(o.size)), "atm_dg_set_schedule")

    # src/dgsem.jl:41-51: MJ, MJI, faceMJ (downloaded once when built on the device)
    @property
    def MJ(self):
        return self.grid.MJ if self.geometry == "host" else self.geometry_array("MJ")

    @property
    def MJI(self):
        return self.grid.MJI if self.geometry == "host" else self.geometry_array("MJI")

    @property
    def faceMJ(self):
        return self.grid.faceMJ if self.geometry == "host" else self.geometry_array("faceMJ")

    def geometry_array(self, name):
        """A geometry array as held by the library (atm_dg_geometry): points, g, MJ,
        MJI, faceMJ, normal, faceix_p or bctag."""
        if name not in self._geom:
            which, dtype, rows, per = _lib.GEOMETRY[name]
            dim = self.grid.dim
            n = (self.grid.Ntot if per == "node" else
                 self.grid.K * self.grid.Nfp if per == "face" else 2 * dim * self.grid.K)
            out = np.empty({"1": 1, "dim": dim, "dim2": dim * dim}[rows] * n, dtype=dtype)
            self.ctx.check(_lib.lib.atm_dg_geometry(self.h, which, out.ctypes.data), "atm_dg_geometry")
            self._geom[name] = out.reshape(-1, n) if rows != "1" else out
        return self._geom[name]

    def __call__(self, dq: Field, q: Field, time: float = 0.0, increment: bool = True):
        self.ctx.check(_lib.lib.atm_dg_tendency(self.h, dq.ptr, q.ptr, float(time), int(increment)),
                       "atm_dg_tendency")
        return dq

    def lsrk_stage(self, q_in: Field, q_out: Field, dq: Field, a: float, bdt: float, time=0.0):
        self.ctx.check(_lib.lib.atm_lsrk_stage(self.h, q_in.ptr, q_out.ptr, dq.ptr, float(a),
                                               float(bdt), float(time)), "atm_lsrk_stage")

    @property
    def epw(self):
        """Elements per workgroup of the element kernel (results do not depend on it)."""
        return int(_lib.lib.atm_dg_get_epw(self.h))

    @epw.setter
    def epw(self, value):
        self.ctx.check(_lib.lib.atm_dg_set_epw(self.h, int(value)), "atm_dg_set_epw")

    def epw_candidates(self):
        """The workgroup shapes that fit (whole waves, <= 512 threads or 768 for the FAST
        kernel's 3-D N = 6 elements with the Roe or Rusanov surface flux, the device's LDS)."""
        out = (ctypes.c_int32 * 8)()
        n = _lib.lib.atm_dg_epw_candidates(self.h, out, 8)
        return [int(out[i]) for i in range(n)]

    def fieldarray(self, values=None):
        f = Field(self.Ns, self.grid.Ntot, self.ctx)
        if values is None:
            f.fill_(0.0)
        else:
            f.upload(values)
        return f

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib.atm_dg_destroy(self.h)
                self.h = None
        except Exception:
            pass
