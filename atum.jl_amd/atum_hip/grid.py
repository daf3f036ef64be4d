"""Reference cell and brick grid (the host-side stand-in for Bennu).

Bennu (an un-vendored dependency of the reference, compat 0.3) provides the
LGL reference cell, `brickgrid`, the metric terms and the face tables that
`DGSEM` consumes (src/dgsem.jl:38-56, :74-95).  Its conventions are
reconstructed from the reference's call sites:

* node ijk = i + Nq*(j + Nq*k); `D[n, i]` = derivative of the i-th Lagrange
  polynomial at node n (src/dgsem.jl:281, :546-547);
* element order: Cartesian (first dimension fastest), or stacked (the last,
  vertical dimension fastest, so that each column is contiguous:
  test/advection/imex_wave_2d.jl:46-47);
* faces 2r+side with side 0 at xi_r = -1 (src/dgsem.jl:706-707); face points
  enumerate the other node indices with the lower dimension fastest;
* a non-periodic box face gets boundary tag 2r+side+1 (tag 1 = the x = x_min
  face, experiment/euler/sod.jl:20-23);
* affine metrics g = d xi / d x = diag(2/h), J = prod(h/2); MJ = w_i w_j w_k J,
  MJI = 1/MJ, faceMJ = (face weights) * (face Jacobian).
"""
from __future__ import annotations

import numpy as np


def _legendre(x, N):
    p0 = np.ones_like(x)
    if N == 0:
        return p0
    p1 = x.copy()
    for k in range(2, N + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1


def lgl(N: int):
    """LGL nodes, weights and the derivative matrix D[r, c] = l_c'(xi_r)."""
    if N < 1:
        raise ValueError("N >= 1 required")
    ld = np.longdouble
    x = np.zeros(N + 1, dtype=ld)
    x[0], x[N] = -1, 1
    if N > 1:
        c = np.zeros(N + 1)
        c[N] = 1
        r = np.sort(np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c)).real)
        xi = r.astype(ld)
        for _ in range(6):  # Newton on P_N' with P_N'' from the Legendre ODE
            P = _legendre(xi, N)
            Pm = _legendre(xi, N - 1)
            dP = N * (Pm - xi * P) / (1 - xi * xi)
            d2P = (2 * xi * dP - N * (N + 1) * P) / (1 - xi * xi)
            xi = xi - dP / d2P
        x[1:N] = xi
    PN = _legendre(x, N)
    PN[0], PN[N] = (-1) ** N, 1
    w = 2 / (N * (N + 1) * PN * PN)
    D = np.zeros((N + 1, N + 1), dtype=ld)
    for r in range(N + 1):
        for c in range(N + 1):
            if r != c:
                D[r, c] = PN[r] / (PN[c] * (x[r] - x[c]))
    D[0, 0] = -ld(N * (N + 1)) / 4
    D[N, N] = ld(N * (N + 1)) / 4
    return x.astype(np.float64), w.astype(np.float64), D.astype(np.float64)


class LobattoCell:
    """Tensor-product LGL cell with N+1 points per direction."""

    def __init__(self, N: int, dim: int):
        self.N, self.Nq, self.dim = N, N + 1, dim
        self.points_1d, self.weights_1d, self.D = lgl(N)
        self.Np = self.Nq ** dim

    def __repr__(self):
        return f"LobattoCell(N={self.N}, dim={self.dim})"


class Grid:
    """A brick of affine hexahedral (quad, interval) elements with face tables.

    With lazy=True nothing is computed up front: the node coordinates are
    formed on first use of `points`, the geometry and face tables on first use
    of any of them (an operator built with geometry="device" never needs them
    on the host)."""

    _GEOMETRY = ("D", "elem_index", "MJ", "MJI", "g", "faceix_m", "faceix_p", "faceMJ", "normal", "bctag")

    def __init__(self, cell: LobattoCell, coordinates, periodic, ordering="cartesian", lazy=False):
        dim = cell.dim
        if len(coordinates) != dim or len(periodic) != dim:
            raise ValueError("one coordinate array and periodic flag per dimension")
        self.cell = cell
        self.dim, self.N, self.Nq, self.Np = dim, cell.N, cell.Nq, cell.Np
        self.coords = [np.asarray(c, dtype=np.float64) for c in coordinates]
        self.Kd = [len(c) - 1 for c in self.coords]
        if min(self.Kd) < 1:
            raise ValueError("at least one element per dimension")
        self.K = int(np.prod(self.Kd))
        self.periodic = tuple(bool(p) for p in periodic)
        if ordering not in ("cartesian", "stacked"):
            raise ValueError("ordering must be 'cartesian' or 'stacked'")
        self.ordering = ordering
        self.Nfpf = self.Np // self.Nq
        self.Nfp = 2 * dim * self.Nfpf
        self.Ntot = self.Np * self.K
        if not lazy:
            self._build()

    def __getattr__(self, name):
        # only called for attributes not yet set: build them on first use
        if name == "points":
            self._build_points()
            return self.__dict__["points"]
        if name in Grid._GEOMETRY:
            self._build()
            return self.__dict__[name]
        raise AttributeError(name)

    # element id of element multi-indices (arrays)
    def elem_id(self, em):
        K = self.Kd
        if self.ordering == "cartesian":
            e = np.zeros_like(em[0])
            for d in reversed(range(self.dim)):
                e = e * K[d] + em[d]
            return e
        V = self.dim - 1
        h = np.zeros_like(em[0])
        stride = 1
        for d in range(self.dim):
            if d != V:
                h = h + em[d] * stride
                stride *= K[d]
        return em[V] + K[V] * h

    def _elements(self):
        """Element multi-indices in element-id order, (dim, K)."""
        if "_em" not in self.__dict__:
            dim, K = self.dim, self.K
            grids = np.meshgrid(*[np.arange(k) for k in self.Kd], indexing="ij")
            em_all = [g.ravel() for g in grids]
            eid = self.elem_id(em_all)
            em = [np.empty(K, dtype=np.int64) for _ in range(dim)]
            for d in range(dim):
                em[d][eid] = em_all[d]
            self._em = em
        return self._em

    def _build_points(self):
        dim, Nq = self.dim, self.Nq
        xi = self.cell.points_1d
        em = self._elements()
        nidx = np.indices((Nq,) * dim).reshape(dim, -1)[::-1]  # nidx[d][ijk], first dim fastest
        points = np.empty((dim, self.Ntot))
        for d in range(dim):
            lo, hi = self.coords[d][em[d]], self.coords[d][em[d] + 1]
            t = xi[nidx[d]]
            x = (1 - t)[:, None] / 2 * lo[None, :] + (1 + t)[:, None] / 2 * hi[None, :]
            points[d] = x.T.ravel()
        self.points = points

    def _build(self):
        dim, Nq, Np, K = self.dim, self.Nq, self.Np, self.K
        xi, w, D = self.cell.points_1d, self.cell.weights_1d, self.cell.D
        self.D = np.ascontiguousarray(D)
        em = self._elements()
        self.elem_index = np.stack(em)  # (dim, K)
        # node multi-indices within an element
        nidx = np.indices((Nq,) * dim).reshape(dim, -1)[::-1]  # nidx[d][ijk], first dim fastest
        lo = np.stack([self.coords[d][em[d]] for d in range(dim)])       # (dim, K)
        hi = np.stack([self.coords[d][em[d] + 1] for d in range(dim)])
        h = hi - lo
        if "points" not in self.__dict__:
            self._build_points()
        J = np.prod(h / 2, axis=0)
        wp = np.ones(Np)
        for d in range(dim):
            wp = wp * w[nidx[d]]
        self.MJ = (wp[None, :] * J[:, None]).ravel()
        self.MJI = 1 / self.MJ
        self.g = np.zeros((dim * dim, self.Ntot))
        for a in range(dim):
            self.g[a + dim * a] = np.repeat(2 / h[a], Np)
        # face tables
        Nfpf, Nfp = self.Nfpf, self.Nfp
        self.faceix_m = np.empty((K, Nfp), dtype=np.int64)
        self.faceix_p = np.empty((K, Nfp), dtype=np.int64)
        self.faceMJ = np.empty((K, Nfp))
        self.normal = np.zeros((dim, K, Nfp))
        self.bctag = np.zeros((K, 2 * dim), dtype=np.int32)
        ar = np.arange(K, dtype=np.int64)
        for r in range(dim):
            others = [d for d in range(dim) if d != r]
            fidx = np.indices((Nq,) * (dim - 1)).reshape(dim - 1, -1)[::-1] if dim > 1 else np.zeros((0, 1), int)
            for side in range(2):
                face = 2 * r + side
                sl = slice(face * Nfpf, (face + 1) * Nfpf)
                im = np.zeros((dim, Nfpf), dtype=np.int64)
                for t, d in enumerate(others):
                    im[d] = fidx[t]
                ip = im.copy()
                im[r] = 0 if side == 0 else Nq - 1
                ip[r] = Nq - 1 if side == 0 else 0
                mult = Nq ** np.arange(dim)
                nm = (im * mult[:, None]).sum(0)
                npn = (ip * mult[:, None]).sum(0)
                nb = [em[d].copy() for d in range(dim)]
                nb[r] = nb[r] + (-1 if side == 0 else 1)
                out = (nb[r] < 0) | (nb[r] >= self.Kd[r])
                if self.periodic[r]:
                    nb[r] = np.mod(nb[r], self.Kd[r])
                    bnd = np.zeros(K, dtype=bool)
                else:
                    bnd = out
                    nb[r] = np.where(out, em[r], nb[r])
                en = self.elem_id(nb)
                self.faceix_m[:, sl] = ar[:, None] * Np + nm[None, :]
                self.faceix_p[:, sl] = np.where(bnd[:, None], ar[:, None] * Np + nm[None, :],
                                                en[:, None] * Np + npn[None, :])
                wf = np.ones(Nfpf)
                fJ = np.ones(K)
                for d in others:
                    wf = wf * w[im[d]]
                    fJ = fJ * (h[d] / 2)
                self.faceMJ[:, sl] = wf[None, :] * fJ[:, None]
                self.normal[r, :, sl] = -1.0 if side == 0 else 1.0
                self.bctag[:, face] = np.where(bnd, face + 1, 0)
        self.faceix_m = self.faceix_m.ravel()
        self.faceix_p = self.faceix_p.ravel()
        self.faceMJ = self.faceMJ.ravel()
        self.normal = self.normal.reshape(dim, -1)
        self.bctag = self.bctag.ravel()

    def morton_order(self):
        """Element ids in Morton (Z-curve) order of their brick multi-indices: the
        processing order that keeps neighbouring elements close in time."""
        em = self._elements()
        key = np.zeros(self.K, dtype=np.int64)
        bits = max(int(k - 1).bit_length() for k in self.Kd) if self.K > 0 else 0
        for b in range(bits):
            for d in range(self.dim):
                key |= ((em[d] >> b) & 1) << (self.dim * b + d)
        return np.argsort(key, kind="stable").astype(np.int32)

    def columns(self):
        """(colelem, KV, ncols): elements of each vertical column, bottom to top."""
        KV = self.Kd[-1]
        em = np.stack(self._elements())
        ncols = self.K // KV
        col = np.zeros(self.K, dtype=np.int64)
        stride = 1
        for d in range(self.dim - 1):
            col = col + em[d] * stride
            stride *= self.Kd[d]
        colelem = np.empty(self.K, dtype=np.int32)
        colelem[col * KV + em[-1]] = np.arange(self.K, dtype=np.int32)
        return colelem, KV, ncols

    def min_node_distance(self, dims=None):
        gap = np.min(np.diff(self.cell.points_1d))
        dims = range(self.dim) if dims is None else dims
        return min(np.min(np.diff(self.coords[d])) / 2 * gap for d in dims)


def brickgrid(cell: LobattoCell, coordinates, periodic=None, ordering="cartesian", lazy=False) -> Grid:
    """Bennu-style `brickgrid(cell, coordinates; periodic, ordering)`."""
    periodic = periodic if periodic is not None else (False,) * cell.dim
    return Grid(cell, coordinates, periodic, ordering, lazy=lazy)
