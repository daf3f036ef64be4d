"""VTK output of DGSEM fields (what the reference's experiments write with
WriteVTK: `vtk_grid(file, grid)` plus the state interpolated to equally spaced
points, `Bennu.toequallyspaced(cell)`, experiment/euler_gravity/risingbubble.jl:81-95).

Each element's nodal values are interpolated from the LGL nodes to N+1 equally
spaced points per direction (exact for the element's polynomials), and every
element is written as N^d linear sub-cells (VTK_LINE / VTK_QUAD / VTK_HEXAHEDRON)
of an XML UnstructuredGrid (.vtu, ASCII); a .pvd collection lists the files of a
run by time.  Host-side NumPy: the snapshot stream calls it on its writer thread.
"""
from __future__ import annotations

import os

import numpy as np

from .grid import lgl

_CELL_TYPE = {1: 3, 2: 9, 3: 12}  # VTK_LINE, VTK_QUAD, VTK_HEXAHEDRON


def equispaced_interpolation(N: int) -> np.ndarray:
    """(N+1, N+1) matrix P with P[i, j] = l_j(t_i): LGL Lagrange basis at equally
    spaced points t_i = -1 + 2i/N."""
    x = lgl(N)[0]
    t = np.linspace(-1.0, 1.0, N + 1)
    P = np.ones((N + 1, N + 1))
    for j in range(N + 1):
        for m in range(N + 1):
            if m != j:
                P[:, j] *= (t - x[m]) / (x[j] - x[m])
    return P


def to_equispaced(values: np.ndarray, N: int, dim: int) -> np.ndarray:
    """Nodal values (Nq^d, K) (node i + Nq(j + Nq k)) to the equally spaced points."""
    Nq = N + 1
    K = values.shape[-1]
    P = equispaced_interpolation(N)
    v = values.reshape((Nq,) * dim + (K,), order="F")
    for d in range(dim):
        v = np.moveaxis(np.tensordot(P, v, axes=([1], [d])), 0, d)
    return v.reshape(Nq ** dim, K, order="F")


def _subcells(N: int, dim: int) -> np.ndarray:
    """Corner indices (into one element's Nq^d points) of its N^d linear sub-cells,
    in VTK corner order."""
    Nq = N + 1

    def idx(i, j=0, k=0):
        return i + Nq * (j + Nq * k)

    cells = []
    if dim == 1:
        for i in range(N):
            cells.append([idx(i), idx(i + 1)])
    elif dim == 2:
        for j in range(N):
            for i in range(N):
                cells.append([idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)])
    else:
        for k in range(N):
            for j in range(N):
                for i in range(N):
                    cells.append([idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                                  idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1),
                                  idx(i, j + 1, k + 1)])
    return np.asarray(cells, dtype=np.int64)


def _ascii(a) -> str:
    return " ".join(repr(float(x)) for x in np.ravel(a)) if np.asarray(a).dtype.kind == "f" else \
        " ".join(str(int(x)) for x in np.ravel(a))


def write_vtu(path: str, points: np.ndarray, fields: dict, N: int, dim: int):
    """Write one .vtu: `points` (dim, Nq^d*K) node coordinates, `fields` name -> (Nq^d*K,)
    nodal values, both in the DGSEM node order."""
    Nq = N + 1
    Np = Nq ** dim
    K = points.shape[-1] // Np
    xyz = np.zeros((3, Np * K))
    for d in range(dim):
        xyz[d] = to_equispaced(np.asarray(points[d]).reshape(Np, K, order="F"), N, dim).reshape(-1, order="F")
    sub = _subcells(N, dim)
    conn = (sub[None, :, :] + (np.arange(K) * Np)[:, None, None]).reshape(-1, sub.shape[1])
    ncell = conn.shape[0]
    offsets = np.arange(1, ncell + 1) * sub.shape[1]
    types = np.full(ncell, _CELL_TYPE[dim])
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        f.write(f'<UnstructuredGrid>\n<Piece NumberOfPoints="{Np * K}" NumberOfCells="{ncell}">\n')
        f.write('<PointData>\n')
        for name, val in fields.items():
            v = to_equispaced(np.asarray(val, dtype=np.float64).reshape(Np, K, order="F"), N, dim)
            f.write(f'<DataArray type="Float64" Name="{name}" format="ascii">{_ascii(v.reshape(-1, order="F"))}'
                    '</DataArray>\n')
        f.write('</PointData>\n<Points>\n')
        f.write(f'<DataArray type="Float64" NumberOfComponents="3" format="ascii">{_ascii(xyz.T)}</DataArray>\n')
        f.write('</Points>\n<Cells>\n')
        f.write(f'<DataArray type="Int64" Name="connectivity" format="ascii">{_ascii(conn)}</DataArray>\n')
        f.write(f'<DataArray type="Int64" Name="offsets" format="ascii">{_ascii(offsets)}</DataArray>\n')
        f.write(f'<DataArray type="UInt8" Name="types" format="ascii">{_ascii(types)}</DataArray>\n')
        f.write('</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n')


def write_pvd(path: str, entries):
    """A ParaView collection of (time, file) entries (files relative to the .pvd)."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
                '<Collection>\n')
        for t, file in entries:
            rel = os.path.relpath(os.path.abspath(file), base)
            f.write(f'<DataSet timestep="{float(t)!r}" group="" part="0" file="{rel}"/>\n')
        f.write('</Collection>\n</VTKFile>\n')
