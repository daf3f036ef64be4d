"""VTK output (host side): interpolation to equally spaced points reproduces the
element polynomials exactly, and the written file holds the mesh's sub-cells,
the interpolated coordinates and fields."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from atum_hip import vtk
from atum_hip.grid import LobattoCell, brickgrid, lgl


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7])
def test_equispaced_interpolation_is_exact_for_polynomials(N):
    x = lgl(N)[0]
    t = np.linspace(-1, 1, N + 1)
    P = vtk.equispaced_interpolation(N)
    for deg in range(N + 1):
        np.testing.assert_allclose(P @ x ** deg, t ** deg, rtol=0, atol=1e-13)
    np.testing.assert_allclose(P.sum(axis=1), 1.0, atol=1e-14)  # partition of unity


def _read(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for da in root.iter("DataArray"):
        name = da.get("Name") or "Points"
        kind = da.get("type")
        arrays[name] = np.array(da.text.split(), dtype=np.float64 if kind.startswith("Float") else np.int64)
    return piece, arrays


@pytest.mark.parametrize("dim,N,Kd", [(1, 3, (4,)), (2, 2, (3, 2)), (3, 3, (2, 1, 2))])
def test_vtu_contents(tmp_path, dim, N, Kd):
    coords = [np.linspace(0.0, 1.0 + d, k + 1) ** 1.0 for d, k in enumerate(Kd)]
    g = brickgrid(LobattoCell(N, dim), coords, (False,) * dim)
    x = g.points
    f = 1.0 + x[0] ** N + (x[-1] ** 2 if dim > 1 else 0.0)  # a polynomial of degree <= N per direction
    path = str(tmp_path / "a.vtu")
    vtk.write_vtu(path, x, {"f": f, "g": -f}, N, dim)
    piece, arr = _read(path)
    K = g.K
    assert int(piece.get("NumberOfPoints")) == K * (N + 1) ** dim
    assert int(piece.get("NumberOfCells")) == K * N ** dim
    pts = arr["Points"].reshape(-1, 3)
    # affine elements: the equally spaced points of each element, and the field at them
    fx = 1.0 + pts[:, 0] ** N + (pts[:, dim - 1] ** 2 if dim > 1 else 0.0)
    np.testing.assert_allclose(arr["f"], fx, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(arr["g"], -fx, rtol=1e-12, atol=1e-12)
    conn = arr["connectivity"].reshape(K * N ** dim, 2 ** dim)
    assert conn.min() == 0 and conn.max() == K * (N + 1) ** dim - 1
    assert np.all(np.diff(arr["offsets"]) == 2 ** dim)
    # every sub-cell has positive extent along each axis (corner order as VTK expects)
    c = pts[conn]
    for d in range(dim):
        assert np.all(c[:, :, d].max(axis=1) > c[:, :, d].min(axis=1))
    assert np.all(c[:, 1, 0] > c[:, 0, 0])


def test_pvd(tmp_path):
    p = str(tmp_path / "run.pvd")
    vtk.write_pvd(p, [(0.0, str(tmp_path / "s0.vtu")), (0.5, str(tmp_path / "s1.vtu"))])
    root = ET.parse(p).getroot()
    ds = root.findall("Collection/DataSet")
    assert [d.get("file") for d in ds] == ["s0.vtu", "s1.vtu"]
    assert [float(d.get("timestep")) for d in ds] == [0.0, 0.5]
