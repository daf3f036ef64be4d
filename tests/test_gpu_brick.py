"""Brick meshes built on the device (atm_dg_create_brick) against the host builder.

The geometry the library forms from the vertex coordinates must equal the host
builder's bit for bit (the host builder is itself bit-identical to the
oracle's, tests/test_host.py), and an operator on device-built geometry must
give bit-identical tendencies.
"""
import numpy as np
import pytest

import problems as P
from gpu_cases import make_state

pytestmark = pytest.mark.gpu

GEOM = ("points", "g", "MJ", "MJI", "faceMJ", "normal", "faceix_p", "bctag")


def _coords(dim, Kd, seed):
    """Vertex coordinates: Julia ranges with one dimension perturbed (non-uniform h)."""
    rng = np.random.default_rng(seed)
    cs = []
    for d in range(dim):
        v = P.julia_range(-1.0 + d, 2.0 + 0.5 * d, Kd[d] + 1)
        if d == 0 and Kd[d] > 1:
            v = v.copy()
            v[1:-1] += 0.2 * (v[1] - v[0]) * rng.uniform(-1, 1, Kd[d] - 1)
        cs.append(v)
    return cs


CASES = [(1, 3, (5,), (True,), "cartesian"), (1, 6, (4,), (False,), "cartesian"),
         (2, 4, (5, 3), (True, False), "cartesian"), (2, 2, (3, 4), (False, True), "stacked"),
         (2, 7, (2, 2), (True, True), "cartesian"), (3, 3, (3, 2, 4), (True, True, False), "stacked"),
         (3, 4, (2, 3, 2), (False, True, False), "cartesian"), (3, 1, (2, 2, 2), (True, False, True), "stacked")]


@pytest.mark.parametrize("dim,N,Kd,periodic,ordering", CASES)
def test_device_geometry_bit_identical(dim, N, Kd, periodic, ordering):
    import atum_hip as A
    cs = _coords(dim, Kd, seed=sum(Kd) + N)
    grid = A.brickgrid(A.LobattoCell(N, dim), cs, periodic, ordering)
    law = A.EulerLaw(dim)
    kw = dict(law=law, surface_numericalflux=A.RusanovFlux(),
              volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()))
    dev = A.DGSEM(grid=A.brickgrid(A.LobattoCell(N, dim), cs, periodic, ordering, lazy=True),
                  geometry="device", **kw)
    host = A.DGSEM(grid=grid, **kw)
    want = {"points": grid.points, "g": grid.g, "MJ": grid.MJ, "MJI": grid.MJI, "faceMJ": grid.faceMJ,
            "normal": grid.normal, "faceix_p": grid.faceix_p, "bctag": grid.bctag}
    for name in GEOM:
        got = dev.geometry_array(name)
        np.testing.assert_array_equal(got.reshape(want[name].shape), want[name], err_msg=name)
        assert got.dtype == np.asarray(want[name]).dtype
    # the same operator: identical tendencies
    q = make_state("euler", grid.points, np.random.default_rng(N))
    a, b = dev.fieldarray(q), host.fieldarray(q)
    da, db = dev.fieldarray(), host.fieldarray()
    dev(da, a, 0.0, increment=False)
    host(db, b, 0.0, increment=False)
    np.testing.assert_array_equal(da.numpy(), db.numpy())


def test_device_geometry_gravity_and_lsrk():
    """Euler-gravity with aux from the (host-formed) points, reflecting walls and
    10 LSRK54 steps: device-built and host-built operators agree bit for bit."""
    import atum_hip as A
    Kd, N = 3, 3
    v = P.julia_range(0, 2000, Kd + 1)
    law = A.EulerGravityLaw(3, grav=9.81, problem=A.Problem(boundary="reflect"))
    kw = dict(law=law, surface_numericalflux=A.RoeFlux(),
              volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()), precision="fast")
    host = A.DGSEM(grid=A.brickgrid(A.LobattoCell(N, 3), (v, v, v), (False,) * 3), **kw)
    dev = A.DGSEM(grid=A.brickgrid(A.LobattoCell(N, 3), (v, v, v), (False,) * 3, lazy=True),
                  geometry="device", **kw)
    q0 = P.rising_bubble(host.grid.points)
    out = []
    for dg in (host, dev):
        q = dg.fieldarray(q0)
        A.LSRK54(dg, q, 0.05).run(q, 10)
        dg.ctx.sync()
        out.append(q.numpy())
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(dev.MJ, host.MJ)
