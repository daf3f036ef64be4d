"""Pin the CPU oracle against the reference's own golden values.

Every level-1 accuracy anchor in the reference's test suite is reproduced by
running the same set-up through the oracle (N=4, K=5, Float64) and comparing
the MJ-weighted L2 error with the value the reference test hard-codes, at the
reference test's own tolerance (default isapprox rtol = sqrt(eps)).
"""
import json
import os

import numpy as np
import pytest

import problems as P
from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
N = 4
K = 5


def check(name, err):
    g = GOLD[name]
    assert err == pytest.approx(g["value"], rel=g["rtol"]), (name, err, g)


@pytest.mark.parametrize("dim,cfl", [(1, 0.5), (2, 0.25), (3, 0.25)])
@pytest.mark.parametrize("form", ["weak", "central"])
def test_advection_wave(dim, cfl, form):
    """test/advection/wave_{1,2,3}d.jl: Rusanov surface, LSRK54 to t = 0.7."""
    v = P.julia_range(-1, 1, K + 1)
    g = O.Grid(N, [v] * dim, [True] * dim)
    u = (1.0,) * dim
    dg = O.DG("advection", g, surface_flux="rusanov", u=u,
              volume_form="weak" if form == "weak" else "fd", volume_flux="central")
    q = P.advection_wave(g.points, 0.0, u)
    dt = cfl * P.julia_step(-1, 1, K + 1) / N / np.sqrt(dim)
    dg.solve_lsrk(q, dt, 0.7)
    check(f"advection_wave_{dim}d_{form}", dg.weightednorm(q - P.advection_wave(g.points, 0.7, u)))


@pytest.mark.parametrize("dim,cfl", [(1, 0.5), (2, 0.25), (3, 0.25)])
@pytest.mark.parametrize("form", ["weak", "ec"])
def test_euler_wave(dim, cfl, form):
    """test/euler/wave_{1,2,3}d.jl (3-D runs a single step, timeend = dt)."""
    v = P.julia_range(-1, 1, K + 1)
    g = O.Grid(N, [v] * dim, [True] * dim)
    dg = O.DG("euler", g, surface_flux="rusanov",
              volume_form="weak" if form == "weak" else "fd", volume_flux="ec")
    q = P.euler_wave(g.points, 0.0)
    dt = cfl * P.julia_step(-1, 1, K + 1) / N / np.sqrt(1.4)
    tend = dt if dim == 3 else 0.7
    dg.solve_lsrk(q, dt, tend)
    check(f"euler_wave_{dim}d_{form}", dg.weightednorm(q - P.euler_wave(g.points, tend)))


@pytest.mark.parametrize("form", ["weak", "ec"])
def test_isentropic_vortex(form):
    """test/euler/isentropicvortex.jl: [0,20]x[-5,5], 40 LSRK54 steps to t = 5."""
    g = O.Grid(N, [P.julia_range(0, 20, K + 1), P.julia_range(-5, 5, K + 1)], [True, True])
    dg = O.DG("euler", g, surface_flux="rusanov",
              volume_form="weak" if form == "weak" else "fd", volume_flux="ec")
    q = P.isentropic_vortex(g.points, 0.0)
    nsteps, _ = dg.solve_lsrk(q, 0.25 * P.julia_step(-5, 5, K + 1) / N, 5.0)
    assert nsteps == 40
    check(f"isentropic_vortex_{form}", dg.weightednorm(q - P.isentropic_vortex(g.points, 5.0)))


@pytest.mark.parametrize("form", ["weak", "central"])
@pytest.mark.parametrize("split", [True, False])
def test_imex_wave(form, split):
    """test/advection/imex_wave_2d.jl: ARK23 with a probed banded vertical operator."""
    v = P.julia_range(-1, 1, K + 1)
    g = O.Grid(N, [v, v], [True, False], stacked=True)
    u = (1.0, 0.3)
    kw = dict(volume_form="weak" if form == "weak" else "fd", volume_flux="central",
              surface_flux="rusanov", u=u, bc="zero_inflow")
    dg = O.DG("advection", g, directions=(1,) if split else (1, 2), **kw)
    lin = O.DG("advection", g, directions=(2,), **kw)
    q = P.imex_gaussian(g.points, 0.0, u)
    dt = 0.25 * P.julia_step(-1, 1, K + 1) / N / np.hypot(1.0, 0.3)
    O.solve_ark23(dg, lin, q, dt, 1.0, split_rhs=split)
    check(f"imex_wave_2d_{form}", dg.weightednorm(q - P.imex_gaussian(g.points, 1.0, u)))


def test_linear_euler_gravity():
    """test/euler_gravity/linear.jl: linear model, ARK23(nothing, dg), 5 days, rtol 1e-3."""
    v = P.julia_range(0, 10e3, K + 1)
    g = O.Grid(N, [v, v], [True, False], stacked=True)
    aux = P.flat_aux(g.points, 9.81, P.linear_test_reference(g.points))
    lin = O.DG("linear_euler_gravity", g, directions=(2,), surface_flux="rusanov", bc="reflect",
               aux=aux, grav=9.81)
    q = P.linear_test_initial(aux)
    qref = q.copy()
    dt = 100.0 * g.min_node_distance(dims=(1,)) / 330
    O.solve_ark23(None, lin, q, dt, 5 * 24 * 60 * 60.0, split_rhs=True, adjust_final=False)
    check("linear_euler_gravity_2d", lin.weightednorm(q - qref))


def _entropy_run(law, ic, vflux, sflux, cfl, speed, **kw):
    v = P.julia_range(-1, 1, K + 1)
    g = O.Grid(N, [v], [True])
    dg = O.DG(law, g, volume_form="fd", volume_flux=vflux, surface_flux=sflux, **kw)
    q = np.ascontiguousarray(ic(g.points))
    e0 = dg.entropyintegral(q)
    dg.solve_rlsrk(q, cfl * g.min_node_distance() / speed, 5.0)
    return (dg.entropyintegral(q) - e0) / abs(e0)


def test_entropy_conservation_euler():
    """test/euler/entropy_conservation_1d.jl:71 (RLSRK54, EC volume + EC surface)."""
    d = _entropy_run("euler", P.euler_square, "ec", "ec", 0.5, np.sqrt(1.4))
    assert abs(d) <= GOLD["entropy_euler_1d"]["bound_eps"] * np.finfo(float).eps


def test_entropy_conservation_shallow_water():
    """test/shallow_water/entropy_conservation_1d.jl:70 (grav = 10, the law's default)."""
    d = _entropy_run("shallow_water", P.sw_square, "ec", "ec", 0.25, np.sqrt(10.0), grav=10.0)
    assert abs(d) <= GOLD["entropy_shallow_water_1d"]["bound_eps"] * np.finfo(float).eps


def test_entropy_conservation_advection():
    """test/advection/entropy_conservation_1d.jl:68 (Central volume + Central surface)."""
    d = _entropy_run("advection", lambda x: P.square_rho(x)[None, :], "central", "central", 0.25,
                     1.0, u=(1.0,))
    assert abs(d) <= GOLD["entropy_advection_1d"]["bound_eps"] * np.finfo(float).eps
