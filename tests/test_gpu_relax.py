"""Device measures and relaxation LSRK against the oracle (src/measures.jl,
src/odesolvers.jl:69-146).

Tolerances:
  * entropy integral: 1e-13 relative; entropy product: 1e-12 relative (the
    per-node terms differ by last-bit log/pow rounding, the sums are nearly
    exact on both sides: double-double here, long double in the oracle);
  * weighted norms: 1e-14 relative against the oracle (p = 2) and against
    numpy (p = 1, inf exactly);
  * RLSRK: the reference's own entropy bounds (|d eta| <= 20 eps / 30 eps after
    t = 5) and the final state within 1e-10 relative max-norm of the oracle's
    RLSRK run; the relaxation factor within 1e-10.
"""
import json
import os

import numpy as np
import pytest

import problems as P
from gpu_cases import Case, relmax

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
EPS = np.finfo(float).eps

CASES = [("euler", 1, "ec", "ec"), ("euler", 2, "ec", "rusanov"), ("euler", 3, "weak", "roe"),
         ("euler_gravity", 3, "ec", "roe"), ("shallow_water", 2, "ec", "rusanov"),
         ("advection", 2, "central", "rusanov")]


def _case(law, dim, form, sflux, seed):
    kw = dict(bc="reflect", periodic=(False,) * dim, gravity_domain=True) if law == "euler_gravity" else {}
    return Case(law, dim, 3, 3, form=form, sflux=sflux, seed=seed, **kw)


@pytest.mark.parametrize("law,dim,form,sflux", CASES)
def test_entropy_integral_and_product(law, dim, form, sflux):
    import atum_hip as A
    c = _case(law, dim, form, sflux, seed=11 + dim)
    ref = c.odg.entropyintegral(c.q)
    got = A.entropyintegral(c.dg, c.dg.fieldarray(c.q))
    assert abs(got - ref) <= 1e-13 * abs(ref)
    k = c.rng.standard_normal(c.q.shape)
    refp = c.odg.entropyproduct(c.q, k)
    gotp = A.entropyproduct(c.dg, c.dg.fieldarray(c.q), c.dg.fieldarray(k))
    assert abs(gotp - refp) <= 1e-12 * abs(refp)
    # reproducible bit for bit
    assert A.entropyintegral(c.dg, c.q) == got


def test_entropy_unsupported_for_linear_law():
    import atum_hip as A
    c = Case("linear_euler_gravity", 2, 3, 2, sflux="rusanov", bc="reflect", periodic=(False, False),
             gravity_domain=True)
    with pytest.raises(A.AtumError):
        A.entropyintegral(c.dg, c.q)


@pytest.mark.parametrize("law,dim", [("euler", 2), ("euler_gravity", 3), ("advection", 1)])
def test_weightednorm(law, dim):
    import atum_hip as A
    c = _case(law, dim, "weak" if law != "euler" else "ec", "rusanov", seed=3)
    q = c.q - c.q.mean(axis=1, keepdims=True)
    assert abs(A.weightednorm(c.dg, q) - c.odg.weightednorm(q)) <= 1e-14 * c.odg.weightednorm(q)
    MJ = c.grid.MJ
    comp1 = np.array([np.sum(MJ * np.abs(f)) for f in q])
    got1 = A.weightednorm(c.dg, q, 1, componentwise=True)
    np.testing.assert_allclose(got1, comp1, rtol=1e-13, atol=0)
    assert A.weightednorm(c.dg, q, 1) == pytest.approx(comp1.sum(), rel=1e-13)
    assert A.weightednorm(c.dg, q, np.inf) == np.abs(q).max()
    np.testing.assert_array_equal(A.weightednorm(c.dg, q, np.inf, componentwise=True), np.abs(q).max(axis=1))


def _entropy_square(law, ic, vflux, sflux, cfl, speed, **kw):
    """The 1-D square-wave entropy tests (test/*/entropy_conservation_1d.jl), N = 4, K = 5,
    RLSRK54 to t = 5, on the device; the oracle runs the same problem."""
    import atum_hip as A
    from oracle import oracle as O
    N, K = 4, 5
    v = P.julia_range(-1, 1, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 1), (v,), (True,))
    if law == "euler":
        alaw = A.EulerLaw(1)
    elif law == "shallow_water":
        alaw = A.ShallowWaterLaw(1, grav=kw["grav"])
    else:
        alaw = A.AdvectionLaw(1, kw["u"])
    flux = {"ec": A.EntropyConservativeFlux(), "central": A.CentralFlux()}
    dg = A.DGSEM(law=alaw, grid=grid, volume_form=A.FluxDifferencingForm(flux[vflux]),
                 surface_numericalflux=flux[sflux])
    q0 = np.ascontiguousarray(ic(grid.points))
    dt = cfl * grid.min_node_distance() / speed
    q = dg.fieldarray(q0)
    solver = A.RLSRK54(dg, q, dt)
    nsteps = A.solve_(q, 5.0, solver)
    got = q.numpy()
    og = O.Grid(N, [v], [True])
    odg = O.DG(law, og, volume_form="fd", volume_flux=vflux, surface_flux=sflux, **kw)
    ref = q0.copy()
    on, ogam = odg.solve_rlsrk(ref, dt, 5.0)
    assert nsteps == on
    e0 = odg.entropyintegral(q0)
    return (odg.entropyintegral(got) - e0) / abs(e0), got, ref, solver.gamma, ogam


@pytest.mark.parametrize("name,args", [
    ("entropy_euler_1d", ("euler", P.euler_square, "ec", "ec", 0.5, np.sqrt(1.4), {})),
    ("entropy_shallow_water_1d", ("shallow_water", P.sw_square, "ec", "ec", 0.25, np.sqrt(10.0),
                                  {"grav": 10.0})),
    ("entropy_advection_1d", ("advection", lambda x: P.square_rho(x)[None, :], "central", "central",
                              0.25, 1.0, {"u": (1.0,)})),
])
def test_rlsrk_entropy_conservation(name, args):
    """test/{euler,shallow_water,advection}/entropy_conservation_1d.jl through the HIP path."""
    law, ic, vf, sf, cfl, speed, kw = args
    d, got, ref, gam, ogam = _entropy_square(law, ic, vf, sf, cfl, speed, **kw)
    assert abs(d) <= GOLD[name]["bound_eps"] * EPS
    assert relmax(got, ref) <= 1e-10
    assert abs(gam - ogam) <= 1e-10


def test_rlsrk_dissipative_vortex():
    """RLSRK54 with a dissipative surface flux (gamma != 1): 20 steps of the
    isentropic vortex, EC volume + Rusanov, against the oracle."""
    import atum_hip as A
    from oracle import oracle as O
    K, N = 8, 4
    v = P.julia_range(-5, 5, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    dg = A.DGSEM(law=A.EulerLaw(2), grid=grid,
                 volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RusanovFlux())
    q0 = P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0))
    dt = 0.1 * (10 / K) / N
    q = dg.fieldarray(q0)
    solver = A.RLSRK54(dg, q, dt)
    A.solve_(q, 20 * dt, solver)
    odg = O.DG("euler", O.Grid(N, [v, v], [True, True]), volume_form="fd", volume_flux="ec",
               surface_flux="rusanov")
    ref = q0.copy()
    _, ogam = odg.solve_rlsrk(ref, dt, 20 * dt)
    assert abs(ogam - 1) > 1e-12  # the relaxation is active
    assert abs(solver.gamma - ogam) <= 1e-10
    assert relmax(q.numpy(), ref) <= 1e-10
    # entropy does not grow (Rusanov dissipates)
    assert A.entropyintegral(dg, q) <= A.entropyintegral(dg, q0) + 1e-13 * abs(A.entropyintegral(dg, q0))
