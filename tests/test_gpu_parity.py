"""Parity of the HIP library (through the C ABI) with the CPU oracle.

Tolerances (FP64 throughout):
  * tendency: relative max-norm <= 1e-12 (the kernels keep the reference's
    summation order and are built without FMA contraction);
  * state after 100 steps: relative max-norm <= 1e-10;
  * integer work (face tables, banded index maps): bit-exact (host tests);
  * golden errors of the reference's tests reproduced at the reference's rtol.
"""
import json
import math
import os

import numpy as np
import pytest

import problems as P
from gpu_cases import Case, relmax

pytestmark = pytest.mark.gpu

TENDENCY_TOL = 1e-12
STATE_TOL = 1e-10

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))


def _gpu_tendency(case, q, increment=False, dq0=None):
    qf = case.dg.fieldarray(q)
    dqf = case.dg.fieldarray(dq0 if dq0 is not None else np.zeros_like(q))
    case.dg(dqf, qf, 0.0, increment=increment)
    case.dg.ctx.sync()
    return dqf.numpy()


def _check_tendency(case):
    q = case.q
    ref = case.odg.tendency(q)
    got = _gpu_tendency(case, q)
    assert relmax(got, ref) <= TENDENCY_TOL, relmax(got, ref)
    dq0 = case.rng.standard_normal(q.shape) * np.abs(ref).max()
    ref2 = case.odg.tendency(q, dq0.copy(), increment=True)
    got2 = _gpu_tendency(case, q, True, dq0)
    assert relmax(got2, ref2) <= TENDENCY_TOL, relmax(got2, ref2)


EULER_SURF = ["rusanov", "central", "ec", "kg", "roe", "matrix"]


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("form", ["weak", "ec", "kg", "central"])
@pytest.mark.parametrize("sflux", EULER_SURF)
def test_euler_tendency(dim, form, sflux):
    _check_tendency(Case("euler", dim, 4, 3, form=form, sflux=sflux, seed=dim))


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_euler_all_degrees(N):
    _check_tendency(Case("euler", 2, N, 3, form="ec", sflux="roe", seed=N))
    _check_tendency(Case("euler", 3, N, 2, form="weak", sflux="rusanov", seed=N))


@pytest.mark.parametrize("form,sflux,N", [("ec", "roe", 4), ("ec", "rusanov", 3), ("weak", "rusanov", 4),
                                          ("kg", "matrix", 2), ("central", "rusanov", 3),
                                          ("ec", "matrix", 4), ("weak", "roe", 3)])
@pytest.mark.parametrize("dim", [2, 3])
def test_euler_gravity_tendency(dim, form, sflux, N):
    """Walls on every face (reflective), isothermal background, gravity aux."""
    _check_tendency(Case("euler_gravity", dim, N, 3, form=form, sflux=sflux, bc="reflect",
                         periodic=(False,) * dim, gravity_domain=True, seed=7))


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("sflux", ["roe", "rusanov"])
def test_euler_gravity_ec_every_degree(sflux, N):
    """Reference precision, 3-D EC flux differencing with walls: the kernels specialised at every
    degree for the Roe and Rusanov surface fluxes (the bubble and IMEX configurations)."""
    _check_tendency(Case("euler_gravity", 3, N, 2, form="ec", sflux=sflux, bc="reflect",
                         periodic=(False,) * 3, gravity_domain=True, seed=11 + N))


def test_euler_gravity_coriolis():
    _check_tendency(Case("euler_gravity", 3, 3, 2, form="weak", sflux="rusanov", bc="reflect",
                         periodic=(True, True, False), gravity_domain=True, source="coriolis",
                         omega=7.29212e-5 * 50, seed=8))


@pytest.mark.parametrize("dim,N", [(2, 4), (3, 3), (3, 4)])
@pytest.mark.parametrize("sflux", ["rusanov", "central"])
def test_linear_euler_gravity_tendency(dim, N, sflux):
    _check_tendency(Case("linear_euler_gravity", dim, N, 3, form="weak", sflux=sflux, bc="reflect",
                         periodic=(True,) * (dim - 1) + (False,), directions=(dim,),
                         gravity_domain=True, ordering="stacked", seed=9))


@pytest.mark.parametrize("N", [3, 4, 5, 6, 7])
def test_shallow_water_tendency_ensemble_kernels(N):
    """The compiled-for-N weak/Rusanov shallow-water kernels (N = 3..7)."""
    _check_tendency(Case("shallow_water", 2, N, 3, form="weak", sflux="rusanov", seed=N))


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("form,sflux", [("weak", "roe"), ("ec", "ec"), ("ec", "roe"),
                                        ("central", "rusanov"), ("weak", "central")])
def test_shallow_water_tendency(dim, form, sflux):
    _check_tendency(Case("shallow_water", dim, 3, 3, form=form, sflux=sflux, bc="reflect",
                         periodic=(True,) * (dim - 1) + (False,), seed=11))


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("form", ["weak", "central"])
@pytest.mark.parametrize("sflux,bc", [("rusanov", "copy"), ("central", "zero_inflow"),
                                      ("rusanov", "zero_inflow")])
def test_advection_tendency(dim, form, sflux, bc):
    _check_tendency(Case("advection", dim, 4, 3, form=form, sflux=sflux, bc=bc,
                         periodic=(True,) * (dim - 1) + (bc == "copy",), seed=12))


def test_prescribed_boundary_sod():
    """experiment/euler/sod.jl: tag 1 (x = 0) gets the left state, tag 2 the right."""
    left = [1.0, 0.0, 1.0 / 0.4]
    right = [0.125, 0.0, 0.1 / 0.4]
    _check_tendency(Case("euler", 1, 4, 6, form="ec", sflux="rusanov", bc="prescribed",
                         periodic=(False,), bc_states=[left, right], domain=[(0.0, 1.0)], seed=13))


def test_partial_directions():
    """directions = (1,) and (2,) of a 2-D operator (the IMEX split)."""
    for dirs in [(1,), (2,)]:
        _check_tendency(Case("euler", 2, 4, 3, form="ec", sflux="rusanov", directions=dirs, seed=14))
        _check_tendency(Case("advection", 2, 4, 3, form="weak", sflux="rusanov", directions=dirs,
                             bc="zero_inflow", periodic=(True, False), seed=15))


@pytest.mark.parametrize("law,dim,form,sflux", [("euler", 2, "ec", "rusanov"),
                                                ("euler_gravity", 3, "ec", "roe"),
                                                ("shallow_water", 2, "weak", "rusanov")])
def test_fused_lsrk_stage(law, dim, form, sflux):
    """dq = a*dq + L(q); q_out = q + bdt*dq in one kernel (src/odesolvers.jl:51-53)."""
    kw = dict(bc="reflect", periodic=(False,) * dim, gravity_domain=True) if law == "euler_gravity" else {}
    c = Case(law, dim, 4, 3, form=form, sflux=sflux, seed=16, **kw)
    q = c.q
    dq0 = c.rng.standard_normal(q.shape) * 1e-3 * np.abs(q).max()
    a, bdt = -0.4178904745, 0.00371
    qo, dqo = q.copy(), dq0.copy()
    c.odg.lsrk_stage(qo, dqo, a, bdt)
    qf, qout, dqf = c.dg.fieldarray(q), c.dg.fieldarray(), c.dg.fieldarray(dq0)
    c.dg.lsrk_stage(qf, qout, dqf, a, bdt)
    c.dg.ctx.sync()
    assert relmax(dqf.numpy(), dqo) <= TENDENCY_TOL
    assert relmax(qout.numpy(), qo) <= TENDENCY_TOL


def _vortex_case(form):
    import atum_hip as A
    from oracle import oracle as O
    K, N = 8, 4
    v = P.julia_range(-5, 5, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    vform = A.WeakForm() if form == "weak" else A.FluxDifferencingForm(A.EntropyConservativeFlux())
    dg = A.DGSEM(law=A.EulerLaw(2), grid=grid, volume_form=vform,
                 surface_numericalflux=A.RusanovFlux())
    og = O.Grid(N, [v, v], [True, True])
    odg = O.DG("euler", og, surface_flux="rusanov", volume_form="weak" if form == "weak" else "fd",
               volume_flux="ec")
    q = P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0))
    return dg, odg, q


@pytest.mark.parametrize("form", ["ec", "weak"])
def test_vortex_100_steps(form):
    """Isentropic vortex on [-5,5]^2, 8x8, N=4, LSRK54 at CFL 0.1: state after
    100 steps (500 fused stages, hipGraph replay)."""
    import atum_hip as A
    dg, odg, q = _vortex_case(form)
    dt = 0.1 * (10 / 8) / 4
    qo = q.copy()
    odg.lsrk_steps(qo, dt, 100)
    qf = dg.fieldarray(q)
    solver = A.LSRK54(dg, qf, dt)
    solver.run(qf, 100)
    dg.ctx.sync()
    assert relmax(qf.numpy(), qo) <= STATE_TOL
    # mass and energy conserved to round-off on the periodic domain
    got = qf.numpy()
    for s in (0, 3):
        assert abs(np.sum(dg.MJ * got[s]) - np.sum(dg.MJ * q[s])) <= 1e-12 * np.sum(dg.MJ * np.abs(q[s]))


def test_graph_and_eager_runs_agree():
    """hipGraph replay of step pairs gives bitwise the same state as eager launches,
    for an even and an odd number of steps."""
    import atum_hip as A
    dg, _, q = _vortex_case("ec")
    dt = 0.03125
    for n in (6, 7):
        a = dg.fieldarray(q)
        b = dg.fieldarray(q)
        A.LSRK54(dg, a, dt, use_graph=True).run(a, n)
        A.LSRK54(dg, b, dt, use_graph=False).run(b, n)
        dg.ctx.sync()
        assert np.array_equal(a.numpy(), b.numpy())


def test_lsrk144_run():
    import atum_hip as A
    dg, odg, q = _vortex_case("ec")
    dt = 0.05
    qo = q.copy()
    odg.lsrk_steps(qo, dt, 20, scheme="LSRK144")
    qf = dg.fieldarray(q)
    A.LSRK144(dg, qf, dt).run(qf, 20)
    dg.ctx.sync()
    assert relmax(qf.numpy(), qo) <= STATE_TOL


def test_golden_isentropic_vortex_on_gpu():
    """test/euler/isentropicvortex.jl:83,88 through the HIP path."""
    import atum_hip as A
    K, N = 5, 4
    for form, key in (("weak", "isentropic_vortex_weak"), ("ec", "isentropic_vortex_ec")):
        vx, vy = P.julia_range(0, 20, K + 1), P.julia_range(-5, 5, K + 1)
        grid = A.brickgrid(A.LobattoCell(N, 2), (vx, vy), (True, True))
        vform = A.WeakForm() if form == "weak" else A.FluxDifferencingForm(A.EntropyConservativeFlux())
        dg = A.DGSEM(law=A.EulerLaw(2), grid=grid, volume_form=vform,
                     surface_numericalflux=A.RusanovFlux())
        q = dg.fieldarray(P.isentropic_vortex(grid.points, 0.0))
        n = A.solve_(q, 5.0, A.LSRK54(dg, q, 0.25 * P.julia_step(-5, 5, K + 1) / N))
        assert n == 40
        err = A.weightednorm(dg, q.numpy() - P.isentropic_vortex(grid.points, 5.0))
        assert err == pytest.approx(GOLD[key]["value"], rel=GOLD[key]["rtol"])


@pytest.mark.parametrize("dim", [2, 3])
def test_golden_euler_wave_on_gpu(dim):
    import atum_hip as A
    K, N = 5, 4
    for form in ("weak", "ec"):
        v = P.julia_range(-1, 1, K + 1)
        grid = A.brickgrid(A.LobattoCell(N, dim), [v] * dim, [True] * dim)
        vform = A.WeakForm() if form == "weak" else A.FluxDifferencingForm(A.EntropyConservativeFlux())
        dg = A.DGSEM(law=A.EulerLaw(dim), grid=grid, volume_form=vform,
                     surface_numericalflux=A.RusanovFlux())
        q = dg.fieldarray(P.euler_wave(grid.points, 0.0))
        dt = 0.25 * P.julia_step(-1, 1, K + 1) / N / math.sqrt(1.4)
        tend = dt if dim == 3 else 0.7
        A.solve_(q, tend, A.LSRK54(dg, q, dt))
        err = A.weightednorm(dg, q.numpy() - P.euler_wave(grid.points, tend))
        g = GOLD[f"euler_wave_{dim}d_{form}"]
        assert err == pytest.approx(g["value"], rel=g["rtol"])


def _imex_case(form, split, K=5, N=4):
    import atum_hip as A
    v = P.julia_range(-1, 1, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, False), "stacked")
    u = (1.0, 0.3)
    law = A.AdvectionLaw(2, u, problem=A.Problem(boundary="zero_inflow"))
    vform = A.WeakForm() if form == "weak" else A.FluxDifferencingForm(A.CentralFlux())
    dg = A.DGSEM(law=law, grid=grid, volume_form=vform, surface_numericalflux=A.RusanovFlux(),
                 directions=(1,) if split else (1, 2))
    lin = A.DGSEM(law=law, grid=grid, volume_form=vform, surface_numericalflux=A.RusanovFlux(),
                  directions=(2,))
    return grid, u, dg, lin


@pytest.mark.parametrize("form", ["weak", "central"])
@pytest.mark.parametrize("split", [True, False])
def test_golden_imex_wave_on_gpu(form, split):
    """test/advection/imex_wave_2d.jl:94,98: ARK23 + banded solve on the device."""
    import atum_hip as A
    K, N = 5, 4
    grid, u, dg, lin = _imex_case(form, split, K, N)
    q = dg.fieldarray(P.imex_gaussian(grid.points, 0.0, u))
    dt = 0.25 * P.julia_step(-1, 1, K + 1) / N / math.hypot(1.0, 0.3)
    A.solve_(q, 1.0, A.ARK23(dg, lin, q, dt, split_rhs=split))
    err = A.weightednorm(dg, q.numpy() - P.imex_gaussian(grid.points, 1.0, u))
    g = GOLD[f"imex_wave_2d_{form}"]
    assert err == pytest.approx(g["value"], rel=g["rtol"])


def test_banded_matches_oracle():
    """Probed band of the linear vertical operator and its LU (device) against the
    oracle's; the solve against the oracle's banded solve."""
    import atum_hip as A
    from oracle import oracle as O
    c = Case("linear_euler_gravity", 3, 3, (3, 2, 4), form="weak", sflux="rusanov", bc="reflect",
             periodic=(True, True, False), directions=(3,), gravity_domain=True, ordering="stacked")
    dt = 5.0
    gam = 1 - 1 / math.sqrt(2)
    qf = c.dg.fieldarray(c.q)
    ark = A.ARK23(None, c.dg, qf, dt)
    band_gpu = ark.band_host()
    band_orc = c.odg.banded_factor(c.odg.banded_assemble(), -dt * gam)
    assert band_gpu.shape == band_orc.shape
    assert relmax(band_gpu, band_orc) <= 1e-12
    b = c.rng.standard_normal(c.q.shape)
    xo = c.odg.banded_solve(band_orc, b)
    bf = c.dg.fieldarray(b)
    c.dg.ctx.check(A._lib.lib.atm_banded_solve(ark.band, bf.ptr, bf.ptr), "solve")
    c.dg.ctx.sync()
    assert relmax(bf.numpy(), xo) <= 1e-11
    del O


def test_ark23_imex_euler_gravity_steps():
    """ARK23 unsplit with EC explicit + linear vertical implicit operators (the
    IMEX configuration at reduced size): 10 steps against the oracle."""
    import atum_hip as A
    from oracle import oracle as O
    dim, N = 3, 3
    ex = Case("euler_gravity", dim, N, (2, 2, 3), form="ec", sflux="rusanov", bc="reflect",
              periodic=(True, True, False), gravity_domain=True, ordering="stacked", seed=21)
    lin = Case("linear_euler_gravity", dim, N, (2, 2, 3), form="weak", sflux="rusanov",
               bc="reflect", periodic=(True, True, False), directions=(3,), gravity_domain=True,
               ordering="stacked", seed=21)
    q = ex.q
    dt = 0.1  # within the horizontal acoustic CFL of the explicit part
    qo = q.copy()
    O.solve_ark23(ex.odg, lin.odg, qo, dt, 100 * dt, split_rhs=False, max_steps=10)
    assert np.all(np.isfinite(qo))
    qf = ex.dg.fieldarray(q)
    ark = A.ARK23(ex.dg, lin.dg, qf, dt, split_rhs=False)
    ark.run(qf, 10)
    ex.dg.ctx.sync()
    assert relmax(qf.numpy(), qo) <= STATE_TOL


def test_linear_euler_gravity_golden_on_gpu():
    """test/euler_gravity/linear.jl:138,147 (5 days, 4129 ARK23 steps, rtol 1e-3)."""
    import atum_hip as A
    K, N = 5, 4
    v = P.julia_range(0, 10e3, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, False), "stacked")
    prob = A.Problem(boundary="reflect", referencestate=lambda x: P.linear_test_reference(x))
    law = A.LinearEulerGravityLaw(A.EulerGravityLaw(2, grav=9.81, problem=prob))
    dg = A.DGSEM(law=law, grid=grid, surface_numericalflux=A.RusanovFlux(), directions=(2,))
    q0 = P.linear_test_initial(dg.auxstate)
    q = dg.fieldarray(q0)
    dt = 100.0 * grid.min_node_distance(dims=(1,)) / 330
    n = A.solve_(q, 5 * 24 * 60 * 60.0, A.ARK23(None, dg, q, dt), adjust_final=False)
    assert n == 4129
    err = A.weightednorm(dg, q.numpy() - q0)
    g = GOLD["linear_euler_gravity_2d"]
    assert err == pytest.approx(g["value"], rel=g["rtol"])


@pytest.mark.parametrize("varying", [False, True])
def test_banded_dedupe_and_solve(varying):
    """Lines with bit-identical bands share one factor; with a horizontally varying
    reference state every line keeps its own.  Solves match the oracle either way."""
    import atum_hip as A
    from oracle import oracle as O
    dim, N = 3, 2
    vx, vz = P.julia_range(0, 2000, 3), P.julia_range(0, 1000, 4)
    grid = A.brickgrid(A.LobattoCell(N, dim), (vx, vx, vz), (True, True, False), "stacked")

    def ref(x):
        rho, p = P.isothermal_reference(x)
        if varying:
            rho = rho * (1 + 0.01 * x[0] / 2000 + 0.0037 * x[1] / 2000)
        return rho, p

    prob = A.Problem(boundary="reflect", referencestate=ref)
    law = A.LinearEulerGravityLaw(A.EulerGravityLaw(dim, grav=9.81, problem=prob))
    dg = A.DGSEM(law=law, grid=grid, surface_numericalflux=A.RusanovFlux(), directions=(dim,))
    og = O.Grid(N, [vx, vx, vz], [True, True, False], stacked=True)
    odg = O.DG("linear_euler_gravity", og, directions=(dim,), surface_flux="rusanov", bc="reflect",
               aux=dg.auxstate, grav=9.81)
    dt = 3.0
    q0 = P.linear_test_initial(dg.auxstate)
    qf = dg.fieldarray(q0)
    ark = A.ARK23(None, dg, qf, dt)
    nu = A._lib.lib.atm_banded_nunique(ark.band)
    # a line's band is fixed by its horizontal node position (x, y): one factor per
    # distinct position when the reference state varies horizontally, else one in all
    colelem, KV, ncols = grid.columns()
    Nh = grid.Np // grid.Nq
    pos = {(grid.points[0, grid.Np * colelem[c * KV] + h], grid.points[1, grid.Np * colelem[c * KV] + h])
           for c in range(ncols) for h in range(Nh)}
    assert nu == (len(pos) if varying else 1)
    assert len(pos) > 1
    band_orc = odg.banded_factor(odg.banded_assemble(), -dt * (1 - 1 / math.sqrt(2)))
    assert relmax(ark.band_host(), band_orc) <= 1e-12
    b = np.random.default_rng(3).standard_normal(q0.shape)
    bf = dg.fieldarray(b)
    dg.ctx.check(A._lib.lib.atm_banded_solve(ark.band, bf.ptr, bf.ptr), "solve")
    dg.ctx.sync()
    assert relmax(bf.numpy(), odg.banded_solve(band_orc, b)) <= 1e-11
