"""Batched launches: every member of a batch ends bitwise where it would alone."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def _sw_member(A, K, N, amp, w, c, precision="reference"):
    v = P.julia_range(0.0, 1.0, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    dg = A.DGSEM(law=A.ShallowWaterLaw(2, grav=9.81), grid=grid,
                 surface_numericalflux=A.RusanovFlux(), precision=precision)
    return dg, P.sw_bump(grid.points, amp, w, c), P.sw_ensemble_dt(K, N, amp)


@pytest.mark.parametrize("nsteps", [3, 4])
def test_batch_equals_alone_shallow_water(nsteps):
    import atum_hip as A
    params = P.sw_ensemble_params(n=12, seed=7, kmin=2, kmax=9)
    members = [_sw_member(A, *p) for p in params]
    dgs = [m[0] for m in members]
    qs = [dg.fieldarray(m[1]) for dg, m in zip(dgs, members)]
    dts = [m[2] for m in members]
    ens = A.LSRKEnsemble(dgs, qs, dts)
    assert len({p[1] for p in params}) > 1
    assert ens.batch.ngroups == 1  # every degree in one launch per stage
    ens.run(nsteps)
    for dg, m, q in zip(dgs, members, qs):
        alone = dg.fieldarray(m[1])
        A.LSRK54(dg, alone, m[2]).run(alone, nsteps)
        dg.ctx.sync()
        assert np.array_equal(q.numpy(), alone.numpy())


@pytest.mark.parametrize("sflux", ["rusanov", "roe"])
def test_batch_fast_euler_all_degrees_one_launch(sflux):
    """2-D Euler vortices of N = 3..7 (FAST EC volume) advance in ONE launch per stage;
    members whose own shape is smaller than the launch's workgroup take more
    elements per workgroup; each ends bitwise where it would alone."""
    import atum_hip as A
    members = []
    for N, K in ((3, 5), (4, 4), (5, 3), (6, 3), (7, 2)):
        v = P.julia_range(-5, 5, K + 1)
        g = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
        flux = A.RusanovFlux() if sflux == "rusanov" else A.RoeFlux()
        dg = A.DGSEM(law=A.EulerLaw(2), grid=g, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                     surface_numericalflux=flux, precision="fast")
        members.append((dg, P.isentropic_vortex(g.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0)), 0.02))
    dgs = [m[0] for m in members]
    qs = [dg.fieldarray(m[1]) for dg, m in zip(dgs, members)]
    ens = A.LSRKEnsemble(dgs, qs, [m[2] for m in members])
    assert ens.batch.ngroups == 1
    ens.run(4)
    for (dg, q0, dt), q in zip(members, qs):
        alone = dg.fieldarray(q0)
        A.LSRK54(dg, alone, dt).run(alone, 4)
        dg.ctx.sync()
        assert np.array_equal(q.numpy(), alone.numpy())


def test_batch_mixed_laws_and_modes():
    """Vortex (EC, fast and reference), bubble (EC + Roe) and shallow water in one batch."""
    import atum_hip as A
    members = []
    v = P.julia_range(-5, 5, 4)
    g = A.brickgrid(A.LobattoCell(4, 2), (v, v), (True, True))
    for prec in ("fast", "reference"):
        dg = A.DGSEM(law=A.EulerLaw(2), grid=g, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                     surface_numericalflux=A.RusanovFlux(), precision=prec)
        members.append((dg, P.isentropic_vortex(g.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0)), 0.05))
    vb = P.julia_range(0, 2000, 3)
    gb = A.brickgrid(A.LobattoCell(4, 3), (vb, vb, vb), (False,) * 3)
    law = A.EulerGravityLaw(3, grav=9.81, problem=A.Problem(boundary="reflect"))
    dgb = A.DGSEM(law=law, grid=gb, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                  surface_numericalflux=A.RoeFlux(), precision="fast")
    members.append((dgb, P.rising_bubble(gb.points), 0.1))
    members.append(_sw_member(A, 3, 5, 0.05, 0.1, (0.3, 0.6)))
    dgs = [m[0] for m in members]
    qs = [dg.fieldarray(m[1]) for dg, m in zip(dgs, members)]
    ens = A.LSRKEnsemble(dgs, qs, [m[2] for m in members])
    ens.run(5)
    for (dg, q0, dt), q in zip(members, qs):
        alone = dg.fieldarray(q0)
        A.LSRK54(dg, alone, dt).run(alone, 5)
        dg.ctx.sync()
        assert np.array_equal(q.numpy(), alone.numpy())


def _bubble_member(A, N, Kd, sflux="roe"):
    v = P.julia_range(0, 2000, Kd + 1)
    g = A.brickgrid(A.LobattoCell(N, 3), (v, v, v), (False,) * 3)
    law = A.EulerGravityLaw(3, grav=9.81, problem=A.Problem(boundary="reflect"))
    flux = A.RoeFlux() if sflux == "roe" else A.RusanovFlux()
    dg = A.DGSEM(law=law, grid=g, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=flux, precision="fast")
    return dg, P.rising_bubble(g.points), (0.25 if N < 6 else 0.125) * (2000 / Kd) / N / 330


@pytest.mark.parametrize("sflux", ["roe", "rusanov"])
def test_batch_fast_euler_gravity_3d_all_degrees_one_launch(sflux):
    """3-D bubbles of N = 3..7 (FAST EC volume) in ONE launch per stage: the 3-D N = 6
    member, which alone runs two elements in a 768-thread workgroup, takes one element
    per workgroup here; the smaller members fill the launch's 512 threads; each member
    ends bitwise where it would alone."""
    import atum_hip as A
    members = [_bubble_member(A, N, Kd, sflux) for N, Kd in ((3, 3), (4, 2), (5, 2), (6, 2), (7, 1))]
    dgs = [m[0] for m in members]
    qs = [dg.fieldarray(m[1]) for dg, m in zip(dgs, members)]
    ens = A.LSRKEnsemble(dgs, qs, [m[2] for m in members])
    assert ens.batch.ngroups == 1
    ens.run(6)
    for (dg, q0, dt), q in zip(members, qs):
        alone = dg.fieldarray(q0)
        A.LSRK54(dg, alone, dt).run(alone, 6)
        dg.ctx.sync()
        assert np.array_equal(q.numpy(), alone.numpy()), dg.grid.N
