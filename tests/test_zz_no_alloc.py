"""atm_lsrk_run_noalloc: the LSRK time loop with its ping-pong state in the
library's fixed device array (no allocation in the call) gives bit for bit what
atm_lsrk_run gives with a caller-supplied scratch field, on two calls in a row,
the second with a smaller state; states larger than the fixed array are refused."""
import ctypes

import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def _bubble(A, Kd=6, N=4, geometry="host"):
    v = P.julia_range(0, 2000, Kd + 1)
    grid = A.brickgrid(A.LobattoCell(N, 3), (v, v, v), (False,) * 3, lazy=(geometry == "device"))
    law = A.EulerGravityLaw(3, grav=9.81, problem=A.Problem(boundary="reflect"))
    dg = A.DGSEM(law=law, grid=grid, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RoeFlux(), precision="fast", geometry=geometry)
    return dg, P.rising_bubble(grid.points), 0.25 * (2000 / Kd) / N / 330


def _vortex(A, K=4, N=3):
    v = P.julia_range(-5, 5, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    dg = A.DGSEM(law=A.EulerLaw(2), grid=grid, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RusanovFlux())
    return dg, P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0)), 0.1 * (10 / K) / N


def _both(A, dg, q0, dt, scheme, nsteps):
    L = A._lib.lib
    qa, dqa = dg.fieldarray(q0), dg.fieldarray()
    rc = L.atm_lsrk_run_noalloc(dg.h, scheme, qa.ptr, dqa.ptr, float(dt), 0.0, int(nsteps))
    dg.ctx.check(rc, "atm_lsrk_run_noalloc")
    qb, dqb, scr = dg.fieldarray(q0), dg.fieldarray(), dg.fieldarray()
    dg.ctx.check(L.atm_lsrk_run(dg.h, scheme, qb.ptr, dqb.ptr, scr.ptr, float(dt), 0.0, int(nsteps), 1),
                 "atm_lsrk_run")
    dg.ctx.sync()
    return (qa.numpy(), dqa.numpy()), (qb.numpy(), dqb.numpy())


def test_noalloc_equals_lsrk_run_twice_second_smaller():
    import atum_hip as A
    big = _bubble(A)
    small = _vortex(A)
    assert small[1].size < big[1].size
    # first call: the larger state, an odd step count (the state ends in the fixed
    # array and is copied back), LSRK54
    (qa, dqa), (qb, dqb) = _both(A, *big, scheme=0, nsteps=3)
    assert np.array_equal(qa, qb) and np.array_equal(dqa, dqb)
    assert not np.array_equal(qa, big[1])
    # second call: a smaller state on another operator, LSRK144, an even step count
    (qa, dqa), (qb, dqb) = _both(A, *small, scheme=1, nsteps=4)
    assert np.array_equal(qa, qb) and np.array_equal(dqa, dqb)
    assert not np.array_equal(qa, small[1])


def test_noalloc_refuses_states_beyond_the_fixed_array():
    import atum_hip as A
    L = A._lib.lib
    # 26^3 elements, N = 4, 5 states: 10,985,000 doubles > 8,388,608
    dg, _, dt = _bubble(A, Kd=26, geometry="device")
    n = dg.Ns * dg.grid.Ntot
    assert n > 8388608
    q, dq = ctypes.c_void_p(), ctypes.c_void_p()
    dg.ctx.check(L.atm_alloc(dg.ctx.h, n * 8, ctypes.byref(q)), "atm_alloc")
    dg.ctx.check(L.atm_alloc(dg.ctx.h, n * 8, ctypes.byref(dq)), "atm_alloc")
    try:
        rc = L.atm_lsrk_run_noalloc(dg.h, 0, q, dq, float(dt), 0.0, 1)
        assert rc == -1  # ATM_ERR_ARG, before any launch
        msg = L.atm_last_error(dg.ctx.h).decode()
        assert "fixed scratch" in msg and "8388608" in msg
    finally:
        L.atm_free(dg.ctx.h, q)
        L.atm_free(dg.ctx.h, dq)
