"""Edge cases of the tendency and step path, through the C ABI against the oracle:
a single element that is its own neighbour across every periodic face, grids
one element wide in one direction, the smallest element on walls, and geometry
descriptors whose index tables the library must refuse on the host before any
kernel could follow them.

Tolerances as in the parity tests: tendency relative max-norm <= 1e-12 (FAST
mode on these generic states too), states after 20 LSRK54 steps <= 1e-10."""
import copy

import numpy as np
import pytest

from gpu_cases import Case, relmax

pytestmark = pytest.mark.gpu

TENDENCY_TOL = 1e-12
STATE_TOL = 1e-10


def _fast(case):
    import atum_hip as A
    d = case.dg
    return A.DGSEM(law=d.law, grid=case.grid, volume_form=d.volume_form,
                   surface_numericalflux=d.surface_numericalflux, auxstate=case.aux, precision="fast")


def _check(case, dg=None, steps=20, dt=1e-3):
    import atum_hip as A
    dg = dg or case.dg
    q = case.q
    ref = case.odg.tendency(q)
    qf = dg.fieldarray(q)
    dqf = dg.fieldarray()
    dg(dqf, qf, 0.0, increment=False)
    dg.ctx.sync()
    assert relmax(dqf.numpy(), ref) <= TENDENCY_TOL, relmax(dqf.numpy(), ref)
    qo = q.copy()
    case.odg.lsrk_steps(qo, dt, steps)
    qs = dg.fieldarray(q)
    A.LSRK54(dg, qs, dt).run(qs, steps)
    dg.ctx.sync()
    assert relmax(qs.numpy(), qo) <= STATE_TOL, relmax(qs.numpy(), qo)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("form,sflux", [("ec", "roe"), ("weak", "rusanov"), ("kg", "matrix")])
def test_single_periodic_element(dim, form, sflux):
    """K = 1, periodic: every face's neighbour is the element's own opposite face."""
    c = Case("euler", dim, 3, 1, form=form, sflux=sflux, seed=40 + dim)
    _check(c)
    if form == "ec":
        _check(c, _fast(c))


@pytest.mark.parametrize("N", [1, 7])
def test_single_element_walls(N):
    """One 3-D Euler-gravity element with reflective walls on all six faces, at the
    lowest and highest supported degree."""
    c = Case("euler_gravity", 3, N, 1, form="ec", sflux="roe", bc="reflect", periodic=(False,) * 3,
             gravity_domain=True, seed=50 + N)
    _check(c, dt=1e-2)
    _check(c, _fast(c), dt=1e-2)


@pytest.mark.parametrize("Kd,periodic", [([1, 6], (True, False)), ([6, 1], (False, True))])
def test_one_element_across(Kd, periodic):
    """A strip one element wide: the periodic direction wraps onto the same element,
    the other has walls (shallow water, weak form, as in the ensemble)."""
    c = Case("shallow_water", 2, 4, Kd, form="weak", sflux="rusanov", bc="reflect", periodic=periodic,
             seed=60 + Kd[0])
    _check(c)


def test_stacked_single_column():
    """A stacked grid of one column (1 x 1 x 5 elements), EC + Roe, walls at top and bottom."""
    c = Case("euler_gravity", 3, 3, [1, 1, 5], form="ec", sflux="roe", bc="reflect",
             periodic=(True, True, False), ordering="stacked", gravity_domain=True, seed=70)
    _check(c, dt=1e-2)


def _refused(case, name, mutate):
    import atum_hip as A
    g = copy.copy(case.grid)
    arr = getattr(case.grid, name).copy()
    mutate(arr)
    setattr(g, name, arr)
    d = case.dg
    with pytest.raises(A.AtumError):
        A.DGSEM(law=d.law, grid=g, volume_form=d.volume_form, surface_numericalflux=d.surface_numericalflux,
                auxstate=case.aux)


def test_bad_tables_refused_then_good_operator_works():
    """An out-of-range neighbour index, a face-point table in another order and a
    negative boundary tag are refused at creation; the context stays usable."""
    c = Case("euler", 2, 3, 3, form="ec", sflux="roe", seed=80)
    Ntot = c.grid.Np * c.grid.K

    def out_of_range(a):
        a[7] = Ntot

    def negative(a):
        a[3] = -1

    def swapped(a):
        a[0], a[1] = a[1], a[0]

    _refused(c, "faceix_p", out_of_range)
    _refused(c, "faceix_p", negative)
    _refused(c, "faceix_m", swapped)
    _refused(c, "bctag", negative)
    _check(c, steps=5)
