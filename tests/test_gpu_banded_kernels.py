"""Every solve kernel of the banded vertical solve (block form on the matrix cores
or the vector units, register-blocked, LDS and plain substitution) against the
oracle's banded solve of the same factors, for the degrees where each one runs
(Nq = 4..8, including the lines of Nq 7 and 8 that only the substitution kernels
take); unavailable kernels are refused; the tuner keeps one that works."""
import math

import numpy as np
import pytest

from gpu_cases import Case, relmax

pytestmark = pytest.mark.gpu

KERNELS = ("mfma", "block", "reg", "lds", "plain")


def _setup(N, KV=3):
    import atum_hip as A
    c = Case("linear_euler_gravity", 3, N, (2, 2, KV), form="weak", sflux="rusanov", bc="reflect",
             periodic=(True, True, False), directions=(3,), gravity_domain=True, ordering="stacked", seed=N)
    scale = -5.0 * (1 - 1 / math.sqrt(2))
    band = A.BandedOperator(c.dg, scale)
    band_orc = c.odg.banded_factor(c.odg.banded_assemble(), scale)
    return c, band, band_orc


@pytest.mark.parametrize("N", [3, 4, 5, 6, 7])
def test_every_available_kernel_against_oracle(N):
    import atum_hip as A
    c, band, band_orc = _setup(N)
    b = c.rng.standard_normal(c.q.shape)
    ref = c.odg.banded_solve(band_orc, b)
    ran = []
    for k in KERNELS:
        try:
            band.kernel = k
        except A.AtumError:
            assert k in ("mfma", "block", "reg") and N >= 6  # instantiated for Nq <= 6
            continue
        assert band.kernel == k
        x = c.dg.fieldarray()
        bf = c.dg.fieldarray(b)
        band.solve(x, bf)
        c.dg.ctx.sync()
        assert relmax(x.numpy(), ref) <= 1e-11, k
        ran.append(k)
    assert {"lds", "plain"} <= set(ran)
    if N <= 5:
        assert set(ran) == set(KERNELS)
    band.kernel = "auto"
    assert band.kernel == ran[0]  # the automatic choice is the first available


@pytest.mark.parametrize("N", [3, 7])
def test_autotune_keeps_a_working_kernel(N):
    c, band, band_orc = _setup(N)
    b = c.rng.standard_normal(c.q.shape)
    bf = c.dg.fieldarray(b)
    x = c.dg.fieldarray()
    best, times = band.autotune(x, bf, reps=3)
    assert best in times and band.kernel == best
    assert all(t > 0 for t in times.values())
    band.solve(x, bf)
    c.dg.ctx.sync()
    assert relmax(x.numpy(), c.odg.banded_solve(band_orc, b)) <= 1e-11


def test_ark23_steps_with_each_solve_kernel():
    """ARK23 with the banded factors solved by each kernel: the same states to round-off."""
    import atum_hip as A
    c, _, _ = _setup(3)
    dt = 5.0
    out = {}
    for k in KERNELS:
        q = c.dg.fieldarray(c.q)
        band = A.BandedOperator(c.dg, -dt * (1 - 1 / math.sqrt(2)))
        band.kernel = k
        A.ARK23(None, c.dg, q, dt, banded=band).run(q, 5)
        c.dg.ctx.sync()
        out[k] = q.numpy()
    for k in KERNELS[1:]:
        assert relmax(out[k], out["mfma"]) <= 1e-11, k
    assert np.all(np.isfinite(out["mfma"]))


def test_tuner_remembers_the_banded_kernel():
    import atum_hip as A
    c, band, _ = _setup(4)
    cache = A.TuneCache()
    best, times = A.autotune_banded(band, reps=2, cache=cache)
    assert times and best in times
    band2 = A.BandedOperator(c.dg, band.scale)
    hit, t2 = A.autotune_banded(band2, cache=cache)
    assert hit == best and t2 == {} and band2.kernel == best


@pytest.mark.parametrize("KV", [16, 18])
def test_block_form_solves_of_tall_columns(KV):
    """Columns of 16 elements (the block form's completely unrolled sequence, at its bound) and
    of 18 (its loop form) against the oracle."""
    c, band, band_orc = _setup(3, KV)
    b = c.rng.standard_normal(c.q.shape)
    ref = c.odg.banded_solve(band_orc, b)
    for k in ("mfma", "block"):
        band.kernel = k
        x = c.dg.fieldarray()
        band.solve(x, c.dg.fieldarray(b))
        c.dg.ctx.sync()
        assert relmax(x.numpy(), ref) <= 1e-11, (k, KV)


@pytest.mark.parametrize("N", [3, 4])
def test_every_kernel_solves_in_place(N):
    """x may alias the right-hand side (atm_banded_solve): each kernel solving in place gives the
    oracle's solution."""
    import atum_hip as A
    c, band, band_orc = _setup(N)
    b = c.rng.standard_normal(c.q.shape)
    ref = c.odg.banded_solve(band_orc, b)
    for k in KERNELS:
        try:
            band.kernel = k
        except A.AtumError:
            continue
        xb = c.dg.fieldarray(b)
        band.solve(xb, xb)
        c.dg.ctx.sync()
        assert relmax(xb.numpy(), ref) <= 1e-11, k


@pytest.mark.parametrize("dim,N,KV", [(2, 1, 5), (2, 2, 3), (3, 1, 4)])
def test_small_blocks_against_oracle(dim, N, KV):
    """Blocks of 8 to 12 rows (the matrix-core solve's smallest tiles, whose fragment reads run
    furthest past a block's data) with every available kernel against the oracle."""
    import atum_hip as A
    K = (3,) * (dim - 1) + (KV,)
    c = Case("linear_euler_gravity", dim, N, K, form="weak", sflux="rusanov", bc="reflect",
             periodic=(True,) * (dim - 1) + (False,), directions=(dim,), gravity_domain=True,
             ordering="stacked", seed=N)
    scale = -5.0 * (1 - 1 / math.sqrt(2))
    band = A.BandedOperator(c.dg, scale)
    ref = c.odg.banded_solve(c.odg.banded_factor(c.odg.banded_assemble(), scale),
                             b := c.rng.standard_normal(c.q.shape))
    ran = []
    for k in KERNELS:
        try:
            band.kernel = k
        except A.AtumError:
            continue
        x = c.dg.fieldarray()
        band.solve(x, c.dg.fieldarray(b))
        c.dg.ctx.sync()
        assert relmax(x.numpy(), ref) <= 1e-11, k
        ran.append(k)
    assert "mfma" in ran


def test_mfma_workgroup_shapes_give_the_same_bits():
    """The MFMA solve with 4 waves per workgroup (64 lines sharing each block's operators in
    LDS) and with 1 (16 lines) gives the same bits, matching the oracle; a shape whose LDS does
    not fit is refused; the tuner picks one of the two and the cache restores it."""
    import atum_hip as A
    c, band, band_orc = _setup(3)
    b = c.rng.standard_normal(c.q.shape)
    ref = c.odg.banded_solve(band_orc, b)
    band.kernel = "mfma"
    out = {}
    for w in (4, 1):
        band.waves = w
        assert band.waves == w
        x = c.dg.fieldarray()
        band.solve(x, c.dg.fieldarray(b))
        c.dg.ctx.sync()
        out[w] = x.numpy()
    assert np.array_equal(out[4], out[1])
    assert relmax(out[4], ref) <= 1e-11
    band.waves = 0
    assert band.waves == 4  # the largest that fits
    cache = A.TuneCache()
    best, times = A.autotune_banded(band, reps=2, cache=cache)
    assert best == "mfma" and band.waves in (1, 4)
    band2 = A.BandedOperator(c.dg, band.scale)
    hit, t2 = A.autotune_banded(band2, cache=cache)
    assert hit == "mfma" and t2 == {} and band2.waves == band.waves
    # a column of 18 elements (n = 360) leaves no room for four waves' forward solutions in LDS
    tall, tband, _ = _setup(3, 18)
    with pytest.raises(A.AtumError):
        tband.waves = 4
    assert tband.waves == 1
