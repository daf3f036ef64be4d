"""The three flux-differencing summation orders of the reference
(FluxDifferencingForm kernel_type, src/dgsem.jl:135-212) in REFERENCE precision:
each against the oracle's kernel of the same order, tendency within 1e-12
relative max-norm; the orders differ in their last bits (so the choice is not
ignored); combinations the reference or the FAST kernel do not define are
refused at creation."""
import numpy as np
import pytest

from gpu_cases import Case, relmax

pytestmark = pytest.mark.gpu

CASES = [  # law, dim, N, volume flux, surface flux, extra Case arguments
    ("euler", 2, 4, "ec", "rusanov", {}),
    ("euler", 3, 3, "ec", "roe", {}),
    ("euler", 1, 5, "kg", "matrix", {}),
    ("advection", 2, 3, "central", "rusanov", {}),
    ("shallow_water", 2, 5, "ec", "ec", {}),
    ("euler_gravity", 3, 4, "ec", "roe", dict(bc="reflect", periodic=(False,) * 3, gravity_domain=True)),
    ("euler_gravity", 2, 6, "kg", "rusanov", dict(bc="reflect", periodic=(True, False), gravity_domain=True)),
]


def _with_kernel(c, kernel_type, directions=None, precision="reference"):
    import atum_hip as A
    vf = A.FluxDifferencingForm(c.dg.volume_form.volume_numericalflux, kernel_type=kernel_type)
    return A.DGSEM(law=c.dg.law, grid=c.grid, volume_form=vf, surface_numericalflux=c.dg.surface_numericalflux,
                   directions=directions if directions is not None else c.dg.directions, auxstate=c.aux,
                   precision=precision)


def _tend(dg, q, dq0=None):
    qf = dg.fieldarray(q)
    dqf = dg.fieldarray(dq0 if dq0 is not None else np.zeros_like(q))
    dg(dqf, qf, 0.0, increment=dq0 is not None)
    dg.ctx.sync()
    return dqf.numpy()


@pytest.mark.parametrize("kernel_type", ["per_dir_symmetric", "naive"])
@pytest.mark.parametrize("law,dim,N,vf,sf,kw", CASES)
def test_kernel_type_against_oracle(kernel_type, law, dim, N, vf, sf, kw):
    from oracle import oracle as O
    c = Case(law, dim, N, 2 if dim == 3 else 3, form=vf, sflux=sf, seed=N + dim, **kw)
    dg = _with_kernel(c, kernel_type)
    okw = dict(surface_flux=sf, volume_form="fd", volume_flux=vf, fd_kernel=kernel_type, u=c.u,
               bc=kw.get("bc", "copy"), aux=c.aux)
    odg = O.DG(law, c.ogrid, **okw)
    ref = odg.tendency(c.q)
    got = _tend(dg, c.q)
    assert relmax(got, ref) <= 1e-12
    dq0 = c.rng.standard_normal(c.q.shape)
    ref2 = odg.tendency(c.q, dq0.copy(), increment=True)
    assert relmax(_tend(dg, c.q, dq0), ref2) <= 1e-12


def test_kernel_types_differ_in_last_bits():
    """On the same state the three orders agree to round-off but not bit for bit:
    the kernel type reaches the kernel."""
    c = Case("euler", 3, 3, 2, form="ec", sflux="roe", seed=5)
    outs = {kt: _tend(_with_kernel(c, kt), c.q) for kt in ("per_dir", "per_dir_symmetric", "naive")}
    for a in outs:
        for b in outs:
            if a < b:
                assert not np.array_equal(outs[a], outs[b]), (a, b)
                assert relmax(outs[a], outs[b]) <= 1e-12


def test_kernel_type_lsrk_steps_against_oracle():
    """Fused LSRK54 stages with the symmetric order: 20 steps against the oracle."""
    import atum_hip as A
    from oracle import oracle as O
    c = Case("euler", 2, 4, 3, form="ec", sflux="rusanov", seed=9)
    dg = _with_kernel(c, "per_dir_symmetric")
    odg = O.DG("euler", c.ogrid, surface_flux="rusanov", volume_form="fd", volume_flux="ec",
               fd_kernel="per_dir_symmetric")
    dt = 0.01
    ref = c.q.copy()
    odg.lsrk_steps(ref, dt, 20)
    q = dg.fieldarray(c.q)
    A.LSRK54(dg, q, dt).run(q, 20)
    dg.ctx.sync()
    assert relmax(q.numpy(), ref) <= 1e-10


def test_kernel_type_refusals():
    import atum_hip as A
    c = Case("euler", 2, 3, 2, form="ec", sflux="rusanov", seed=1)
    with pytest.raises(A.AtumError, match="all directions"):
        _with_kernel(c, "naive", directions=(1,))
    with pytest.raises(A.AtumError, match="FAST"):
        _with_kernel(c, "per_dir_symmetric", precision="fast")
    with pytest.raises(ValueError):
        A.FluxDifferencingForm(A.EntropyConservativeFlux(), kernel_type="other")
    # partial directions are fine for the per-direction orders
    dg = _with_kernel(c, "per_dir_symmetric", directions=(2,))
    assert dg.h
