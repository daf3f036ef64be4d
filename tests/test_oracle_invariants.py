"""Pin the CPU oracle by what the mathematics fixes.

These cover what the reference's golden values leave open (Roe, Matrix and
Kennedy-Gruber fluxes, Euler-gravity, LSRK144, the banded solver) with checks
chosen so that a dropped term, a wrong sign or a transposed operand fails:
closed-form LGL operators, free-stream preservation (flux consistency),
discrete conservation, semi-discrete entropy conservation / stability,
brute-force Jacobians, dense linear algebra and temporal convergence orders.
"""
import itertools

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

import problems as P
from oracle import oracle as O


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_lgl_closed_forms(N):
    xi, w, D = O.lgl(N)
    # interior nodes are the roots of P_N'; end points are -1, 1
    dP = npleg.legder(np.eye(N + 1)[N])
    assert np.allclose(npleg.legval(xi[1:-1], dP), 0, atol=1e-12)
    assert xi[0] == -1 and xi[-1] == 1
    # weights 2 / (N (N+1) P_N(x)^2); quadrature exact to degree 2N-1
    PN = npleg.legval(xi, np.eye(N + 1)[N])
    assert np.allclose(w, 2 / (N * (N + 1) * PN ** 2), rtol=1e-14)
    for k in range(2 * N):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert np.dot(w, xi ** k) == pytest.approx(exact, abs=1e-14)
    # D differentiates polynomials of degree <= N exactly; D[0,0] = -N(N+1)/4
    for k in range(N + 1):
        assert np.allclose(D @ xi ** k, k * xi ** max(k - 1, 0) if k else 0, atol=1e-12)
    assert D[0, 0] == pytest.approx(-N * (N + 1) / 4) and D[N, N] == pytest.approx(N * (N + 1) / 4)
    # summation by parts: W D + (W D)^T = diag(-1, 0, ..., 0, 1)
    B = np.diag(w) @ D + (np.diag(w) @ D).T
    Bx = np.zeros_like(B)
    Bx[0, 0], Bx[-1, -1] = -1, 1
    assert np.allclose(B, Bx, atol=1e-13)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("stacked", [False, True])
def test_brickgrid_tables(dim, stacked):
    N = 3
    coords = [np.array([0.0, 0.5, 1.5, 2.0]), np.array([-1.0, 0.0, 3.0]), np.array([1.0, 2.0, 2.5, 4.0])]
    coords = coords[:dim]
    per = [True, False, True][:dim]
    g = O.Grid(N, coords, per, stacked=stacked)
    vol = np.prod([c[-1] - c[0] for c in coords])
    assert g.MJ.sum() == pytest.approx(vol, rel=1e-13)
    assert np.allclose(g.MJ * g.MJI, 1)
    Nfpf = g.Nfp // (2 * dim)
    F = g.faceix_m.reshape(g.K, 2 * dim, Nfpf)
    Fp = g.faceix_p.reshape(g.K, 2 * dim, Nfpf)
    tags = g.bctag.reshape(g.K, 2 * dim)
    fMJ = g.faceMJ.reshape(g.K, 2 * dim, Nfpf)
    nrm = g.normal.reshape(dim, g.K, 2 * dim, Nfpf)
    pts = g.points
    for e, f in itertools.product(range(g.K), range(2 * dim)):
        r, side = divmod(f, 2)
        # minus nodes sit on the face; normal is outward along r
        assert np.all(nrm[r, e, f] == (1.0 if side else -1.0))
        if tags[e, f] == 0:
            # neighbour nodes coincide with minus nodes (modulo periodic shift) in the other dims
            for d in range(dim):
                if d != r:
                    assert np.allclose(pts[d, F[e, f]], pts[d, Fp[e, f]])
            # the neighbour's matching face points back
            en = Fp[e, f][0] // g.Np
            fo = 2 * r + (1 - side)
            assert sorted(Fp[en, fo]) == sorted(F[e, f])
        else:
            assert not per[r] and tags[e, f] == f + 1
    # total boundary area of each non-periodic face of the box
    for r in range(dim):
        if per[r]:
            continue
        for side in range(2):
            f = 2 * r + side
            sel = tags[:, f] == f + 1
            area = np.prod([coords[d][-1] - coords[d][0] for d in range(dim) if d != r])
            assert fMJ[sel, f].sum() == pytest.approx(area, rel=1e-13)


def _random_state(law, dim, x, rng, aux=None):
    n = x.shape[1]
    rho = 1.5 + 0.3 * rng.random(n)
    rows = [rho]
    if law == "advection":
        return rho[None, :].copy()
    for d in range(dim):
        rows.append(rho * 0.2 * rng.standard_normal(n))
    if law == "shallow_water":
        rows.append(0.1 * rng.random(n))
    else:
        e = 3.0 + 0.5 * rng.random(n)
        if law in ("euler_gravity", "linear_euler_gravity"):
            e = e + rho * aux[0]
        rows.append(e)
    return np.stack(rows)


LAW_FLUXES = {
    "advection": (["central"], ["central", "rusanov"]),
    "euler": (["central", "ec", "kg"], ["central", "rusanov", "ec", "kg", "roe", "matrix"]),
    "euler_gravity": (["central", "ec", "kg"], ["central", "rusanov", "ec", "kg", "roe", "matrix"]),
    "linear_euler_gravity": (["central"], ["central", "rusanov"]),
    "shallow_water": (["central", "ec"], ["central", "rusanov", "ec", "roe"]),
}


def _mkdg(law, g, vf, sf, **kw):
    aux = None
    if law in ("euler_gravity", "linear_euler_gravity"):
        # a geopotential that is smooth across periodic boundaries
        z = g.points[-1]
        rho_ref, p_ref = P.isothermal_reference(g.points * 100.0)
        rows = [2.0 * np.sin(np.pi * z)]
        rows += [np.zeros_like(z)] * (g.dim - 1) + [2.0 * np.pi * np.cos(np.pi * z)]
        aux = np.stack(rows + [rho_ref, p_ref])
    form = "weak" if vf is None else "fd"
    dg = O.DG(law, g, volume_form=form, volume_flux=vf or "central", surface_flux=sf, aux=aux,
              grav=9.81, **kw)
    return dg, aux


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("law", ["advection", "euler", "shallow_water"])
def test_free_stream(dim, law):
    """A constant state has zero tendency for every consistent volume/surface flux pair."""
    v = P.julia_range(-1, 1, 4)
    g = O.Grid(2, [v] * dim, [True] * dim)
    vfs, sfs = LAW_FLUXES[law]
    q0 = _random_state(law, dim, np.zeros((dim, 1)), np.random.default_rng(1))
    q = np.repeat(q0, g.Ntot, axis=1)
    for vf in [None] + vfs:
        for sf in sfs:
            dg, _ = _mkdg(law, g, vf, sf, u=(1.0, 0.4, -0.7))
            dq = dg.tendency(q)
            assert np.abs(dq).max() < 1e-11 * max(1.0, np.abs(q).max()), (vf, sf)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("law", ["advection", "euler", "euler_gravity", "shallow_water"])
def test_discrete_conservation(dim, law):
    """Periodic domain: sum MJ * dq is zero for mass (and energy where conserved)."""
    v = P.julia_range(-1, 1, 4)
    g = O.Grid(3, [v] * dim, [True] * dim)
    rng = np.random.default_rng(2)
    vfs, sfs = LAW_FLUXES[law]
    for vf in [None] + vfs:
        for sf in sfs:
            dg, aux = _mkdg(law, g, vf, sf, u=(1.0, 0.4, -0.7))
            q = _random_state(law, dim, g.points, rng, aux)
            dq = dg.tendency(q)
            scale = np.sum(g.MJ * np.abs(dq[0])) + 1e-300
            assert abs(np.sum(g.MJ * dq[0])) < 1e-13 * scale, (vf, sf)
            if law in ("euler", "shallow_water"):
                se = np.sum(g.MJ * np.abs(dq[-1]))
                assert abs(np.sum(g.MJ * dq[-1])) < 1e-13 * se, (vf, sf)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("law", ["euler", "euler_gravity", "shallow_water"])
def test_semidiscrete_entropy(dim, law):
    """EC volume + EC surface conserves entropy: <v(q), L(q)>_MJ = 0 (src/measures.jl:28-33);
    the Matrix surface flux (Euler laws) is entropy stable: <v(q), L(q)> < 0."""
    v = P.julia_range(-1, 1, 4)
    g = O.Grid(3, [v] * dim, [True] * dim)
    rng = np.random.default_rng(3)
    dg, aux = _mkdg(law, g, "ec", "ec")
    q = _random_state(law, dim, g.points, rng, aux)
    dq = dg.tendency(q)
    scale = np.sum(g.MJ * np.abs(dq).sum(0))
    assert abs(dg.entropyproduct(q, dq)) < 1e-14 * scale
    if law != "shallow_water":
        dgm, _ = _mkdg(law, g, "ec", "matrix")
        dqm = dgm.tendency(q)
        assert dgm.entropyproduct(q, dqm) < -1e-3 * scale


# (the reference defines no Matrix flux for shallow water)
@pytest.mark.parametrize("law,sf", [(law, sf) for law in ("euler", "euler_gravity", "shallow_water")
                                    for sf in ("roe", "rusanov", "matrix")
                                    if not (law == "shallow_water" and sf == "matrix")])
def test_surface_flux_antisymmetry(law, sf):
    """Two elements, periodic: swapping the sides of a face flips the numerical flux, so the
    face contributions of the two elements cancel in the sum of MJ*dq for every state."""
    v = np.array([0.0, 1.0, 2.0])
    g = O.Grid(2, [v], [True])
    rng = np.random.default_rng(4)
    dg, aux = _mkdg(law, g, "central", sf)
    q = _random_state(law, 1, g.points, rng, aux)
    if law == "euler_gravity":
        # Roe uses the minus-side geopotential for both pressures; keep Phi constant
        dg, aux = _mkdg(law, O.Grid(2, [v], [True]), "central", sf)
    dq = dg.tendency(q)
    tot = (g.MJ[None, :] * dq).sum(1)
    assert np.allclose(tot[:2], 0, atol=1e-12 * np.abs(dq).max())


def test_band_equals_dense_jacobian():
    """The probed band equals the brute-force Jacobian of the linear tendency (one unit
    vector at a time) and the banded LU solve equals a dense solve."""
    v = P.julia_range(0, 10e3, 4)
    g = O.Grid(2, [v, v], [True, False], stacked=True)
    aux = P.flat_aux(g.points, 9.81, P.linear_test_reference(g.points))
    lin = O.DG("linear_euler_gravity", g, directions=(2,), surface_flux="rusanov", bc="reflect",
               aux=aux, grav=9.81)
    n, bw, nl = lin.band_shape()
    band = lin.banded_assemble()
    colelem, KV, ncols = g.colelem()
    Ns, Nq, Nh = lin.Ns, g.Nq, g.Np // g.Nq
    size = Ns * g.Ntot
    # dense Jacobian
    J = np.zeros((size, size))
    e = np.zeros((Ns, g.Ntot))
    for m in range(size):
        e.ravel()[m] = 1.0
        J[:, m] = lin.tendency(e).ravel()
        e.ravel()[m] = 0.0

    def gidx(line, r):
        c, h = divmod(line, Nh)
        s, rem = r % Ns, r // Ns
        k, ez = rem % Nq, rem // Nq
        return s * g.Ntot + h + Nh * k + g.Np * colelem[c * KV + ez]

    idx = np.array([[gidx(ln, r) for r in range(n)] for ln in range(nl)])
    # every nonzero of J couples unknowns of the same line within the band
    covered = np.zeros_like(J, dtype=bool)
    for ln in range(nl):
        sub = J[np.ix_(idx[ln], idx[ln])]
        for r in range(n):
            for c in range(max(0, r - bw), min(n, r + bw + 1)):
                assert band[ln, r, bw + c - r] == sub[r, c]
        covered[np.ix_(idx[ln], idx[ln])] = True
    assert np.all(J[~covered] == 0)
    # banded LU solve of (I - c L) x = b equals the dense solve
    scale = -37.0
    fac = lin.banded_factor(band.copy(), scale)
    b = np.random.default_rng(5).standard_normal((Ns, g.Ntot))
    x = lin.banded_solve(fac, b)
    xd = np.linalg.solve(np.eye(size) + scale * J, b.ravel()).reshape(Ns, g.Ntot)
    assert np.allclose(x, xd, rtol=1e-10, atol=1e-12 * np.abs(xd).max())


def test_linear_operator_is_linear():
    v = P.julia_range(0, 10e3, 3)
    g = O.Grid(3, [v, v], [True, False], stacked=True)
    aux = P.flat_aux(g.points, 9.81, P.linear_test_reference(g.points))
    lin = O.DG("linear_euler_gravity", g, directions=(2,), surface_flux="rusanov", bc="reflect",
               aux=aux, grav=9.81)
    rng = np.random.default_rng(6)
    a, b = rng.standard_normal((2, lin.Ns, g.Ntot))
    lhs = lin.tendency(2.0 * a - 3.0 * b)
    rhs = 2.0 * lin.tendency(a) - 3.0 * lin.tendency(b)
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12 * np.abs(rhs).max())


@pytest.mark.parametrize("scheme,order", [("LSRK54", 4), ("LSRK144", 4)])
def test_lsrk_temporal_order(scheme, order):
    """Error against a fine-dt run shrinks as dt^4 (both coefficient tables)."""
    v = P.julia_range(-1, 1, 5)
    g = O.Grid(4, [v], [True])
    dg = O.DG("euler", g, volume_form="fd", volume_flux="ec", surface_flux="rusanov")
    q0 = P.euler_wave(g.points, 0.0)
    T = 0.4
    ref = q0.copy()
    dg.solve_lsrk(ref, T / 512, T, scheme=scheme)
    errs = []
    for n in (8, 16, 32):
        q = q0.copy()
        dg.solve_lsrk(q, T / n, T, scheme=scheme)
        errs.append(np.abs(q - ref).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > order - 0.5), rates


def test_ark23_temporal_order():
    v = P.julia_range(-1, 1, 5)
    g = O.Grid(3, [v, v], [True, False], stacked=True)
    u = (1.0, 0.3)
    kw = dict(surface_flux="rusanov", u=u, bc="zero_inflow")
    dg = O.DG("advection", g, directions=(1,), **kw)
    lin = O.DG("advection", g, directions=(2,), **kw)
    q0 = P.imex_gaussian(g.points, 0.0, u)
    T = 0.25
    ref = q0.copy()
    O.solve_ark23(dg, lin, ref, T / 1024, T)
    errs = []
    for n in (16, 32, 64):
        q = q0.copy()
        O.solve_ark23(dg, lin, q, T / n, T)
        errs.append(np.abs(q - ref).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 1.7), rates


def test_fd_kernel_variants_agree():
    """:per_dir, :per_dir_symmetric and :naive compute the same sum (round-off apart)."""
    v = P.julia_range(-1, 1, 3)
    g = O.Grid(3, [v, v, v], [True, True, True])
    rng = np.random.default_rng(7)
    out = []
    for kern in ("per_dir", "per_dir_symmetric", "naive"):
        dg = O.DG("euler", g, volume_form="fd", volume_flux="ec", surface_flux="roe", fd_kernel=kern)
        q = _random_state("euler", 3, g.points, np.random.default_rng(7))
        out.append(dg.tendency(q))
    del rng
    for o in out[1:]:
        assert np.abs(o - out[0]).max() <= 1e-12 * np.abs(out[0]).max()
