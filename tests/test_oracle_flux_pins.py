"""Checks that pin the oracle's Roe, Matrix and Kennedy-Gruber fluxes.

The reference's test suite holds no golden value for these fluxes, so each is
pinned here by a property that the mathematics fixes and that a plausible
transcription mistake (swapped wave strengths, a dropped absolute value, a
wrong eigenvalue scale, a dropped shear term, another mean of the sound speed,
a wrong gravity fluctuation weight) breaks.  tests/test_oracle_mutants.py
plants such mistakes in copies of the oracle and checks that this file fails
on every one of them.

  * Roe (src/balancelaws/euler.jl:91-138, euler_gravity.jl:149-198,
    shallow_water.jl:68-110): with the Roe averages of velocity and density,
    the mass and momentum rows (all rows for shallow water) of A(q+ - q-) equal
    F(q+).n - F(q-).n exactly, for any sound speed.  So when all wave speeds
    are positive the flux is F(q-).n, when all are negative F(q+).n; a pure
    shear jump is dissipated as |u.n| (q+ - q-) / 2 in every row; a pressure
    jump at zero normal velocity as (p+ - p-) / (2 c) in the mass row, c the
    Roe-weighted mean of the two sound speeds.
  * Roe, Matrix, Rusanov, EC, KG: to first order in q+ - q-, an upwind-type
    flux is F(q).n + (A - |A|)(q+ - q-)/2 with |A| = R |Lambda| R^-1 from an
    eigendecomposition (numpy) of the flux Jacobian A (central differences of
    the oracle's physical flux); Rusanov replaces |A| by the largest wave
    speed, EC and KG by 0.  The Matrix flux uses |u| in place of |u|^2
    (src/balancelaws/euler.jl:219), which is first-order exact when |u| = 1,
    the speed used here.
  * Matrix: (v+ - v-).(EC - Matrix) >= 0 pointwise (entropy dissipation).
  * KG (src/balancelaws/euler.jl:171-196): kinetic-energy preservation, so at
    uniform pressure on a periodic mesh the discrete kinetic energy rate
    sum MJ (u.d(rho u) - |u|^2/2 d rho) is zero to round-off.
  * KG with gravity (src/balancelaws/euler_gravity.jl:238-269): a hydrostatic
    state with rho linear and p quadratic in z (exactly differentiated at
    N >= 2) has zero tendency, which holds only with the fluctuation weight
    rho_avg / 2.
"""
import math

import numpy as np
import pytest

import problems as P
from oracle import oracle as O

GAMMA = 1.4
GRAV = 9.81


# ---------------------------------------------------------------------------
# states


def _unit(rng, dim):
    v = rng.standard_normal(dim)
    return v / np.linalg.norm(v)


def _euler_state(rho, u, p, law="euler", phi=0.0):
    u = np.asarray(u, dtype=float)
    rhou = rho * u
    rhoe = p / (GAMMA - 1) + rho * (u @ u) / 2
    if law == "euler_gravity":
        rhoe = rhoe + rho * phi
    return np.concatenate([[rho], rhou, [rhoe]])


def _sw_state(h, u, theta):
    u = np.asarray(u, dtype=float)
    return np.concatenate([[h], h * u, [h * theta]])


def _aux(law, dim, phi=500.0 * GRAV):
    if law == "euler_gravity":
        return np.concatenate([[phi], np.zeros(dim - 1), [GRAV], [1.0, 1.0]])
    return None


def _state(law, dim, rho, u, pr, phi=500.0 * GRAV, theta=0.3):
    if law == "shallow_water":
        return _sw_state(rho, u, theta)
    return _euler_state(rho, u, pr, law, phi)


def _sound(law, rho, p):
    return math.sqrt(GRAV * rho) if law == "shallow_water" else math.sqrt(GAMMA * p / rho)


def _nflux(law, dim, q, aux, n):
    return O.point_flux(law, dim, "flux", q, aux, gamma=GAMMA, grav=GRAV).T @ n


def _surface(law, dim, flux, n, qm, qp, aux):
    return O.point_flux(law, dim, "surface", qm, aux, qp, aux, flux=flux, n=n, gamma=GAMMA, grav=GRAV)


LAWS_DIMS = [("euler", 1), ("euler", 2), ("euler", 3), ("euler_gravity", 2), ("euler_gravity", 3),
             ("shallow_water", 1), ("shallow_water", 2)]


# ---------------------------------------------------------------------------
# Roe: exact identities


@pytest.mark.parametrize("law,dim", LAWS_DIMS)
@pytest.mark.parametrize("sign", [1, -1])
def test_roe_supersonic_is_upwind(law, dim, sign):
    """All wave speeds of one sign: the flux is the upwind physical flux in the
    rows where the Roe averages make A(q+ - q-) = F(q+).n - F(q-).n exact."""
    rng = np.random.default_rng(11 + dim + (7 if sign < 0 else 0))
    aux = _aux(law, dim)
    rows = slice(None) if law == "shallow_water" else slice(0, dim + 1)
    for _ in range(40):
        n = _unit(rng, dim)
        states = []
        for _side in range(2):
            rho = rng.uniform(0.5, 2.0)
            pr = rng.uniform(0.5, 2.0)
            c = _sound(law, rho, pr)
            # normal speed 2.5-4 sound speeds, plus a random tangential part
            u = sign * rng.uniform(2.5, 4.0) * c * n
            if dim > 1:
                t = rng.standard_normal(dim)
                t -= (t @ n) * n
                u = u + rng.uniform(0, 1.5) * c * t
            states.append(_state(law, dim, rho, u, pr, theta=rng.uniform(-1, 1)))
        qm, qp = states
        nf = _surface(law, dim, "roe", n, qm, qp, aux)
        up = _nflux(law, dim, qm if sign > 0 else qp, aux, n)
        scale = np.abs(up).max() + np.abs(_nflux(law, dim, qp if sign > 0 else qm, aux, n)).max()
        assert np.abs(nf[rows] - up[rows]).max() <= 1e-13 * scale


@pytest.mark.parametrize("law,dim", [("euler", 2), ("euler", 3), ("euler_gravity", 3), ("shallow_water", 2)])
def test_roe_shear_jump(law, dim):
    """Equal density and pressure, a jump in tangential velocity only: the
    dissipation is |u.n| (q+ - q-) / 2 in every row (the shear wave is an exact
    eigenvector, whatever the sound speed)."""
    rng = np.random.default_rng(21 + dim)
    aux = _aux(law, dim)
    for _ in range(40):
        n = _unit(rng, dim)
        t = rng.standard_normal(dim)
        t -= (t @ n) * n
        rho, pr = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        base = rng.uniform(-2, 2) * n + rng.uniform(-1, 1) * t
        du = rng.uniform(0.1, 1.0) * t
        qm = _state(law, dim, rho, base, pr, theta=0.4)
        qp = _state(law, dim, rho, base + du, pr, theta=0.4)
        central = (_nflux(law, dim, qm, aux, n) + _nflux(law, dim, qp, aux, n)) / 2
        nf = _surface(law, dim, "roe", n, qm, qp, aux)
        want = central - abs(base @ n) * (qp - qm) / 2
        # with gravity the energy carries rho*Phi >> p, and the pressure jump
        # (zero here) is recovered from it to ~1e-12 of rho*e only
        tol = 2e-12 if law == "euler_gravity" else 1e-13
        assert np.abs(nf - want).max() <= tol * np.abs(central).max()


@pytest.mark.parametrize("law,dim", [("euler", 1), ("euler", 2), ("euler", 3), ("euler_gravity", 3)])
def test_roe_acoustic_mass_dissipation(law, dim):
    """Zero normal velocity on both sides, different temperatures: the mass row
    is -(p+ - p-) / (2 c) with c = (sqrt(rho-) c- + sqrt(rho+) c+) / (sqrt(rho-) + sqrt(rho+)),
    and the normal momentum is the mean pressure."""
    rng = np.random.default_rng(31 + dim)
    aux = _aux(law, dim, FIRST_ORDER_PHI)
    for _ in range(40):
        n = _unit(rng, dim)
        sides = []
        for _s in range(2):
            rho, pr = rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0)
            u = np.zeros(dim)
            if dim > 1:
                t = rng.standard_normal(dim)
                u = t - (t @ n) * n
            sides.append((rho, pr, u))
        (rm, pm, um), (rp, pp, up) = sides
        qm = _state(law, dim, rm, um, pm, phi=FIRST_ORDER_PHI)
        qp = _state(law, dim, rp, up, pp, phi=FIRST_ORDER_PHI)
        nf = _surface(law, dim, "roe", n, qm, qp, aux)
        cm, cp = _sound(law, rm, pm), _sound(law, rp, pp)
        c = (math.sqrt(rm) * cm + math.sqrt(rp) * cp) / (math.sqrt(rm) + math.sqrt(rp))
        assert nf[0] == pytest.approx(-(pp - pm) / (2 * c), rel=1e-12, abs=1e-15)
        mom_n = nf[1:dim + 1] @ n
        assert mom_n == pytest.approx((pm + pp) / 2, rel=1e-12)


# ---------------------------------------------------------------------------
# first-order structure of every surface flux


def _jacobian(law, dim, q, aux, n):
    m = q.size
    A = np.zeros((m, m))
    for j in range(m):
        h = 1e-6 * max(1.0, abs(q[j]))
        e = np.zeros(m)
        e[j] = h
        A[:, j] = (_nflux(law, dim, q + e, aux, n) - _nflux(law, dim, q - e, aux, n)) / (2 * h)
    return A


def _abs_matrix(A):
    """|A| = A sign(A), the matrix sign function by Newton's iteration (robust for
    the repeated eigenvalue u.n of the flux Jacobian), and the spectral radius."""
    lam = np.linalg.eigvals(A)
    assert np.abs(lam.imag).max() <= 1e-8 * np.abs(lam).max()
    S = A.copy()
    for _ in range(100):
        S1 = (S + np.linalg.inv(S)) / 2
        done = np.abs(S1 - S).max() <= 1e-15 * np.abs(S1).max()
        S = S1
        if done:
            break
    assert np.abs(S @ S - np.eye(len(A))).max() <= 1e-8
    return A @ S, np.abs(lam.real).max()


def _base_state(law, dim, regime, rng):
    """A state with |u| = 1 and its normal: subsonic, or supersonic either way."""
    n = _unit(rng, dim)
    rho = rng.uniform(0.7, 1.5)
    if dim == 1:
        u = n.copy()
    else:
        t = rng.standard_normal(dim)
        t -= (t @ n) * n
        t /= np.linalg.norm(t)
        ang = 0.35 if regime != "subsonic" else 1.0
        u = math.cos(ang) * n + math.sin(ang) * t
    if regime == "supersonic_minus":
        u = -u
    if law == "shallow_water":
        c = 1.9 if regime == "subsonic" else 0.45
        rho = c * c / GRAV
        pr = None
    else:
        c = 1.9 if regime == "subsonic" else 0.45
        pr = rho * c * c / GAMMA
    return _state(law, dim, rho, u, pr, phi=FIRST_ORDER_PHI, theta=0.3), n


# a geopotential of the size of the other energies keeps the Jacobian well conditioned
FIRST_ORDER_PHI = 2.0
FIRST_ORDER = [("euler", 1, "roe"), ("euler", 2, "roe"), ("euler", 3, "roe"),
               ("euler", 1, "matrix"), ("euler", 2, "matrix"), ("euler", 3, "matrix"),
               ("euler_gravity", 2, "roe"), ("euler_gravity", 3, "roe"), ("euler_gravity", 3, "matrix"),
               ("shallow_water", 1, "roe"), ("shallow_water", 2, "roe"),
               ("euler", 3, "rusanov"), ("shallow_water", 2, "rusanov"),
               ("euler", 3, "ec"), ("euler", 2, "kg"), ("euler_gravity", 3, "kg"), ("shallow_water", 2, "ec")]


@pytest.mark.parametrize("law,dim,flux", FIRST_ORDER)
@pytest.mark.parametrize("regime", ["subsonic", "supersonic_plus", "supersonic_minus"])
def test_first_order_dissipation(law, dim, flux, regime):
    rng = np.random.default_rng([dim, len(law), len(flux), len(regime), FIRST_ORDER.index((law, dim, flux))])
    aux = _aux(law, dim, FIRST_ORDER_PHI)
    for _ in range(4):
        q, n = _base_state(law, dim, regime, rng)
        A = _jacobian(law, dim, q, aux, n)
        absA, lmax = _abs_matrix(A)
        if flux in ("ec", "kg"):
            diss = np.zeros_like(A)
        elif flux == "rusanov":
            diss = lmax * np.eye(q.size)
        else:
            diss = absA
        r = rng.standard_normal(q.size) * np.abs(q).clip(0.2)
        f0 = _nflux(law, dim, q, aux, n)
        errs = []
        for delta in (2e-3, 1e-3, 5e-4):
            nf = _surface(law, dim, flux, n, q, q + delta * r, aux)
            lin = f0 + (A - diss) @ (delta * r) / 2
            errs.append(np.abs(nf - lin).max())
        first = np.abs(diss @ r).max() + np.abs(A @ r).max()
        if errs[0] <= 1e-13 * first * 2e-3:
            continue  # exact to round-off (a flux linear in the jump, e.g. upwind shallow water)
        # second order: halving the jump quarters the residual ...
        assert errs[1] <= errs[0] / 3.2 and errs[2] <= errs[1] / 3.2, errs
        # ... and the residual is small against the first-order terms
        assert errs[2] <= 5e-2 * first * 5e-4


# ---------------------------------------------------------------------------
# Matrix: pointwise entropy dissipation


@pytest.mark.parametrize("law,dim", [("euler", 1), ("euler", 2), ("euler", 3), ("euler_gravity", 3)])
def test_matrix_dissipates_entropy_pointwise(law, dim):
    rng = np.random.default_rng(41 + dim)
    aux = _aux(law, dim)
    for _ in range(300):
        n = _unit(rng, dim)
        st = []
        for _s in range(2):
            rho, pr = rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0)
            u = rng.standard_normal(dim) * rng.uniform(0, 2.5)
            st.append(_state(law, dim, rho, u, pr))
        qm, qp = st
        vm = O.point_flux(law, dim, "entropyvariables", qm, aux, gamma=GAMMA, grav=GRAV)
        vp = O.point_flux(law, dim, "entropyvariables", qp, aux, gamma=GAMMA, grav=GRAV)
        ec = _surface(law, dim, "ec", n, qm, qp, aux)
        mt = _surface(law, dim, "matrix", n, qm, qp, aux)
        prod = (vp - vm) @ (ec - mt)
        assert prod >= -1e-13 * np.abs(vp - vm).max() * np.abs(ec).max()


# ---------------------------------------------------------------------------
# Kennedy-Gruber: kinetic energy and hydrostatic balance on whole meshes


@pytest.mark.parametrize("dim", [2, 3])
def test_kg_preserves_kinetic_energy(dim):
    """Periodic mesh, uniform pressure, random smooth density and velocity: with
    the KG volume and surface fluxes the kinetic energy does not change."""
    v = P.julia_range(-1, 1, 4)
    g = O.Grid(3, [v] * dim, [True] * dim)
    rng = np.random.default_rng(51 + dim)
    x = g.points

    def smooth():
        f = np.zeros(g.Ntot)
        for _ in range(3):
            k = rng.integers(1, 3, size=dim)
            ph = rng.random(dim) * 2 * np.pi
            term = np.ones(g.Ntot)
            for d in range(dim):
                term = term * np.sin(np.pi * k[d] * x[d] + ph[d])
            f += rng.standard_normal() * term
        return f

    rho = 1.5 + 0.3 * smooth()
    u = [0.5 * smooth() for _ in range(dim)]
    pr = np.full(g.Ntot, 1.7)
    q = np.stack([rho] + [rho * ud for ud in u] + [pr / (GAMMA - 1) + rho * sum(ud * ud for ud in u) / 2])
    dg = O.DG("euler", g, volume_form="fd", volume_flux="kg", surface_flux="kg", gamma=GAMMA)
    dq = dg.tendency(q)
    ke_rate = g.MJ * (sum(u[d] * dq[1 + d] for d in range(dim)) - sum(ud * ud for ud in u) / 2 * dq[0])
    scale = np.sum(g.MJ * (np.abs(rho) * sum(np.abs(ud) for ud in u)) * np.abs(dq[1:dim + 1]).sum(0))
    assert abs(ke_rate.sum()) <= 1e-13 * scale
    # the same state with a pressure perturbation does change the kinetic energy
    q2 = q.copy()
    q2[-1] += (0.2 * smooth()) / (GAMMA - 1)
    dq2 = dg.tendency(q2)
    ke2 = g.MJ * (sum(u[d] * dq2[1 + d] for d in range(dim)) - sum(ud * ud for ud in u) / 2 * dq2[0])
    assert abs(ke2.sum()) > 1e-6 * scale


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("N", [2, 4])
def test_kg_gravity_hydrostatic_balance(dim, N):
    """rho = r0 - a z, p = p0 - g (r0 z - a z^2 / 2), u = 0, Phi = g z: KG volume and
    surface fluxes give a zero tendency (walls at the bottom and top)."""
    vh = P.julia_range(0, 1000, 3)
    vz = P.julia_range(0, 1000, 4)
    coords = [vh] * (dim - 1) + [vz]
    g = O.Grid(N, coords, [True] * (dim - 1) + [False])
    z = g.points[-1]
    r0, a, p0 = 1.2, 4e-5, 1.0e5
    rho = r0 - a * z
    pr = p0 - GRAV * (r0 * z - a * z * z / 2)
    phi = GRAV * z
    aux = np.stack([phi] + [np.zeros_like(z)] * (dim - 1) + [np.full_like(z, GRAV)] +
                   [rho, pr])
    q = np.stack([rho] + [np.zeros_like(z)] * dim + [pr / (GAMMA - 1) + rho * phi])
    dg = O.DG("euler_gravity", g, volume_form="fd", volume_flux="kg", surface_flux="kg", bc="reflect",
              aux=aux, gamma=GAMMA, grav=GRAV)
    dq = dg.tendency(q)
    # scale: the pressure-gradient term, p0 / element height
    scale = p0 / (vz[1] - vz[0])
    assert np.abs(dq).max() <= 1e-11 * scale
