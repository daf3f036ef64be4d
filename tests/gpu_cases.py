"""Shared set-ups for the GPU parity tests: the same grid, auxiliary state and
input state are handed to the HIP library (through atum_hip) and to the CPU
oracle."""
from __future__ import annotations

import numpy as np

import problems as P
from oracle import oracle as O

FORMS = {"weak": None, "central": "central", "ec": "ec", "kg": "kg"}


def make_state(law, x, rng, aux=None, amp=0.05):
    """A smooth state with random modal content (positive density/pressure)."""
    dim = x.shape[0]
    n = x.shape[1]
    span = [max(np.ptp(x[d]), 1e-12) for d in range(dim)]
    xs = [(x[d] - x[d].min()) / span[d] for d in range(dim)]

    def smooth():
        f = np.zeros(n)
        for _ in range(3):
            k = rng.integers(1, 3, size=dim)
            ph = rng.random(dim) * 2 * np.pi
            term = np.ones(n)
            for d in range(dim):
                term = term * np.sin(2 * np.pi * k[d] * xs[d] + ph[d])
            f += rng.standard_normal() * term
        return f / 3

    if law == "advection":
        return (2.0 + smooth())[None, :]
    if law in ("euler_gravity", "linear_euler_gravity"):
        rho0, p0 = aux[dim + 1], aux[dim + 2]
        rho = rho0 * (1 + amp * smooth())
        u = [amp * 10 * smooth() for _ in range(dim)]
        p = p0 * (1 + amp * smooth())
        rows = [rho] + [rho * ud for ud in u]
        ke = sum((rho * ud) ** 2 for ud in u) / (2 * rho)
        rows.append(p / 0.4 + ke + rho * aux[0])
        return np.stack(rows)
    rho = 1.5 + 0.5 * smooth()
    u = [0.5 * smooth() for _ in range(dim)]
    rows = [rho] + [rho * ud for ud in u]
    if law == "shallow_water":
        rows.append(rho * smooth())
    else:
        p = 2.0 + 0.5 * smooth()
        rows.append(P.euler_energy(1.4, rho, [rho * ud for ud in u], p))
    return np.stack(rows)


class Case:
    """One DGSEM configuration built twice: HIP library and oracle."""

    def __init__(self, law, dim, N, K, form="weak", sflux="rusanov", bc="copy", periodic=None,
                 directions=None, ordering="cartesian", domain=None, source=None, omega=0.0,
                 bc_states=None, u=None, grav=9.81, seed=0, gravity_domain=False):
        import atum_hip as A
        self.law_name = law
        periodic = periodic if periodic is not None else (True,) * dim
        if domain is None:
            domain = [(0.0, 2000.0)] * dim if gravity_domain else [(-1.0, 1.0)] * dim
        Kd = K if isinstance(K, (list, tuple)) else [K] * dim
        coords = [P.julia_range(a, b, k + 1) for (a, b), k in zip(domain, Kd)]
        cell = A.LobattoCell(N, dim)
        self.grid = A.brickgrid(cell, coords, periodic, ordering)
        self.ogrid = O.Grid(N, coords, periodic, stacked=(ordering == "stacked"))
        x = self.grid.points
        rng = np.random.default_rng(seed)
        u = u if u is not None else tuple([1.0, -0.4, 0.7][:dim])
        self.u = u
        prob = A.Problem(boundary=bc, source=source, omega=omega,
                         bc_states=None if bc_states is None else np.asarray(bc_states))
        self.aux = None
        if law == "advection":
            alaw = A.AdvectionLaw(dim, u, problem=prob)
        elif law == "euler":
            alaw = A.EulerLaw(dim, problem=prob)
        elif law == "shallow_water":
            alaw = A.ShallowWaterLaw(dim, grav=grav, problem=prob)
        else:
            prob.referencestate = lambda xx: P.isothermal_reference(xx, grav=grav)
            eg = A.EulerGravityLaw(dim, grav=grav, problem=prob)
            alaw = eg if law == "euler_gravity" else A.LinearEulerGravityLaw(eg)
            self.aux = alaw.auxiliary(x)
        vf = FORMS[form]
        vform = A.WeakForm() if vf is None else A.FluxDifferencingForm(_flux(A, vf))
        self.dg = A.DGSEM(law=alaw, grid=self.grid, volume_form=vform,
                          surface_numericalflux=_flux(A, sflux), directions=directions,
                          auxstate=self.aux)
        self.odg = O.DG(law, self.ogrid, surface_flux=sflux,
                        volume_form="weak" if vf is None else "fd", volume_flux=vf or "central",
                        directions=directions, u=u, bc=bc, bc_states=bc_states, source=source,
                        omega=omega, aux=self.aux, grav=grav)
        self.q = make_state(law, x, rng, self.aux)
        self.rng = rng


def _flux(A, name):
    return {"central": A.CentralFlux(), "rusanov": A.RusanovFlux(), "ec": A.EntropyConservativeFlux(),
            "kg": A.KennedyGruberFlux(), "roe": A.RoeFlux(), "matrix": A.MatrixFlux()}[name]


def relmax(a, b):
    """Relative max-norm difference ||a - b||_inf / ||b||_inf."""
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)
