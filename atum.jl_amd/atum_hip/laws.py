"""Balance laws and numerical fluxes: the plug-in interface of the reference.

Law objects carry the constants the reference bakes into type parameters
(src/balancelaw.jl:6-15) and the problem plug-ins it dispatches on
(`boundarystate`, `source!`, `referencestate`).  Arbitrary Julia methods cannot
cross a C ABI, so problem plug-ins are chosen from the enumerated set the
reference's tests and experiments use; `referencestate` (which only feeds the
auxiliary state) is any host callable.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np


# --- numerical fluxes (src/numericalfluxes.jl:31-47) ------------------------
class AbstractNumericalFlux:
    name = ""

    def __repr__(self):
        return type(self).__name__ + "()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)


class CentralFlux(AbstractNumericalFlux):
    name = "central"


class RusanovFlux(AbstractNumericalFlux):
    name = "rusanov"


class EntropyConservativeFlux(AbstractNumericalFlux):
    name = "ec"


class KennedyGruberFlux(AbstractNumericalFlux):
    name = "kg"


class RoeFlux(AbstractNumericalFlux):
    name = "roe"


class MatrixFlux(AbstractNumericalFlux):
    name = "matrix"


# --- problems (AbstractProblem, src/balancelaw.jl:8-10) ----------------------
@dataclass
class Problem:
    """Problem plug-ins.

    boundary:  "copy" (src/balancelaw.jl:21), "reflect"
               (experiment/euler_gravity/risingbubble.jl:10-15), "zero_inflow"
               (test/advection/imex_wave_2d.jl:19-26) or "prescribed" with
               `bc_states[tag-1]` (experiment/euler/sod.jl:20-23).
    source:    None or "coriolis" with rotation rate `omega`
               (experiment/euler_gravity/baroclinicwave.jl:27-34).
    referencestate: callable x (dim, n) -> (rho_ref, p_ref) for the gravity laws.
    """
    boundary: str = "copy"
    bc_states: Optional[np.ndarray] = None
    source: Optional[str] = None
    omega: float = 0.0
    referencestate: Optional[Callable] = None


DummyProblem = Problem


class AbstractBalanceLaw:
    law = ""
    naux_min = 0

    def __init__(self, dim: int, problem: Optional[Problem] = None):
        if dim not in (1, 2, 3):
            raise ValueError("dim must be 1, 2 or 3")
        self.dim = dim
        self.problem = problem or Problem()

    @property
    def nstates(self) -> int:
        return self.dim + 2

    def auxiliary(self, x):
        """auxiliary(law, x) (src/balancelaw.jl:17): no aux state by default."""
        return None

    def __repr__(self):
        return f"{type(self).__name__}(dim={self.dim})"


class AdvectionLaw(AbstractBalanceLaw):
    """src/balancelaws/advection.jl:8-14."""
    law = "advection"

    def __init__(self, dim, u=None, problem=None):
        super().__init__(dim, problem)
        self.u = tuple(float(v) for v in (u if u is not None else (1.0,) * dim))
        if len(self.u) != dim:
            raise ValueError("velocity needs one component per dimension")

    @property
    def nstates(self):
        return 1


class EulerLaw(AbstractBalanceLaw):
    """src/balancelaws/euler.jl:9-17."""
    law = "euler"

    def __init__(self, dim, gamma=1.4, problem=None):
        super().__init__(dim, problem)
        self.gamma = float(gamma)

    def pressure(self, rho, rhou, rhoe):
        ke = sum(m * m for m in rhou)
        return (self.gamma - 1) * (rhoe - ke / (2 * rho))

    def energy(self, rho, rhou, p):
        ke = sum(m * m for m in rhou)
        return p / (self.gamma - 1) + ke / (2 * rho)

    def soundspeed(self, rho, p):
        return np.sqrt(self.gamma * p / rho)


class EulerGravityLaw(AbstractBalanceLaw):
    """src/balancelaws/euler_gravity.jl:10-23; aux = [Φ, ∇Φ, ρ_ref, p_ref] (:42-55)."""
    law = "euler_gravity"

    def __init__(self, dim, gamma=1.4, grav=9.81, sphere=False, pde_level_balance=False,
                 problem=None):
        super().__init__(dim, problem)
        self.gamma, self.grav = float(gamma), float(grav)
        self.sphere, self.pde_level_balance = bool(sphere), bool(pde_level_balance)
        if self.sphere:
            raise NotImplementedError("spherical geopotential needs a cubed-sphere grid")

    naux_min = property(lambda self: 1 + self.dim)

    def auxiliary(self, x):
        z = x[-1]
        rows = [self.grav * z]
        for d in range(self.dim):
            rows.append(np.full_like(z, self.grav if d == self.dim - 1 else 0.0))
        ref = self.problem.referencestate
        if ref is not None:
            rho, p = ref(x)
            rows += [rho, p]
        return np.stack(rows)

    def pressure(self, rho, rhou, rhoe, phi):
        ke = sum(m * m for m in rhou)
        return (self.gamma - 1) * (rhoe - ke / (2 * rho) - rho * phi)


class LinearEulerGravityLaw(AbstractBalanceLaw):
    """src/balancelaws/linear_euler_gravity.jl: linearisation about the reference state."""
    law = "linear_euler_gravity"

    def __init__(self, parent: EulerGravityLaw):
        super().__init__(parent.dim, parent.problem)
        self.parent = parent
        self.gamma, self.grav = parent.gamma, parent.grav
        self.pde_level_balance = False

    naux_min = property(lambda self: 3 + self.dim)

    def auxiliary(self, x):
        return self.parent.auxiliary(x)


class ShallowWaterLaw(AbstractBalanceLaw):
    """src/balancelaws/shallow_water.jl:9-17 (default grav = 10)."""
    law = "shallow_water"

    def __init__(self, dim, grav=10.0, problem=None):
        super().__init__(dim, problem)
        self.grav = float(grav)
