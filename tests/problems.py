"""Problem set-ups of the reference tests and experiments, restated in numpy.

Each function takes the node coordinates as an array of shape (dim, Ntot) and
returns the state as (Ns, Ntot), one row per state component, in the order of
the reference's state vectors.  Used to drive both the CPU oracle and the HIP
library from the same inputs.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def julia_range(a, b, length):
    """Values of Julia's ``range(a, stop=b, length=n)`` for Float64 endpoints.

    Julia computes the elements in twice precision, which yields the correctly
    rounded value of ``a + i*(b-a)/(n-1)``; the exact rational is rounded here.
    """
    fa, fb = Fraction(a), Fraction(b)
    n = length - 1
    return np.array([float(fa + (fb - fa) * i / n) for i in range(length)])


def julia_step(a, b, length):
    return float((Fraction(b) - Fraction(a)) / (length - 1))


def euler_energy(gamma, rho, rhou, p):
    """src/balancelaws/euler.jl:36-39: p/(γ-1) + ρu⃗'ρu⃗/2ρ."""
    ke = rhou[0] * rhou[0]
    for d in range(1, len(rhou)):
        ke = ke + rhou[d] * rhou[d]
    return p / (gamma - 1) + ke / (2 * rho)


# --- test/euler/isentropicvortex.jl:16-32 -------------------------------------
def isentropic_vortex(x, t, gamma=1.4, center=(5.0, 0.0), u0=(1.0, 0.0), beta=5.0):
    r0 = x[0] - center[0] - u0[0] * t
    r1 = x[1] - center[1] - u0[1] * t
    f = beta * np.exp(1 - (r0 * r0 + r1 * r1))
    rho = (1 - (gamma - 1) * f ** 2 / (16 * gamma * np.pi ** 2)) ** (1 / (gamma - 1))
    u = u0[0] + f / (2 * np.pi) * (-r1)
    v = u0[1] + f / (2 * np.pi) * r0
    rhou = [rho * u, rho * v]
    p = rho ** gamma
    rhoe = euler_energy(gamma, rho, rhou, p)
    return np.stack([rho, rhou[0], rhou[1], rhoe])


# --- test/euler/wave_{1,2,3}d.jl:16-24 ----------------------------------------
def euler_wave(x, t, gamma=1.4):
    dim = x.shape[0]
    u = [1.0] * dim
    sx = x[0]
    for d in range(1, dim):
        sx = sx + x[d]
    rho = 2 + np.sin(np.pi * (sx - float(sum(u)) * t))
    rhou = [rho * ud for ud in u]
    rhoe = euler_energy(gamma, rho, rhou, 1.0)
    return np.stack([rho] + rhou + [rhoe])


# --- test/advection/wave_{1,2,3}d.jl:16-19 ------------------------------------
def advection_wave(x, t, u):
    sx = x[0]
    for d in range(1, x.shape[0]):
        sx = sx + x[d]
    return (2 + np.sin(np.pi * (sx - float(sum(u)) * t)))[None, :]


# --- test/advection/imex_wave_2d.jl:28-36 -------------------------------------
def imex_gaussian(x, t, u):
    beta = 5
    mu = 0.2
    r = [np.mod(x[d] - u[d] * t + 1, 2) - 1 for d in range(x.shape[0])]
    rr = (-r[0]) * r[0]
    for d in range(1, len(r)):
        rr = rr + (-r[d]) * r[d]
    return (beta * np.exp(rr / mu ** 2))[None, :]


# --- test/*/entropy_conservation_1d.jl: square waves ---------------------------
def square_rho(x):
    return np.where(np.abs(x[0]) < 0.5, 2.0, 1.0)


def euler_square(x, gamma=1.4):
    rho = square_rho(x)
    rhou = [rho * 1.0]
    return np.stack([rho, rhou[0], euler_energy(gamma, rho, rhou, 1.0)])


def sw_square(x):
    rho = square_rho(x)
    return np.stack([rho, rho * 1.0, np.zeros_like(rho)])


# --- gravity problems ----------------------------------------------------------
CV_D = 719.0


def flat_aux(x, grav, ref=None):
    """src/balancelaws/euler_gravity.jl:42-55 (flat geopotential) + reference state."""
    dim = x.shape[0]
    z = x[-1]
    rows = [grav * z]
    for d in range(dim):
        rows.append(np.full_like(z, grav if d == dim - 1 else 0.0))
    if ref is not None:
        rows += list(ref)
    return np.stack(rows)


def linear_test_reference(x, gamma=1.4, grav=9.81):
    """test/euler_gravity/linear.jl:20-43: constant-θ hydrostatic state (ρ, p)."""
    z = x[-1]
    phi = grav * z
    cp_d = gamma * CV_D
    R_d = cp_d - CV_D
    theta = 300.0
    p0 = 1e5
    pi_exner = 1 - grav / (cp_d * theta) * z
    rho = p0 / (R_d * theta) * pi_exner ** (CV_D / R_d)
    T = theta * pi_exner
    rhoe = rho * (CV_D * T + phi)
    p = (gamma - 1) * (rhoe - rho * phi)  # EulerGravity pressure with ρu = 0
    return rho, p


def linear_test_initial(aux, gamma=1.4):
    """test/euler_gravity/linear.jl:45-53."""
    dim = aux.shape[0] - 3
    phi = aux[0]
    rho, p = aux[dim + 1], aux[dim + 2]
    rhoe = p / (gamma - 1) + rho * phi
    z = np.zeros_like(rho)
    return np.stack([rho] + [z] * dim + [rhoe])


def isothermal_reference(x, gamma=1.4, grav=9.81, T_ref=300.0, p_s=1e5):
    """Isothermal hydrostatic state (experiment/euler_gravity/acousticwave.jl:26-45
    pattern, flat geometry): ρ_ref = ρ_s exp(-g z / (R T)), p_ref = ρ_ref R T."""
    z = x[-1]
    cp_d = gamma * CV_D
    R_d = cp_d - CV_D
    delta = grav / (R_d * T_ref)
    rho_s = p_s / (T_ref * R_d)
    rho = rho_s * np.exp(-delta * z)
    p = rho * R_d * T_ref
    return rho, p


def rising_bubble(x, gamma=1.4, grav=9.81, center=(1000.0, 1000.0, 350.0), radius=250.0,
                  amplitude=2.0, T_ref=300.0, p_s=1e5):
    """Warm bubble on an isothermal hydrostatic background (3-D; walls).

    The background is the isothermal state; a Gaussian potential-temperature
    perturbation θ' = A exp(-r²/R²) at fixed pressure lowers the density
    (cf. experiment/euler_gravity/risingbubble.jl:17-51 for the 2-D original).
    """
    dim = x.shape[0]
    cp_d = gamma * CV_D
    R_d = cp_d - CV_D
    rho0, p = isothermal_reference(x, gamma, grav, T_ref, p_s)
    r2 = np.zeros_like(x[0])
    for d in range(dim):
        r2 = r2 + (x[d] - center[d]) ** 2
    T0 = p / (rho0 * R_d)
    exner = (p / p_s) ** (R_d / cp_d)
    theta = T0 / exner + amplitude * np.exp(-r2 / radius ** 2)
    T = theta * exner
    rho = p / (R_d * T)
    phi = grav * x[-1]
    rhoe = rho * (CV_D * T + phi)
    z = np.zeros_like(rho)
    return np.stack([rho] + [z] * dim + [rhoe])


# --- shallow-water ensemble (256 instances of uneven size and degree) ---------
def sw_ensemble_params(n=256, seed=20260101, kmin=8, kmax=64, nmin=3, nmax=7):
    """Per instance: K (elements per side), N, bump amplitude, width and centre,
    drawn with NumPy PCG64 from `seed`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        K = int(rng.integers(kmin, kmax + 1))
        N = int(rng.integers(nmin, nmax + 1))
        A = float(rng.uniform(0.01, 0.1))
        w = float(rng.uniform(0.05, 0.2))
        c = (float(rng.uniform(0, 1)), float(rng.uniform(0, 1)))
        out.append((K, N, A, w, c))
    return out


def sw_bump(x, A, w, c):
    """h = 1 + A exp(-|x - c|^2 / w^2), at rest, zero tracer."""
    r2 = (x[0] - c[0]) ** 2 + (x[1] - c[1]) ** 2
    h = 1 + A * np.exp(-r2 / w ** 2)
    z = np.zeros_like(h)
    return np.stack([h, z, z, z])


def sw_ensemble_dt(K, N, A, grav=9.81):
    """dt = (1/8) dx / N / sqrt(g h_max) (experiment/shallow_water/bickleyjet.jl:53-54 rule)."""
    return 0.125 * (1.0 / K) / N / np.sqrt(grav * (1 + A))


def theta_noise_state(grid, aux, seed, amplitude=0.1, gamma=1.4, p_s=1e5):
    """Isothermal background (aux rows dim+1, dim+2 = rho_ref, p_ref) at rest plus a
    seeded N(0, amplitude^2) potential-temperature noise at fixed pressure."""
    dim = grid.dim
    rho0, p0 = aux[dim + 1], aux[dim + 2]
    rng = np.random.Generator(np.random.PCG64(seed))
    cp, cv = gamma * CV_D, CV_D
    R = cp - cv
    T0 = p0 / (rho0 * R)
    exner = (p0 / p_s) ** (R / cp)
    theta = T0 / exner + amplitude * rng.standard_normal(grid.Ntot)
    T = theta * exner
    rho = p0 / (R * T)
    z = np.zeros_like(rho)
    return np.stack([rho] + [z] * dim + [rho * (cv * T + aux[0])])
