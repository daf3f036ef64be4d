"""Request streams for the serving loop: Poisson arrivals of vortex, shallow-water
and IMEX Euler-gravity requests (50/30/20), each with its own mesh, N and step count."""
from __future__ import annotations

import numpy as np

import problems as P


def _vortex(K, N):
    def build(cache):
        import atum_hip as A
        key = ("vortex", K, N)
        if key not in cache:
            v = P.julia_range(-5, 5, K + 1)
            grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True), lazy=True)
            dg = A.DGSEM(law=A.EulerLaw(2), grid=grid,
                         volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                         surface_numericalflux=A.RusanovFlux(), precision="fast", geometry="device")
            cache[key] = (grid, dg)
        grid, dg = cache[key]
        q0 = P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0))
        return dg, q0, 0.1 * (10 / K) / N
    return build


def _sw(K, N, amp, w, c):
    def build(cache):
        import atum_hip as A
        key = ("sw", K, N)
        if key not in cache:
            v = P.julia_range(0.0, 1.0, K + 1)
            grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True), lazy=True)
            dg = A.DGSEM(law=A.ShallowWaterLaw(2, grav=9.81), grid=grid, surface_numericalflux=A.RusanovFlux(),
                         geometry="device")
            cache[key] = (grid, dg)
        grid, dg = cache[key]
        return dg, P.sw_bump(grid.points, amp, w, c), P.sw_ensemble_dt(K, N, amp)
    return build


def _imex(KH, KV, N, seed):
    def build(cache):
        import atum_hip as A
        key = ("imex", KH, KV, N)
        if key not in cache:
            vx = P.julia_range(0.0, 20e3, KH + 1)
            vz = P.julia_range(0.0, 10e3, KV + 1)
            grid = A.brickgrid(A.LobattoCell(N, 3), (vx, vx, vz), (True, True, False), "stacked", lazy=True)
            prob = A.Problem(boundary="reflect", referencestate=lambda x: P.isothermal_reference(x))
            law = A.EulerGravityLaw(3, grav=9.81, problem=prob)
            ex = A.DGSEM(law=law, grid=grid, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                         surface_numericalflux=A.RusanovFlux(), precision="fast", geometry="device")
            lin = A.DGSEM(law=A.LinearEulerGravityLaw(law), grid=grid, surface_numericalflux=A.RusanovFlux(),
                          directions=(3,), auxstate=ex.auxstate, geometry="device")
            cache[key] = (grid, ex, lin)
        grid, ex, lin = cache[key]
        q0 = P.theta_noise_state(grid, ex.auxstate, seed)
        dt = 0.2 * grid.min_node_distance(dims=(0, 1)) / 330.0
        return ex, lin, q0, dt
    return build


def make_stream(duration=60.0, rate=50.0, seed=20260103, mix=(0.5, 0.3, 0.2), steps=(50, 500),
                vortex_k=(8, 128), sw_k=(8, 64), imex_kh=(8, 32), imex_kv=10, nrange=(3, 7)):
    """Poisson(rate) arrivals over `duration` seconds (PCG64 seed)."""
    from atum_hip.serve import Request
    rng = np.random.Generator(np.random.PCG64(seed))
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        u = rng.random()
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        if u < mix[0]:
            K = int(rng.integers(vortex_k[0], vortex_k[1] + 1))
            N = int(rng.integers(nrange[0], nrange[1] + 1))
            r = Request(rid, "vortex", ("vortex", K, N), nsteps, t, _vortex(K, N),
                        footprint=3 * 4 * K * K * (N + 1) ** 2 * 8)
        elif u < mix[0] + mix[1]:
            K = int(rng.integers(sw_k[0], sw_k[1] + 1))
            N = int(rng.integers(nrange[0], nrange[1] + 1))
            amp, w = float(rng.uniform(0.01, 0.1)), float(rng.uniform(0.05, 0.2))
            c = (float(rng.uniform(0, 1)), float(rng.uniform(0, 1)))
            r = Request(rid, "sw", ("sw", K, N), nsteps, t, _sw(K, N, amp, w, c),
                        footprint=3 * 4 * K * K * (N + 1) ** 2 * 8)
        else:
            KH = int(rng.integers(imex_kh[0], imex_kh[1] + 1))
            r = Request(rid, "imex", ("imex", KH, imex_kv, 3), nsteps, t, _imex(KH, imex_kv, 3, int(rng.integers(1 << 30))),
                        footprint=12 * 5 * KH * KH * imex_kv * 64 * 8)
        out.append(r)
        rid += 1
    return out


def oracle_result(r, cache=None):
    """The oracle's run of request `r` (same initial state, time step and step
    count; the operator the request's kind names).  Test infrastructure only."""
    from oracle import oracle as O
    built = r.build({} if cache is None else cache)
    if r.kind == "imex":
        _, KH, KV, N = r.key
        ex, _lin, q0, dt = built
        vx = P.julia_range(0.0, 20e3, KH + 1)
        vz = P.julia_range(0.0, 10e3, KV + 1)
        og = O.Grid(N, [vx, vx, vz], [True, True, False], stacked=True)
        kw = dict(bc="reflect", aux=ex.auxstate, grav=9.81)
        oex = O.DG("euler_gravity", og, volume_form="fd", volume_flux="ec", surface_flux="rusanov", **kw)
        olin = O.DG("linear_euler_gravity", og, surface_flux="rusanov", directions=(3,), **kw)
        q = np.array(q0, dtype=np.float64)
        if r.steps:
            O.solve_ark23(oex, olin, q, dt, 1e30, split_rhs=False, max_steps=r.steps)
        return q
    kind, K, N = r.key
    _dg, q0, dt = built
    q = np.array(q0, dtype=np.float64)
    if kind == "vortex":
        v = P.julia_range(-5, 5, K + 1)
        odg = O.DG("euler", O.Grid(N, [v, v], [True, True]), volume_form="fd", volume_flux="ec",
                   surface_flux="rusanov")
    else:
        v = P.julia_range(0.0, 1.0, K + 1)
        odg = O.DG("shallow_water", O.Grid(N, [v, v], [True, True]), surface_flux="rusanov", grav=9.81)
    if r.steps:
        odg.lsrk_steps(q, dt, r.steps)
    return q
