#!/usr/bin/env python3
"""Benchmark: FP64 DOF-updates/s per RK stage of the DGSEM tendency + LSRK54
step on one MI355X, on the 3-D Euler-with-gravity rising warm bubble:
16x16x16 hex elements on [0,2000]^3 m, N = 4, isothermal hydrostatic background
(T = 300 K, g = 9.81), Gaussian 2 K potential-temperature bubble of radius 250 m
at z = 350 m, no-flux walls, entropy-conservative volume flux, Roe surface flux.

A step is one LSRK54 time step (5 fused stage kernels: volume + surface +
stage update) over the whole mesh; inputs are resident in HBM before timing.
Timing: HIP events recorded on the library's own stream around exactly
--steps steps (replayed hipGraphs of two steps each).

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "atum.jl_amd"), ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def progress(msg):
    """A progress line on stderr (the JSON result stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_stage(dim, Nq, K, Ns, naux):
    """Canonical traffic of one stage in the reference data model: per node q and dq
    read+write (32*Ns), metric + MJ ((d^2+1)*8), aux (naux*8); per face point normal +
    faceMJ ((d+1)*8) and two int32 indices (8); per element 2d int32 tags."""
    Np = Nq ** dim
    Nfp = 2 * dim * Nq ** (dim - 1)
    per_elem = Np * (32 * Ns + (dim * dim + 1) * 8 + naux * 8) + Nfp * ((dim + 1) * 8 + 8) + 2 * dim * 4
    return per_elem * K


FP64_PEAK_TFS = 78.6  # MI355X FP64 (vector and matrix)


def ec_volume_flops_per_stage(dim, Nq, K, Ns, gravity):
    """FP64 operations of the EC flux-differencing volume term as the reference evaluates
    it (+, -, *, /, sqrt, log = 1 each): per node, direction and partner node (the
    node itself included) one two-point flux, 3d^2 + 16d + 48 (+15 with gravity), and
    its contraction with the metric weights, Ns (6d + 2)."""
    pair = 3 * dim * dim + 16 * dim + 48 + (15 if gravity else 0)
    contraction = Ns * (6 * dim + 2)
    return K * Nq ** dim * dim * Nq * (pair + contraction)


def bubble_setup(Kd, N, precision="fast"):
    import numpy as np  # noqa: F401
    import atum_hip as A
    import problems as P

    dim = 3
    v = P.julia_range(0.0, 2000.0, Kd + 1)
    grid = A.brickgrid(A.LobattoCell(N, dim), (v, v, v), (False, False, False))
    law = A.EulerGravityLaw(dim, grav=9.81, problem=A.Problem(boundary="reflect"))
    dg = A.DGSEM(law=law, grid=grid,
                 volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RoeFlux(), precision=precision)
    q0 = P.rising_bubble(grid.points)
    # the experiment/euler_gravity/risingbubble.jl:66-67 rule (1/3 dx / N / 330, for the 2-D
    # bubble) at CFL 1/4: in 3-D its CFL 1/3 is linearly unstable (the 16^3 bubble blows up
    # between steps 100 and 140); the cost of a stage does not depend on dt
    dt = 0.25 * (2000.0 / Kd) / N / 330.0
    return grid, dg, q0, dt


def cpu_baseline(Kd, N, stages):
    """The oracle (FP64 C++ restatement, OpenMP over elements) on the same workload,
    a bounded sample of `stages` LSRK54 stages."""
    import numpy as np
    import problems as P
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = O.set_threads(threads)
    v = P.julia_range(0.0, 2000.0, Kd + 1)
    g = O.Grid(N, [v, v, v], [False] * 3)
    aux = P.flat_aux(g.points, 9.81)
    dg = O.DG("euler_gravity", g, volume_form="fd", volume_flux="ec", surface_flux="roe",
              bc="reflect", aux=aux, grav=9.81)
    q = P.rising_bubble(g.points)
    dq = np.zeros_like(q)
    dt = 0.25 * (2000.0 / Kd) / N / 330.0  # as the timed run
    a = [0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
         -3550918686646 / 2091501179385, -1275806237668 / 842570457699]
    b = [1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
         1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
         2277821191437 / 14882151754819]
    t0 = time.perf_counter()
    for s in range(stages):
        dg.lsrk_stage(q, dq, a[s % 5], b[s % 5] * dt)
    el = time.perf_counter() - t0
    dof = q.size
    return {"value": dof * stages / el, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": f"{stages} LSRK54 stages of the same {Kd}^3 N={N} bubble on the host "
                      f"({el:.1f} s), FP64 C++ oracle with OpenMP over elements"}


def ensemble_workload(args, steps=None):
    """256 shallow-water instances (K ~ U{8..64}^2 elements, N ~ U{3..7}, weak form + Rusanov,
    periodic unit square, seeded Gaussian bumps), LSRK54 with per-instance dt: batched launches
    (one per N group and stage, hipGraph) against launching each instance's kernels alone."""
    import numpy as np
    import atum_hip as A
    import problems as P

    nsteps = args.steps if steps is None else steps

    params = P.sw_ensemble_params(n=args.instances)
    dgs, q0s, dts = [], [], []
    for (K, N, amp, w, c) in params:
        v = P.julia_range(0.0, 1.0, K + 1)
        grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
        dgs.append(A.DGSEM(law=A.ShallowWaterLaw(2, grav=9.81), grid=grid,
                           surface_numericalflux=A.RusanovFlux()))
        q0s.append(P.sw_bump(grid.points, amp, w, c))
        dts.append(P.sw_ensemble_dt(K, N, amp))
    ctx = dgs[0].ctx
    dof = sum(q.size for q in q0s)
    stages = 5 * nsteps
    tuned = None
    if args.ensemble_tune:  # one workgroup shape per launch group, chosen on the batched run itself
        tuned = A.autotune_batch(dgs, q0s, dts)
    qs = [dg.fieldarray(q) for dg, q in zip(dgs, q0s)]
    ens = A.LSRKEnsemble(dgs, qs, dts)
    ens.run(2)
    ctx.sync()
    ctx.record(0)
    ens.run(nsteps)
    ctx.record(1)
    ms_b = ctx.elapsed_ms(0, 1)
    batched = [q.numpy() for q in qs]
    if not all(np.all(np.isfinite(b)) for b in batched):
        raise SystemExit("bench: non-finite ensemble state")
    per = {}
    for graph in (False, True):
        qs2 = [dg.fieldarray(q) for dg, q in zip(dgs, q0s)]
        solvers = [A.LSRK54(dg, q, dt, use_graph=graph) for dg, q, dt in zip(dgs, qs2, dts)]
        for s_, q in zip(solvers, qs2):
            s_.run(q, 2)
        ctx.sync()
        ctx.record(2)
        for s_, q in zip(solvers, qs2):
            s_.run(q, nsteps)
        ctx.record(3)
        per[graph] = ctx.elapsed_ms(2, 3)
        ctx.sync()
        same = all(np.array_equal(b, q.numpy()) for b, q in zip(batched, qs2))
        if not same:
            raise SystemExit("bench: batched and per-request states differ")
    value = dof * stages / (ms_b / 1e3)
    return {
        "metric": "aggregate FP64 DOF-updates/s per RK stage (batched ensemble)",
        "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": nsteps, "warmup": 2,
        "ms_per_step": ms_b / nsteps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(dgs)} shallow-water instances, K~U{{8..64}}^2, N~U{{3..7}}, "
                               "weak form + Rusanov, LSRK54, per-instance dt (PCG64 seed 20260101)",
                   "dof": dof, "kernel_groups": ens.batch.ngroups,
                   "elements_per_workgroup": {f"N={k[2]}": v[0] for k, v in tuned.items()} if tuned else None},
        "per_request_eager": {"value": dof * stages / (per[False] / 1e3), "ms_per_step": per[False] / nsteps},
        "per_request_graph": {"value": dof * stages / (per[True] / 1e3), "ms_per_step": per[True] / nsteps},
        "speedup_vs_per_request_eager": per[False] / ms_b,
        "speedup_vs_per_request_graph": per[True] / ms_b,
        "bitwise_equal_to_alone": True,
    }


def ensemble3d_workload(args):
    """3-D Euler-gravity bubbles (FAST EC + Roe, walls) of N = 3..7 with 2^3..5^3 elements
    (PCG64 seed 20260104), LSRK54 with per-instance dt: all degrees in one launch per stage,
    against one batch per degree (one launch per degree and stage) and each instance alone."""
    import numpy as np
    import atum_hip as A
    import problems as P

    rng = np.random.Generator(np.random.PCG64(20260104))
    dgs, q0s, dts = [], [], []
    for _ in range(args.instances3d):
        N = int(rng.integers(3, 8))
        Kd = int(rng.integers(2, 6))
        v = P.julia_range(0.0, 2000.0, Kd + 1)
        g = A.brickgrid(A.LobattoCell(N, 3), (v, v, v), (False,) * 3)
        law = A.EulerGravityLaw(3, grav=9.81, problem=A.Problem(boundary="reflect"))
        dgs.append(A.DGSEM(law=law, grid=g, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                           surface_numericalflux=A.RoeFlux(), precision="fast"))
        q0s.append(P.rising_bubble(g.points))
        dts.append((0.25 if N < 6 else 0.125) * (2000.0 / Kd) / N / 330.0)
    ctx = dgs[0].ctx
    dof = sum(q.size for q in q0s)
    nsteps = args.steps
    stages = 5 * nsteps

    def timed(groups):
        qs_all, ens_all = [], []
        for idx in groups:
            qs = [dgs[i].fieldarray(q0s[i]) for i in idx]
            ens_all.append(A.LSRKEnsemble([dgs[i] for i in idx], qs, [dts[i] for i in idx]))
            qs_all.append(qs)
        for e in ens_all:
            e.run(2)
        ctx.sync()
        ctx.record(0)
        for e in ens_all:
            e.run(nsteps)
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1)
        ctx.sync()
        out = {}
        for idx, qs in zip(groups, qs_all):
            for i, q in zip(idx, qs):
                out[i] = q.numpy()
        return ms, out, [e.batch.ngroups for e in ens_all]

    ms_one, res_one, ng = timed([list(range(len(dgs)))])
    by_deg = {}
    for i, dg in enumerate(dgs):
        by_deg.setdefault(dg.grid.N, []).append(i)
    ms_deg, res_deg, _ = timed(list(by_deg.values()))
    qs2 = [dg.fieldarray(q) for dg, q in zip(dgs, q0s)]
    solvers = [A.LSRK54(dg, q, dt) for dg, q, dt in zip(dgs, qs2, dts)]
    for s_, q in zip(solvers, qs2):
        s_.run(q, 2)
    ctx.sync()
    ctx.record(2)
    for s_, q in zip(solvers, qs2):
        s_.run(q, nsteps)
    ctx.record(3)
    ms_alone = ctx.elapsed_ms(2, 3)
    ctx.sync()
    same = all(np.array_equal(res_one[i], res_deg[i]) and np.array_equal(res_one[i], q.numpy())
               for i, q in enumerate(qs2))
    if not same:
        raise SystemExit("bench: batched, per-degree and alone states differ")
    return {"metric": "aggregate FP64 DOF-updates/s per RK stage (batched 3-D ensemble)",
            "value": dof * stages / (ms_one / 1e3), "unit": "DOF-updates/s", "n_gpus": 1, "steps": nsteps,
            "warmup": 2, "ms_per_step": ms_one / nsteps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{len(dgs)} 3-D Euler-gravity bubbles, N~U{{3..7}}, K~U{{2..5}}^3, FAST EC + Roe, "
                                   "LSRK54 (PCG64 seed 20260104)", "dof": dof, "kernel_groups": ng[0]},
            "one_batch_per_degree": {"value": dof * stages / (ms_deg / 1e3), "ms_per_step": ms_deg / nsteps,
                                     "launches_per_stage": len(by_deg)},
            "alone": {"value": dof * stages / (ms_alone / 1e3), "ms_per_step": ms_alone / nsteps},
            "bitwise_equal": True}


def imex_workload(args):
    """32x32x10 hex elements on 20 km x 20 km x 10 km, N = 3, Euler-gravity on an isothermal
    background + 0.1 K seeded theta noise; ARK23 (split_rhs = false): explicit EC flux
    differencing + Rusanov, implicit vertical linear Euler-gravity with per-line banded LU."""
    import numpy as np
    import atum_hip as A
    import problems as P

    N = 3
    vx = P.julia_range(0.0, 20e3, 33)
    vz = P.julia_range(0.0, 10e3, 11)
    grid = A.brickgrid(A.LobattoCell(N, 3), (vx, vx, vz), (True, True, False), "stacked")
    prob = A.Problem(boundary="reflect", referencestate=lambda x: P.isothermal_reference(x))
    law = A.EulerGravityLaw(3, grav=9.81, problem=prob)
    ex = A.DGSEM(law=law, grid=grid, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RusanovFlux(), precision=args.precision)
    lin = A.DGSEM(law=A.LinearEulerGravityLaw(law), grid=grid, surface_numericalflux=A.RusanovFlux(),
                  directions=(3,), auxstate=ex.auxstate)
    aux = ex.auxstate
    rho0, p0 = aux[4], aux[5]
    rng = np.random.Generator(np.random.PCG64(20260102))
    cp, cv = 1.4 * P.CV_D, P.CV_D
    R = cp - cv
    T0 = p0 / (rho0 * R)
    exner = (p0 / 1e5) ** (R / cp)
    theta = T0 / exner + 0.1 * rng.standard_normal(grid.Ntot)
    T = theta * exner
    rho = p0 / (R * T)
    z = np.zeros_like(rho)
    q0 = np.stack([rho, z, z, z, rho * (cv * T + aux[0])])
    dt = 0.2 * grid.min_node_distance(dims=(0, 1)) / 330.0
    tuned = {}
    if args.tune:  # both operators' workgroup shapes, before the factorisation and the timing
        tuned = {"explicit": A.autotune(ex, q0, reps=20), "linear": A.autotune(lin, q0, reps=20)}
    q = ex.fieldarray(q0)
    ark = A.ARK23(ex, lin, q, dt, split_rhs=False)
    ctx = ex.ctx
    band_tuned = None
    if args.tune:  # the banded solve kernel, timed on a right-hand side of the state's size
        scratch = ex.fieldarray(q0)
        band_tuned = ark.banded.autotune(ex.fieldarray(), scratch, reps=10)
    ark.run(q, 2)
    ctx.sync()
    ctx.record(0)
    ark.run(q, args.steps)
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1)
    if not np.all(np.isfinite(q.numpy())):
        raise SystemExit("bench: non-finite IMEX state")
    dof = q0.size
    return {"metric": "FP64 DOF-updates/s per RK stage (IMEX ARK23)", "value": dof * 3 * args.steps / (ms / 1e3),
            "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": 2,
            "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "IMEX Euler-gravity 32x32x10, N=3, ARK23 unsplit, banded vertical solve",
                       "dof": dof, "dt": dt, "lines": 32 * 32 * 16, "precision": args.precision,
                       "elements_per_workgroup": {"explicit": ex.epw, "linear": lin.epw},
                       "autotune_us_per_launch": {k: v[1] for k, v in tuned.items()} or None,
                       "banded_solve_kernel": ark.banded.kernel, "banded_solve_waves": ark.banded.waves,
                       "banded_autotune_us_per_solve": band_tuned[1] if band_tuned else None}}


def weak_workload(args):
    """One large weak-form instance (the HBM-bound kernel family): 2-D shallow water,
    --weak-K^2 elements on the periodic unit square, N = --N, weak volume + Rusanov,
    LSRK54 fused stages, elements per workgroup tuned before timing."""
    import numpy as np
    import atum_hip as A
    import problems as P

    K1, N = args.weak_K, args.N
    v = P.julia_range(0.0, 1.0, K1 + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True), lazy=True)
    dg = A.DGSEM(law=A.ShallowWaterLaw(2, grav=9.81), grid=grid, surface_numericalflux=A.RusanovFlux(),
                 geometry="device")
    q0 = P.sw_bump(grid.points, 0.05, 0.1, (0.5, 0.5))
    tuned = None
    if args.epw:
        dg.epw = args.epw
    elif args.tune:
        tuned = A.autotune(dg, q0, reps=20)
    dt = P.sw_ensemble_dt(K1, N, 0.05)
    q = dg.fieldarray(q0)
    s_ = A.LSRK54(dg, q, dt)
    s_.run(q, 4)
    dg.ctx.sync()
    dg.ctx.record(0)
    s_.run(q, args.steps)
    dg.ctx.record(1)
    ms = dg.ctx.elapsed_ms(0, 1)
    if not np.all(np.isfinite(q.numpy())):
        raise SystemExit("bench: non-finite shallow-water state")
    dof = q0.size
    stages = 5 * args.steps
    algo = algorithmic_bytes_per_stage(2, N + 1, grid.K, 4, 0)
    kern_s = ms / 1e3 / stages
    achieved = algo / kern_s / 1e9
    return {"metric": "FP64 DOF-updates/s per RK stage (weak form, one instance)",
            "value": dof * stages / (ms / 1e3), "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 4, "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"2-D shallow water {K1}^2 elements, N={N}, weak form + Rusanov, LSRK54",
                       "dof": dof, "elements_per_workgroup": dg.epw,
                       "autotune_us_per_launch": tuned[1] if tuned else None},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "fused weak stage kernel (volume + surface + LSRK update)",
                         "algorithmic_bytes_per_launch": algo, "avg_launch_us": kern_s * 1e6}}


def vortex_workload(args):
    """2-D isentropic vortex, [-5,5]^2 periodic, 8x8, N = 4, EC + Rusanov, LSRK54, CFL 0.1."""
    import numpy as np
    import atum_hip as A
    import problems as P

    K, N = 8, 4
    v = P.julia_range(-5, 5, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    dg = A.DGSEM(law=A.EulerLaw(2), grid=grid, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RusanovFlux(), precision=args.precision)
    q0 = P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0))
    dt = 0.1 * (10 / K) / N
    if args.tune:
        A.autotune(dg, q0, reps=50)
    q = dg.fieldarray(q0)
    s_ = A.LSRK54(dg, q, dt)
    s_.run(q, 4)
    dg.ctx.sync()
    dg.ctx.record(0)
    s_.run(q, args.steps)
    dg.ctx.record(1)
    ms = dg.ctx.elapsed_ms(0, 1)
    dof = q0.size
    if not np.all(np.isfinite(q.numpy())):
        raise SystemExit("bench: non-finite vortex state")
    return {"metric": "FP64 DOF-updates/s per RK stage (isentropic vortex)", "value": dof * 5 * args.steps / (ms / 1e3),
            "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": 4,
            "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "2-D isentropic vortex 8x8, N=4, EC + Rusanov, LSRK54 (CFL 0.1)",
                       "dof": dof, "us_per_stage": ms / (5 * args.steps) * 1e3, "precision": args.precision,
                       "elements_per_workgroup": dg.epw}}


def serve_workload(args):
    """Served stream: Poisson arrivals (rate 50/s, PCG64 seed 20260103) for --serve-seconds,
    50/30/20 vortex (8..128^2, N 3..7) / shallow water (8..64^2, N 3..7) / IMEX Euler-gravity
    (8..32^2 x 10, N 3) requests with U{50..500} steps, continuously batched."""
    import atum_hip as A
    from serve_requests import make_stream

    reqs = make_stream(duration=args.serve_seconds)
    progress(f"serve: {len(reqs)} requests")
    server = A.Server(chunk=args.chunk, keep_results=False, imex_streams=args.imex_streams,
                      max_active=args.serve_max_active,
                      tune=A.TuneCache() if args.serve_tune else None)
    stats = server.run(reqs)
    s = stats.summary()
    kinds = {k: sum(r.kind == k for r in reqs) for k in ("vortex", "sw", "imex")}
    progress(f"serve: served in {stats.wall:.1f} s; per-request sample (every {args.serve_sample}th alone)")
    per = per_request_sample(reqs[::args.serve_sample], server.cache)
    # per kind: time to completion, and what serving the kind's requests one after another
    # would take at the rate its sampled requests reach alone (kinds that fill the device on
    # their own gain little from shared launches; small ones gain most)
    import numpy as np
    per_kind = {}
    c0 = stats.clock0
    for k in ("vortex", "sw", "imex"):
        rk = [r for r in reqs if r.kind == k and r.finished is not None]
        if not rk:
            continue
        lat = np.array([r.finished - (c0 + r.arrival) for r in rk])
        upd = float(sum(r.dof * r.steps * (3 if k == "imex" else 5) for r in rk))
        rate = per["per_kind"].get(k, {}).get("value", 0.0)
        per_kind[k] = {"requests": len(rk), "dof_updates": upd, "p50_s": float(np.percentile(lat, 50)),
                       "p99_s": float(np.percentile(lat, 99)), "alone_value": rate,
                       "alone_seconds_estimate": upd / rate if rate > 0 else None}
    plain_s = sum(v["alone_seconds_estimate"] or 0.0 for v in per_kind.values())
    return {"metric": "served stream: aggregate FP64 DOF-updates/s per RK stage, p50/p99 time-to-completion",
            "value": s["dof_updates_per_s"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": None,
            "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Poisson 50/s for {args.serve_seconds} s, mix 50/30/20 "
                                   "vortex/shallow-water/IMEX, steps U{50..500}", "requests": len(reqs),
                       "kinds": kinds, "chunk_steps": args.chunk, "imex_streams": args.imex_streams,
                       "max_active": args.serve_max_active,
                       "autotune": bool(args.serve_tune)},
            "p50_s": s["p50_s"], "p99_s": s["p99_s"], "mean_s": s["mean_s"], "completed": s["completed"],
            "wall_s": s["wall_s"], "ticks": s["ticks"], "max_active": s["max_active"],
            "breakdown_s": {k: s[k] for k in ("build_s", "admit_s", "enqueue_s", "wait_s", "retire_s", "gpu_span_s")},
            "per_request_sample": per, "per_kind": per_kind,
            "plain_seconds_estimate": plain_s,
            "speedup_vs_per_request": s["dof_updates_per_s"] / per["value"] if per["value"] > 0 else None,
            "footprint": server.footprint_stats()}


def per_request_sample(reqs, cache):
    """The plain way to serve the same requests: each one alone, its solver's launches only
    (LSRK54 with its step graph, or ARK23 with its factorisation built beforehand), timed
    with HIP events around each run; returns the aggregate DOF-updates/s over the sample."""
    import math
    import atum_hip as A

    ctx = A.context()
    dof_updates, seconds = 0.0, 0.0
    kinds = {}
    for r in reqs:
        built = r.build(cache)
        if r.kind == "imex":
            ex, lin, q0, dt = built
            q = lin.fieldarray(q0)
            band = cache.get(("banded", r.key, dt)) or A.BandedOperator(lin, -dt * (1 - 1 / math.sqrt(2)))
            solver = A.ARK23(ex, lin, q, dt, split_rhs=False, banded=band)
            stages = 3
        else:
            dg, q0, dt = built
            q = dg.fieldarray(q0)
            solver = A.LSRK54(dg, q, dt)
            stages = 5
        ctx.sync()
        ctx.record(10)
        solver.run(q, r.steps)
        ctx.record(11)
        el = ctx.elapsed_ms(10, 11) / 1e3
        seconds += el
        dof_updates += q0.size * r.steps * stages
        kd = kinds.setdefault(r.kind, [0, 0.0, 0.0])
        kd[0] += 1
        kd[1] += q0.size * r.steps * stages
        kd[2] += el
    per_kind = {k: {"requests": v[0], "value": v[1] / v[2] if v[2] > 0 else 0.0, "gpu_s": v[2]}
                for k, v in kinds.items()}
    return {"requests": len(reqs), "value": dof_updates / seconds if seconds > 0 else 0.0,
            "unit": "DOF-updates/s", "gpu_s": seconds, "per_kind": per_kind,
            "note": "every k-th request of the stream run alone, one after another (solver launches only)"}


def copy_workload(args):
    """Calibration of the HBM byte counters (and a practical bandwidth reference): three
    device-to-device copies of 1 GiB through the library (atm_copy), timed with HIP events."""
    import ctypes
    import atum_hip._lib as L

    ctx = L.context()
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    ctx.check(L.lib.atm_alloc(ctx.h, n, ctypes.byref(a)), "atm_alloc")
    ctx.check(L.lib.atm_alloc(ctx.h, n, ctypes.byref(b)), "atm_alloc")
    ctx.check(L.lib.atm_copy(ctx.h, b, a, n), "atm_copy")  # warm-up
    ctx.record(0)
    for _ in range(3):
        ctx.check(L.lib.atm_copy(ctx.h, b, a, n), "atm_copy")
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1)
    ctx.sync()
    L.lib.atm_free(ctx.h, a)
    L.lib.atm_free(ctx.h, b)
    gbs = 3 * 2 * n / (ms / 1e3) / 1e9
    return {"metric": "device-to-device copy bandwidth (read + write)", "value": gbs, "unit": "GB/s",
            "n_gpus": 1, "steps": 3, "warmup": 1, "ms_per_step": ms / 3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "copy of 1 GiB, three times (HBM counter calibration)", "bytes": n}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--elements", type=int, default=16, help="elements per direction")
    ap.add_argument("--N", type=int, default=4)
    ap.add_argument("--cpu-stages", type=int, default=450)
    ap.add_argument("--dt-scale", type=float, default=1.0,
                    help="bubble: multiply the rising-bubble time step (its CFL formula is the "
                         "reference's for N = 4; higher degrees need a smaller step to stay stable)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precision", choices=["fast", "reference"], default="fast",
                    help="fast: pair-symmetric EC kernel (round-off-level differences); "
                         "reference: the reference's summation order, no FMA")
    ap.add_argument("--no-reference-order", action="store_true",
                    help="skip the extra timing of the reference-order kernel")
    ap.add_argument("--workload", choices=["bubble", "ensemble", "ensemble3d", "imex", "vortex", "serve", "copy",
                                           "weak"],
                    default="bubble",
                    help="bubble (default, the metric's configuration); the others are extra lines")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--instances3d", type=int, default=64, help="ensemble3d: number of 3-D instances")
    ap.add_argument("--weak-K", type=int, default=256, help="weak workload: elements per direction")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="serve: length of the Poisson arrival stream (the served configuration: 60 s)")
    ap.add_argument("--ensemble-steps", type=int, default=100,
                    help="bubble line: steps of the batched ensemble it carries (the configuration: 100)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="bubble line: leave out the batched shallow-water ensemble object")
    ap.add_argument("--no-serve", action="store_true",
                    help="bubble line: leave out the served-stream object")
    ap.add_argument("--chunk", type=int, default=10, help="steps per serving tick")
    ap.add_argument("--epw", type=int, default=0,
                    help="bubble / weak: use this many elements per workgroup instead of tuning")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="bubble: keep the default elements per workgroup (no autotuning before timing)")
    ap.add_argument("--ensemble-tune", action="store_true",
                    help="ensemble: tune one workgroup shape per launch group on the batched run "
                         "(measured within noise of the default shapes)")
    ap.add_argument("--serve-sample", type=int, default=8,
                    help="serve: every k-th request is also run alone for the per-request baseline")
    ap.add_argument("--serve-tune", action="store_true",
                    help="serve: tune each new operator's workgroup shape on first use (under load; "
                         "measured slower on the default stream)")
    ap.add_argument("--serve-max-active", type=int, default=64,
                    help="serve: requests resident on the device at once (admission limit besides memory)")
    ap.add_argument("--imex-streams", type=int, default=3,
                    help="side streams for the IMEX requests of the served stream (0: the main stream)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; the path would shard by elements with one face "
              "exchange per stage, but no multi-GPU run is made here.", file=sys.stderr)
        sys.exit(2)
    if args.steps % 2:
        args.steps += 1  # whole two-step graphs: the state returns to q every pair

    import numpy as np
    import atum_hip as A

    if args.workload != "bubble":
        fn = {"ensemble": ensemble_workload, "ensemble3d": ensemble3d_workload, "imex": imex_workload,
              "vortex": vortex_workload,
              "serve": serve_workload, "copy": copy_workload, "weak": weak_workload}[args.workload]
        print(json.dumps(fn(args)))
        return

    Kd, N = args.elements, args.N

    tuning = {}

    def timed(precision):
        grid, dg, q0, dt = bubble_setup(Kd, N, precision)
        dt *= args.dt_scale
        ctx = dg.ctx
        if args.epw and precision == args.precision:  # a fixed workgroup shape (e.g. for counter runs)
            dg.epw = args.epw
        elif args.tune:  # elements per workgroup chosen by timing, before the timed region
            best, times = A.autotune(dg, q0, reps=50)
            tuning[precision] = {"epw": best, "us_per_launch": times}
        q = dg.fieldarray(q0)
        solver = A.LSRK54(dg, q, dt, use_graph=not args.no_graph)
        solver.run(q, args.warmup)
        ctx.sync()
        ctx.record(0)
        solver.run(q, args.steps)
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1)
        ctx.sync()
        if not np.all(np.isfinite(q.numpy())):
            raise SystemExit("bench: non-finite state")
        return grid, dg, dt, ms

    progress(f"bubble {Kd}^3 N={N} ({args.precision}): tuning and {args.steps} timed steps")
    grid, dg, dt, ms = timed(args.precision)
    progress(f"bubble: {ms / args.steps:.3f} ms per step")
    Ns, K, Np = dg.Ns, grid.K, grid.Np
    stages = 5 * args.steps
    dof = Ns * Np * K
    value = dof * stages / (ms / 1e3)
    naux = dg.auxstate.shape[0]
    algo = algorithmic_bytes_per_stage(3, N + 1, K, Ns, naux)
    kern_s = ms / 1e3 / stages
    achieved = algo / kern_s / 1e9
    # HBM bytes per launch come from rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs of this
    # kernel on this workload (a profiler pass of its own), read here from their summary
    traffic, traffic_source = None, None
    tpath = os.path.join(ROOT, "profiles", "fd_kernel_bubble16_traffic.json")
    if os.path.exists(tpath) and Kd == 16 and N == 4 and args.precision == "fast":
        try:
            tj = json.load(open(tpath))
            traffic = tj.get("bytes_per_dispatch")
            traffic_source = {"read_from": "profiles/fd_kernel_bubble16_traffic.json", "measured_in_this_run": False,
                              "counters": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, scaled by a copy kernel",
                              "kernel": tj.get("kernel"), "source_commit": tj.get("source_commit")}
        except Exception:
            traffic = None
    out = {
        "metric": "FP64 DOF-updates/s per RK stage",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Euler-gravity rising bubble, {Kd}^3 hex elements, N={N}, "
                               "EC flux differencing + Roe surface, reflective walls, LSRK54",
                   "time_step": "0.25 dx / N / 330 m/s: CFL 1/4 on the reference bubble's rule (its CFL 1/3 "
                                "is unstable in 3-D)",
                   "elements": K, "dof": dof, "stages_per_step": 5, "dt": dt,
                   "hipgraph": not args.no_graph, "precision": args.precision,
                   "elements_per_workgroup": dg.epw, "autotune": tuning.get(args.precision)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": "fused stage kernel (volume + surface + LSRK update)",
                     "algorithmic_bytes_per_launch": algo, "avg_launch_us": kern_s * 1e6},
    }
    flops = ec_volume_flops_per_stage(3, N + 1, K, Ns, True)
    out["roofline_fp64"] = {"bound": "fp64", "achieved": flops / kern_s / 1e12, "peak": FP64_PEAK_TFS,
                            "unit": "TFLOP/s", "frac": flops / kern_s / 1e12 / FP64_PEAK_TFS,
                            "flops_per_launch": flops,
                            "note": "EC volume operations in the reference's evaluation count (each "
                                    "ordered pair, diagonal included); the fused kernel evaluates each "
                                    "unordered pair once and also does the surface term and the update"}
    if args.precision == "fast" and not args.no_reference_order:
        progress("bubble, reference-order kernel")
        _, _, _, ms_ref = timed("reference")
        out["reference_order"] = {"value": dof * stages / (ms_ref / 1e3), "unit": "DOF-updates/s",
                                  "ms_per_step": ms_ref / args.steps,
                                  "avg_launch_us": ms_ref / 1e3 / stages * 1e6,
                                  "autotune": tuning.get("reference"),
                                  "note": "same run with the reference's summation order and no FMA"}
    if not args.no_cpu_baseline:
        progress(f"CPU baseline: {args.cpu_stages} stages of the oracle")
        out["cpu_baseline"] = cpu_baseline(Kd, N, args.cpu_stages)
    # the rest of the metric: the batched ensemble (one launch per stage for all members) and
    # the served stream (p50/p99 time to completion), each against launching the plain
    # kernels once per request
    if not args.no_ensemble:
        progress(f"ensemble: {args.instances} shallow-water instances, {args.ensemble_steps} steps")
        e = ensemble_workload(args, steps=args.ensemble_steps)
        out["ensemble"] = {k: e[k] for k in ("value", "unit", "steps", "ms_per_step", "config",
                                             "per_request_eager", "per_request_graph",
                                             "speedup_vs_per_request_eager", "speedup_vs_per_request_graph",
                                             "bitwise_equal_to_alone")}
    if not args.no_serve:
        progress(f"served stream: {args.serve_seconds:.0f} s of Poisson arrivals")
        sv = serve_workload(args)
        out["served"] = {k: sv.get(k) for k in ("value", "unit", "config", "p50_s", "p99_s", "mean_s", "completed",
                                            "wall_s", "max_active", "per_request_sample", "per_kind",
                                            "speedup_vs_per_request", "footprint")}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
