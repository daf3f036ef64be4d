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


def algorithmic_bytes_per_stage(dim, Nq, K, Ns, naux):
    """Canonical traffic of one stage in the reference data model: per node q and dq
    read+write (32*Ns), metric + MJ ((d^2+1)*8), aux (naux*8); per face point normal +
    faceMJ ((d+1)*8) and two int32 indices (8); per element 2d int32 tags."""
    Np = Nq ** dim
    Nfp = 2 * dim * Nq ** (dim - 1)
    per_elem = Np * (32 * Ns + (dim * dim + 1) * 8 + naux * 8) + Nfp * ((dim + 1) * 8 + 8) + 2 * dim * 4
    return per_elem * K


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
    dt = (1.0 / 3.0) * (2000.0 / Kd) / N / 330.0  # experiment/euler_gravity/risingbubble.jl:66-67
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
    dt = (1.0 / 3.0) * (2000.0 / Kd) / N / 330.0
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--elements", type=int, default=16, help="elements per direction")
    ap.add_argument("--N", type=int, default=4)
    ap.add_argument("--cpu-stages", type=int, default=150)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precision", choices=["fast", "reference"], default="fast",
                    help="fast: pair-symmetric EC kernel (round-off-level differences); "
                         "reference: the reference's summation order, no FMA")
    ap.add_argument("--no-reference-order", action="store_true",
                    help="skip the extra timing of the reference-order kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the path shards by elements with one face "
              "exchange per stage (see DESIGN.md), but no multi-GPU run is made here.", file=sys.stderr)
        sys.exit(2)
    if args.steps % 2:
        args.steps += 1  # whole two-step graphs: the state returns to q every pair

    import numpy as np
    import atum_hip as A

    Kd, N = args.elements, args.N

    def timed(precision):
        grid, dg, q0, dt = bubble_setup(Kd, N, precision)
        ctx = dg.ctx
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

    grid, dg, dt, ms = timed(args.precision)
    Ns, K, Np = dg.Ns, grid.K, grid.Np
    stages = 5 * args.steps
    dof = Ns * Np * K
    value = dof * stages / (ms / 1e3)
    naux = dg.auxstate.shape[0]
    algo = algorithmic_bytes_per_stage(3, N + 1, K, Ns, naux)
    kern_s = ms / 1e3 / stages
    achieved = algo / kern_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fd_kernel_bubble16_traffic.json")
    if os.path.exists(tpath) and Kd == 16 and N == 4:
        try:
            traffic = json.load(open(tpath)).get("bytes_per_dispatch")
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
                   "elements": K, "dof": dof, "stages_per_step": 5, "dt": dt,
                   "hipgraph": not args.no_graph, "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "fused stage kernel (volume + surface + LSRK update)",
                     "algorithmic_bytes_per_launch": algo, "avg_launch_us": kern_s * 1e6},
    }
    if args.precision == "fast" and not args.no_reference_order:
        _, _, _, ms_ref = timed("reference")
        out["reference_order"] = {"value": dof * stages / (ms_ref / 1e3), "unit": "DOF-updates/s",
                                  "ms_per_step": ms_ref / args.steps,
                                  "avg_launch_us": ms_ref / 1e3 / stages * 1e6,
                                  "note": "same run with the reference's summation order and no FMA"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(Kd, N, args.cpu_stages)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
