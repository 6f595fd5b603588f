"""Benchmark: one RK3 step of NUC3d's right-hand side + TVD-RK3 advance on one
MI355X, on decaying compressible isotropic turbulence at 256^3.

Workload (the configuration the metric is quoted on): 256^3 periodic box
[0,2pi)^3, seeded random-phase solenoidal velocity with E(k) ~ k^4 exp(-2(k/4)^2),
turbulent Mach 0.3 (u_rms = 1, Mach = 0.3/sqrt(3)), Re = 200 (Re_lambda ~ 100),
rho = 1, p = 1/(gamma Mach^2); CRWENO5 in x/y/z, Lax-Friedrichs splitting,
8th-order central viscous terms, TVD-RK3 at CFL 0.5, FP64.

A step is one TVD-RK3 step (three RHS evaluations + stage updates) with the
state resident in HBM.  value = cells * 3 * steps / seconds / 1e6
(Mcell-updates per second per RK stage).

roofline: the dominant kernel (largest total time among the library's
kernels, timed with HIP events on the library's stream during the timed
region).  Its algorithmic bytes per launch are the per-stage compulsory
traffic of 106.7 B per cell (320 B per cell per RK3 step: read u^n and u^s,
write u^{s+1}) split evenly over the three direction sweeps that make up a
stage: 35.6 B per cell per sweep launch.

cpu_baseline: the oracle (a C restatement of the reference algorithm) timed on
this host's cores on one RHS evaluation of the same 256^3 workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nuc3d-cartisian_amd"))

HBM_PEAK_GBS = 8000.0
B_ALG_STAGE = 320.0 / 3.0   # bytes per cell per RK stage


def workload_config(G, n=256, precision=None):
    import numpy as np
    return G.make_config(n=(n, n, n), npr=(1, 1, 1), scheme=G.CRWENO5, vis=G.CD8, riemann=G.LF,
                         model=G.NAVIERSTOKES3D, viscosity=G.SUTHERLAND, mach=0.3 / np.sqrt(3.0), reynolds=200.0,
                         prandtl=0.72, gamma=1.4, cfl=0.5, precision=G.FP64 if precision is None else precision)


def cpu_baseline(q, n, threads):
    """Time the oracle's RHS evaluation of the same state on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    oc = O.make_config(n=(n, n, n), npr=(1, 1, 1), scheme=O.CRWENO5, vis=O.CD8, riemann=O.LF, model=O.NS,
                       mach=0.3 / np.sqrt(3.0), reynolds=200.0, cfl=0.5, nthreads=threads)
    s = O.Solver(oc)
    s.set_state(q)
    t0 = time.perf_counter()
    s.rhs()
    t1 = time.perf_counter()
    del s
    cells = n ** 3
    return {"value": cells / (t1 - t0) / 1e6, "unit": "Mcell-updates/s per RK3 stage", "cores": threads,
            "kind": "port",
            "sample": f"one RHS evaluation (one RK stage) of the same {n}^3 CRWENO5/LF/cd8 state, "
                      f"oracle C restatement, gcc -O3 -ffp-contract=off, OpenMP {threads} threads, "
                      f"{t1 - t0:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-slots", action="store_true", help="print per-kernel times to stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one box per GPU and measures one MI355X; "
              "--gpus above 1 is not supported (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import nuc3d_amd as G

    n = args.n
    cfg = workload_config(G, n)
    q = G.init_hit(cfg, k0=4.0, u_rms=1.0, seed=20261202)
    cells = n ** 3
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        if args.warmup:
            ctx.step_async(args.warmup)
            ctx.sync()
        ctx.profile(True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step_async(args.steps)
        dt, res, tsim = ctx.sync()
        t1 = time.perf_counter()
        prof = ctx.profile_read()
        ctx.profile(False)
        q_end = ctx.download()
    if not np.all(np.isfinite(q_end)):
        print("bench.py: state became non-finite", file=sys.stderr)
        sys.exit(3)
    secs = t1 - t0
    ms_per_step = secs / args.steps * 1e3
    value = cells * 3 * args.steps / secs / 1e6

    # dominant kernel and its roofline
    name, (tot_ms, launches) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_ms = tot_ms / max(launches, 1)
    if name.startswith("k_sweep"):
        alg_bytes = B_ALG_STAGE / 3.0 * cells
    else:
        alg_bytes = B_ALG_STAGE * cells
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": name,
                "avg_launch_ms": round(avg_ms, 4)}
    stage_ms = ms_per_step / 3.0
    if args.profile_slots:
        for k, (t, c) in prof.items():
            print(f"  {k:10s} total {t:9.3f} ms  launches {c:5d}  avg {t / max(c, 1):8.4f} ms", file=sys.stderr)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(q, n, threads)

    out = {
        "metric": "Mcell-updates/s per RK3 stage at 256^3, FP64 vs mixed; serving jobs/s + p50/p99",
        "value": round(value, 3),
        "unit": "Mcell-updates/s per RK3 stage",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-phase solenoidal HIT field, seed 20261202)",
        "config": {"workload": f"decaying compressible isotropic turbulence {n}^3, CRWENO5 x3 + LF splitting + "
                               "cd8 viscous + TVD-RK3, CFL 0.5, FP64",
                   "grid": [n, n, n], "rank_grid": [1, 1, 1], "mach_t": 0.3, "reynolds": 200.0,
                   "stage_ms": round(stage_ms, 4), "dt_last": dt},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
