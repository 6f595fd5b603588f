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

The sweeps are bound by FP64 VALU issue rather than HBM; roofline also
reports valu_issue_frac: the kernel's VALU wave-instructions per launch (from
the committed SQ_INSTS_VALU summary) at one per 4 cycles per SIMD, over its
measured launch time.

cpu_baseline: the oracle (a C restatement of the reference algorithm) timed on
this host's CPU share (OMP_NUM_THREADS, else every core) on two full TVD-RK3
steps of the same 256^3 workload, run as that many emulated MPI ranks (the
box cut into the reference's rank blocks, one thread per rank's share of the
line loops).

mixed: the same workload with NUC_MIXED (FP32 split fluxes, stencils, CRWENO5's
nonlinear weights and viscous stencils; FP64 tridiagonal solves, RK3 and
reductions), timed the same way, with the relative error of every global
statistic (kinetic energy, dissipation, enstrophy, u_rms, Re_lambda, ...)
against the FP64 run from the same initial state after the same steps.

fp32: the same with NUC_FP32 (the mixed path's arithmetic with the conserved
state and the RHS stored in FP32), timed and compared the same way.

weno5js: the same box and initial state with WENO5-JS in place of CRWENO5
(FP64 and mixed), timed the same way; a secondary figure.

configs: rates of the other named workloads (Taylor-Green 64^3 FP64;
forced turbulence 512^3 hccs/AUSM mixed and FP64; the Shu-Osher box
512x128x128 FP64, mixed and FP32), a few timed steps each.

serving: the serving stream -- 256 independent boxes (Taylor-Green or decaying
turbulence, N log-uniform in [32, 192] per side, Mach U(0.1,0.5),
Re U(400,3200), scheme uniform over WENO5-JS/WENO5-Z/CRWENO5, LF, cd6,
CFL 0.5, 300 RK3 steps each, initial states generated on the device) arriving
as a Poisson stream at 20 jobs/s -- through the serving loop, as an FP64 and a
mixed stream: jobs/s and p50/p99 latency (arrival -> result), against the same
jobs run one at a time on plain contexts, and the ceiling any schedule could
reach (the stream's cell-updates at the 256^3 box's rate).
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


def workload_config(G, n=256, precision=None, scheme=None):
    import numpy as np
    return G.make_config(n=(n, n, n), npr=(1, 1, 1), scheme=G.CRWENO5 if scheme is None else scheme,
                         vis=G.CD8, riemann=G.LF,
                         model=G.NAVIERSTOKES3D, viscosity=G.SUTHERLAND, mach=0.3 / np.sqrt(3.0), reynolds=200.0,
                         prandtl=0.72, gamma=1.4, cfl=0.5, precision=G.FP64 if precision is None else precision)


def rank_grid(P, n):
    """P emulated ranks as an (npx, npy, npz) grid whose sizes divide n (the
    reference's MPI Cartesian grid, src/MPICommunicator.cpp:26-56)."""
    dims = [1, 1, 1]
    f, rest = 2, P
    primes = []
    while rest > 1:
        while rest % f == 0:
            primes.append(f)
            rest //= f
        f += 1
    for pr in sorted(primes, reverse=True):
        d = min((i for i in range(3) if n % (dims[i] * pr) == 0), key=lambda i: dims[i], default=None)
        if d is None:
            continue   # a factor that does not divide the grid: fewer ranks than threads
        dims[d] *= pr
    return tuple(dims)


def cpu_baseline(q, n, threads, steps=2):
    """Time the oracle on `steps` full TVD-RK3 steps of the same state on the
    host cores, run as P = threads emulated MPI ranks: the box is cut into the
    reference's P rank blocks (compact-scheme closures per block, as a P-rank
    reference run places them) and every line loop runs on P threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    npr = rank_grid(threads, n)
    oc = O.make_config(n=(n, n, n), npr=npr, scheme=O.CRWENO5, vis=O.CD8, riemann=O.LF, model=O.NS,
                       mach=0.3 / np.sqrt(3.0), reynolds=200.0, cfl=0.5, nthreads=threads)
    s = O.Solver(oc)
    s.set_state(q)
    t0 = time.perf_counter()
    s.step(steps)
    t1 = time.perf_counter()
    del s
    cells = n ** 3
    return {"value": cells * 3 * steps / (t1 - t0) / 1e6, "unit": "Mcell-updates/s per RK3 stage",
            "cores": threads, "kind": "port", "ranks": list(npr),
            "sample": f"{steps} full TVD-RK3 steps ({3 * steps} RHS evaluations + stage updates) of the same "
                      f"{n}^3 CRWENO5/LF/cd8 state as {npr[0] * npr[1] * npr[2]} emulated ranks "
                      f"{npr[0]}x{npr[1]}x{npr[2]}, "
                      f"oracle C restatement, gcc -O3 -ffp-contract=off, OpenMP {threads} threads, "
                      f"{t1 - t0:.1f} s"}


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n^3 from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by a device-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "k_sweep_crweno5_256_fp64_traffic.md")
    if n != 256 or not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("machine-readable: "):
            rec = json.loads(line[len("machine-readable: "):])
            if rec.get("kernel") == kernel:
                return float(rec["traffic_bytes"]), os.path.relpath(path, ROOT)
    return None, None


VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4   # one 64-lane FP64 VALU wave-instruction per 4 cycles per SIMD


def measured_valu(kernel, n):
    """VALU wave-instructions per launch of `kernel` at n^3 from the committed
    rocprofv3 --pmc SQ_INSTS_VALU summary, or None."""
    path = os.path.join(ROOT, "profiles", "k_sweep_crweno5_256_fp64_valu.md")
    if n != 256 or not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("machine-readable: "):
            rec = json.loads(line[len("machine-readable: "):])
            if rec.get("kernel") == kernel:
                return float(rec["valu_wave_insts"]), os.path.relpath(path, ROOT)
    return None, None


TILE_SHAPES = {}
PCIE = {}


def stage_stats(G, cfg, q, steps, warmup):
    """(seconds for `steps` timed steps, stats_ex after warmup+steps, profile)"""
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        t_up = time.perf_counter()
        ctx.upload(q)                  # host -> HBM over PCIe (pageable numpy buffer)
        PCIE["upload_ms"] = (time.perf_counter() - t_up) * 1e3
        # tile shapes from the library's autotuner, before the untimed warmup
        TILE_SHAPES[cfg.precision] = ctx.autotune()
        if warmup:
            ctx.step_async(warmup)
            ctx.sync()
        ctx.profile(True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step_async(steps)
        dt, res, tsim = ctx.sync()
        t1 = time.perf_counter()
        prof = ctx.profile_read()
        ctx.profile(False)
        st = ctx.stats_ex()
        t_dn = time.perf_counter()
        q_end = ctx.download()
        PCIE["download_ms"] = (time.perf_counter() - t_dn) * 1e3
    return t1 - t0, st, prof, q_end, dt


def config_rates(G, steps=3, warmup=1):
    """Rates of the other named workloads (Mcell-updates/s per RK3 stage, state
    resident in HBM, initial states generated on the device or uploaded before
    timing): Taylor-Green 64^3 (WENO5-JS/LF/cd6, FP64, 200 steps, one captured
    graph per step), forced turbulence 512^3 (hccs/AUSM/cd6, Const forcing on
    1 < |k| < 2, mixed and FP64) and the Shu-Osher box 512x128x128
    (WENO5-Z/AUSM/cd6, CFL 0.4, FP64, mixed and FP32)."""
    import numpy as np
    out = {}

    def timed(cfg, init, nsteps, nwarm, graph=False):
        with G.Context(cfg) as ctx:
            init(ctx)
            ctx.autotune()
            if graph:
                ctx.set_graph(True)
            ctx.step_async(nwarm)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step_async(nsteps)
            ctx.sync()
            t1 = time.perf_counter()
            ok = bool(np.isfinite(ctx.stats()).all())
        cells = cfg.nx * cfg.ny * cfg.nz
        return {"value": round(cells * 3 * nsteps / (t1 - t0) / 1e6, 1),
                "ms_per_step": round((t1 - t0) / nsteps * 1e3, 4), "steps": nsteps, "finite": ok}

    print("bench: Taylor-Green 64^3", file=sys.stderr, flush=True)
    c1 = G.make_config(n=(64, 64, 64), scheme=G.WENO5JS, vis=G.CD6, riemann=G.LF, model=G.NAVIERSTOKES3D,
                       viscosity=G.SUTHERLAND, mach=0.1, reynolds=1600.0, prandtl=0.72, cfl=0.5)
    out["tgv_64_weno5js_fp64"] = timed(c1, lambda c: c.init_tgv_device(), 200, 5, graph=True)
    for label, prec in (("mixed", G.MIXED), ("fp64", G.FP64)):
        print(f"bench: forced 512^3 hccs {label}", file=sys.stderr, flush=True)
        c3 = G.make_config(n=(512, 512, 512), scheme=G.HCCS, vis=G.CD6, riemann=G.AUSM, model=G.NAVIERSTOKES3D,
                           viscosity=G.SUTHERLAND, mach=0.5 / np.sqrt(3.0), reynolds=200.0, prandtl=0.72, cfl=0.5,
                           forcing=G.FORCING_CONST, forcing_k=1.5, forcing_amp=0.05, forcing_amp_hel=0.0,
                           precision=prec)
        out[f"forced_512_hccs_{label}"] = timed(c3, lambda c: c.init_hit_device(k0=4.0, u_rms=1.0, seed=20261502),
                                                steps, warmup)
    n4 = (512, 128, 128)
    kw4 = dict(n=n4, L=(10.0, 2.5, 2.5), scheme=G.WENO5Z, vis=G.CD6, riemann=G.AUSM, model=G.NAVIERSTOKES3D,
               viscosity=G.SUTHERLAND, mach=1.0 / np.sqrt(1.4), reynolds=2000.0, prandtl=0.72, cfl=0.4)
    q4 = G.init_shu_osher(G.make_config(**kw4), pert=0.05, seed=20261602)
    for label, prec in (("fp64", G.FP64), ("mixed", G.MIXED), ("fp32", G.FP32)):
        print(f"bench: Shu-Osher 512x128x128 {label}", file=sys.stderr, flush=True)
        out[f"shu_osher_512x128x128_{label}"] = timed(G.make_config(precision=prec, **kw4), lambda c: c.upload(q4),
                                                      5 * steps, warmup)
    return out


STAT_NAMES = ["ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz", "om_x", "om_y", "om_z", "enstrophy", "Q",
              "helicity", "KE", "u_rms", "M_t", "theta_rms", "mu_mean", "Re_lambda", "lambda", "rho_mean",
              "tau_t", "skewness", "eps_s", "eps_d"]


BUDGET_TERMS = ["tke", "production", "transport", "pressure_diffusion", "p1", "p2", "p3", "viscous_diffusion",
                "v1", "v2", "v3", "convection"]


def budget_error(G, n, q, steps=3):
    """Volume means of the TKE budget terms (nuc_average_*, time averages over
    `steps` steps after each step, untimed) in FP64 and in the mixed path, and
    the mixed path's error; terms whose FP64 mean is a divergence (zero on the
    periodic box) are measured against the budget's scale."""
    import numpy as np
    means = {}
    for prec in (G.FP64, G.MIXED):
        with G.Context(workload_config(G, n, prec)) as ctx:
            ctx.upload(q)
            ctx.average_begin(3)
            for _ in range(steps):
                ctx.step(1)
                ctx.average_update()
            means[prec] = ctx.average_means(4)[:len(BUDGET_TERMS)]
            ctx.average_end()
    b64, bmx = means[G.FP64], means[G.MIXED]
    scale = float(np.abs(b64[[0, 1, 3, 7]]).max())
    return {"averaged_steps": steps,
            "fp64_means": {t: float(b64[i]) for i, t in enumerate(BUDGET_TERMS)},
            "rel_err_vs_fp64": {t: float(abs(bmx[i] - b64[i]) / max(abs(b64[i]), scale))
                                for i, t in enumerate(BUDGET_TERMS)}}


# The serving stream: 256 independent jobs, Poisson arrivals at 20 jobs/s,
# Taylor-Green or decaying turbulence (50/50), N log-uniform in [32, 192] per
# side, Mach U(0.1, 0.5), Re U(400, 3200), scheme uniform over
# WENO5-JS/WENO5-Z/CRWENO5, LF, cd6, CFL 0.5, 300 RK3 steps each; FP64 and
# mixed precision as separate streams admitted against a fixed HBM budget.
SERVE_JOBS, SERVE_NMIN, SERVE_NMAX, SERVE_STEPS, SERVE_RATE = 256, 32, 192, 300, 20.0
SERVE_SEED = 20261202 + 500
SERVE_HBM_GIB = 64
SERVE_MAX_ACTIVE = 16
# idle compatible jobs of together at most this many cells advance as one
# pack (one launch per kernel): up to four 40^3 or two 48^3 boxes
SERVE_PACK_CELLS = 4 * 40 ** 3


def serving_jobs(G, njobs, nmin, nmax, rate, seed):
    """Job specs (cfg, initial condition, u_rms, seed, n) and Poisson arrival times."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        n += n % 2
        mach = rng.uniform(0.1, 0.5)
        re = rng.uniform(400.0, 3200.0)
        scheme = [G.WENO5JS, G.WENO5Z, G.CRWENO5][rng.integers(3)]
        cfg = G.make_config(n=(n, n, n), scheme=scheme, vis=G.CD6, riemann=G.LF, model=G.NAVIERSTOKES3D,
                            mach=mach, reynolds=re, cfl=0.5)
        hit = rng.random() >= 0.5
        u_rms = min(1.0, 0.3 / (np.sqrt(3.0) * mach))
        jobs.append((cfg, G.IC_HIT if hit else G.IC_TGV, u_rms, int(rng.integers(1 << 30)), n))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=njobs))
    return jobs, arrivals


def with_precision(jobs, prec):
    import copy
    out = []
    for cfg, ic, u_rms, seed, n in jobs:
        c = copy.copy(cfg)
        c.precision = prec
        out.append((c, ic, u_rms, seed, n))
    return out


def run_stream(G, jobs, arrivals, steps, pack_cells=SERVE_PACK_CELLS, stats=None, label=""):
    """Jobs submitted at their arrival times (None: all at once), initial
    states generated on the device at admission; returns (wall s, latencies s)."""
    import numpy as np
    with G.Server(hbm_budget=SERVE_HBM_GIB << 30, max_active=SERVE_MAX_ACTIVE, pack_cells=pack_cells,
                  autotune=True) as srv:
        ids = []
        t0 = time.perf_counter()
        for k, (cfg, ic, u_rms, seed, _) in enumerate(jobs):
            if arrivals is not None:
                wait = arrivals[k] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            ids.append(srv.submit_ic(cfg, ic, steps, k0=4.0, u_rms=u_rms, seed=seed))
        lat = []
        for k, jid in enumerate(ids):
            if label and k % 32 == 0:
                print(f"bench: {label} stream, {k}/{len(ids)} jobs done, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
            while True:
                r = srv.result(jid, want_state=False)
                if r is not None:
                    break
                time.sleep(0.0005)
            if r[1].state != G.JOB_DONE:
                raise RuntimeError(f"serving job {jid} failed: state {r[1].state} error {r[1].error}")
            lat.append(r[1].done_s - r[1].submit_s)
        wall = time.perf_counter() - t0
        if stats is not None:
            stats.update(srv.stats())
    return wall, np.array(lat)


def run_sequential(G, jobs, steps, label=""):
    """The same jobs one at a time on plain contexts (same device initial
    conditions, nuc_step); returns the wall time."""
    t0 = time.perf_counter()
    for k, (cfg, ic, u_rms, seed, _) in enumerate(jobs):
        if label and k % 32 == 0:
            print(f"bench: {label} sequential, {k}/{len(jobs)} jobs done, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        with G.Context(cfg) as ctx:
            if ic == G.IC_HIT:
                ctx.init_hit_device(k0=4.0, u_rms=u_rms, seed=seed)
            else:
                ctx.init_tgv_device()
            ctx.step(steps)
    return time.perf_counter() - t0


def serving_leg(G, rate_256, njobs=SERVE_JOBS, steps=SERVE_STEPS, nmax=SERVE_NMAX):
    """The serving stream in FP64 and in mixed precision: served throughput and
    latency against sequential plain solves of the same jobs, and the ceiling
    any schedule could reach (the stream's cell-updates at the 256^3 box's rate)."""
    import numpy as np
    jobs, arrivals = serving_jobs(G, njobs, SERVE_NMIN, nmax, SERVE_RATE, SERVE_SEED)
    cell_updates = float(sum(n ** 3 for *_, n in jobs)) * 3 * steps
    out = {"jobs": njobs, "steps_per_job": steps, "grid_n": [SERVE_NMIN, nmax], "poisson_rate_per_s": SERVE_RATE,
           "hbm_budget_gib": SERVE_HBM_GIB, "max_active": SERVE_MAX_ACTIVE, "pack_cells": SERVE_PACK_CELLS,
           "cell_updates": cell_updates}
    for label, prec in (("fp64", G.FP64), ("mixed", G.MIXED)):
        stream = with_precision(jobs, prec)
        # kernels, allocator pools and the autotuner's per-shape cache warmed
        # as in a running server (one tuning per distinct box shape), untimed
        run_stream(G, stream[:2], None, 2)
        seen = set()
        for cfg, ic, u_rms, seed, n in stream:
            if n not in seen:
                seen.add(n)
                with G.Context(cfg) as tctx:
                    tctx.init_tgv_device()
                    tctx.autotune()
        st = {}
        wall, lat = run_stream(G, stream, arrivals, steps, stats=st, label=label)
        seq = run_sequential(G, stream, steps, label=label)
        rate = rate_256[label]
        ceil = njobs / (cell_updates / (rate * 1e6))
        out[label] = {"jobs_per_s": round(njobs / wall, 3), "wall_s": round(wall, 2),
                      "sequential_jobs_per_s": round(njobs / seq, 3), "sequential_wall_s": round(seq, 2),
                      "speedup_vs_sequential": round(seq / wall, 3),
                      "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                      "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
                      "ceiling_jobs_per_s": round(ceil, 3), "fraction_of_ceiling": round(njobs / wall / ceil, 3),
                      "ceiling_speedup_vs_sequential": round(ceil / (njobs / seq), 3),
                      "peak_active": int(st.get("peak_active", 0)), "peak_bytes": int(st.get("peak_bytes", 0)),
                      "single_slices": int(st.get("single_slices", 0)), "pack_slices": int(st.get("pack_slices", 0))}
        print(f"bench: {label} serving {json.dumps(out[label])}", file=sys.stderr, flush=True)
    return out




def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-slots", action="store_true", help="print per-kernel times to stderr")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-weno5", action="store_true", help="skip the WENO5-JS secondary figure")
    ap.add_argument("--no-configs", action="store_true", help="skip the other named workloads' rates")
    ap.add_argument("--serve-jobs", type=int, default=SERVE_JOBS,
                    help="jobs of the serving stream (default: the full 256-job stream)")
    ap.add_argument("--serve-steps", type=int, default=SERVE_STEPS)
    ap.add_argument("--serve-nmax", type=int, default=SERVE_NMAX)
    ap.add_argument("--cpu-steps", type=int, default=2, help="full RK3 steps of the CPU baseline")
    ap.add_argument("--precision", choices=("fp64", "mixed", "fp32"), default="fp64",
                    help="precision of the timed leg (mixed, fp32: for profiling those paths' kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one box per GPU and measures one MI355X; "
              "--gpus above 1 is not supported (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import nuc3d_amd as G

    n = args.n
    cfg = workload_config(G, n, {"fp64": G.FP64, "mixed": G.MIXED, "fp32": G.FP32}[args.precision])
    q = G.init_hit(cfg, k0=4.0, u_rms=1.0, seed=20261202)
    cells = n ** 3
    print(f"bench: FP64 {n}^3, {args.steps} steps", file=sys.stderr, flush=True)
    secs, st64, prof, q_end, dt = stage_stats(G, cfg, q, args.steps, args.warmup)
    if not np.all(np.isfinite(q_end)):
        print("bench.py: state became non-finite", file=sys.stderr)
        sys.exit(3)
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
    traffic, traffic_src = measured_traffic(name, n)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": name,
                "avg_launch_ms": round(avg_ms, 4), "alg_bytes": alg_bytes}
    if traffic:
        roofline["traffic_source"] = traffic_src
        roofline["traffic_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 2)
    # the sweeps' binding resource: FP64 VALU issue against its floor
    valu, valu_src = measured_valu(name, n)
    if valu:
        roofline["valu_issue_frac"] = round(valu / VALU_WAVE_INSTS_PER_S / (avg_ms * 1e-3), 4)
        roofline["valu_source"] = valu_src
    stage_ms = ms_per_step / 3.0
    if args.profile_slots:
        for k, (t, c) in prof.items():
            print(f"  {k:10s} total {t:9.3f} ms  launches {c:5d}  avg {t / max(c, 1):8.4f} ms", file=sys.stderr)

    def reduced_leg(label, prec):
        # the same workload at a reduced precision, timed the same way, with
        # its statistics' errors against the FP64 run after the same steps
        print(f"bench: {label}", file=sys.stderr, flush=True)
        rsecs, strd, rprof, _, _ = stage_stats(G, workload_config(G, n, prec), q, args.steps, args.warmup)
        rel = np.abs(strd - st64) / np.maximum(np.abs(st64), 1e-300)
        eps64, epsrd = st64[25] + st64[26], strd[25] + strd[26]
        if args.profile_slots:
            for k, (t, c) in rprof.items():
                print(f"  {label} {k:10s} avg {t / max(c, 1):8.4f} ms", file=sys.stderr)
        return {"value": round(cells * 3 * args.steps / rsecs / 1e6, 3),
                "ms_per_step": round(rsecs / args.steps * 1e3, 4),
                "speedup_vs_fp64": round(secs / rsecs, 4),
                "rel_err_vs_fp64": {"KE": float(rel[15]), "dissipation": float(abs(epsrd - eps64) / eps64),
                                    **{STAT_NAMES[i]: float(rel[i]) for i in (12, 16, 17, 18, 19, 20, 21, 24, 25, 26)}},
                "after_steps": args.warmup + args.steps}

    mixed = None
    if not args.no_mixed:
        mixed = reduced_leg("mixed", G.MIXED)
        mixed["tke_budget"] = budget_error(G, n, q)
    fp32 = None
    if not args.no_fp32:
        fp32 = reduced_leg("FP32", G.FP32)

    # secondary figure: the same box and initial state with WENO5-JS (explicit
    # five-point reconstruction, no tridiagonal solve) in place of CRWENO5
    weno5 = None
    if not args.no_weno5:
        print("bench: WENO5-JS", file=sys.stderr, flush=True)
        weno5 = {"config": f"decaying compressible isotropic turbulence {n}^3, WENO5-JS x3 + LF splitting + "
                           "cd8 viscous + TVD-RK3, CFL 0.5"}
        for label, prec in (("fp64", G.FP64), ("mixed", G.MIXED)):
            if label == "mixed" and args.no_mixed:
                continue
            wsecs, _, wprof, _, _ = stage_stats(G, workload_config(G, n, prec, G.WENO5JS), q, args.steps,
                                                args.warmup)
            wname, (wtot, wl) = max(wprof.items(), key=lambda kv: kv[1][0])
            weno5[label] = {"value": round(cells * 3 * args.steps / wsecs / 1e6, 3),
                            "ms_per_step": round(wsecs / args.steps * 1e3, 4),
                            "dominant_kernel": wname, "avg_launch_ms": round(wtot / max(wl, 1), 4)}

    configs = None if args.no_configs else config_rates(G)

    serving = None
    if not args.no_serving:
        print("bench: serving stream", file=sys.stderr, flush=True)
        rates = {"fp64": value, "mixed": mixed["value"] if mixed else value}
        serving = serving_leg(G, rates, njobs=args.serve_jobs, steps=args.serve_steps, nmax=args.serve_nmax)

    cpu = None
    if not args.no_cpu_baseline:
        # the GPU host's CPU share (OMP_NUM_THREADS is set to it), else every core
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        print(f"bench: CPU baseline, {threads} threads, {args.cpu_steps} steps", file=sys.stderr, flush=True)
        cpu = cpu_baseline(q, n, threads, steps=args.cpu_steps)

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
        "dtype": {"fp64": "f64", "mixed": "f64 state, f32 stencils", "fp32": "f32 state, f32 stencils"}[args.precision],
        "data": "synthetic (seeded random-phase solenoidal HIT field, seed 20261202)",
        "config": {"workload": f"decaying compressible isotropic turbulence {n}^3, CRWENO5 x3 + LF splitting + "
                               "cd8 viscous + TVD-RK3, CFL 0.5, FP64",
                   "grid": [n, n, n], "rank_grid": [1, 1, 1], "mach_t": 0.3, "reynolds": 200.0,
                   "stage_ms": round(stage_ms, 4), "dt_last": dt,
                   "grad_planes_autotuned": TILE_SHAPES.get(G.FP64),
                   "state_upload_ms": round(PCIE.get("upload_ms", 0.0), 2),
                   "state_download_ms": round(PCIE.get("download_ms", 0.0), 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "fp32": fp32,
        "weno5js": weno5,
        "configs": configs,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
