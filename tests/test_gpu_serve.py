"""The serving loop on the GPU: every job's result is bit-identical to the same
job run alone on a plain context (per-kernel launches, no graphs), whatever
else runs beside it; admission respects the HBM budget; jobs submitted while
others run are admitted as slots free up and overlap them on the device.
"""
import time

import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu


def _job(n, scheme, mach, re, seed, tgv):
    cfg = G.make_config(n=n, scheme=scheme, vis=G.CD6, riemann=G.LF, model=G.NAVIERSTOKES3D, mach=mach,
                        reynolds=re, cfl=0.5)
    q = G.init_tgv(cfg) if tgv else G.init_hit(cfg, k0=3.0, u_rms=0.5, seed=seed)
    return cfg, q


JOBS = [
    ((32, 32, 32), G.WENO5JS, 0.1, 1600.0, 1, True, 7),
    ((48, 40, 32), G.CRWENO5, 0.3, 800.0, 2, False, 5),
    ((24, 24, 24), G.WENO5Z, 0.5, 400.0, 3, False, 9),
    ((64, 64, 64), G.WENO5JS, 0.2, 3200.0, 4, False, 4),
    ((40, 40, 40), G.CRWENO5, 0.4, 1200.0, 5, True, 6),
    ((32, 48, 24), G.WENO5Z, 0.15, 2000.0, 6, True, 8),
]


def _alone(cfg, q, steps):
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        ctx.step(steps)
        return ctx.download()


def _wait(srv, jid, timeout=120):
    t0 = time.time()
    while time.time() - t0 < timeout:
        r = srv.result(jid)
        if r is not None:
            return r
        time.sleep(0.002)
    raise TimeoutError(f"job {jid}")


def test_results_identical_alone_and_in_the_mix():
    print("serving mix", flush=True)
    ref, ids = {}, {}
    with G.Server(hbm_budget=8 << 30, max_active=4) as srv:
        for i, (n, sch, mach, re, seed, tgv, steps) in enumerate(JOBS):
            cfg, q = _job(n, sch, mach, re, seed, tgv)
            ids[i] = srv.submit(cfg, q, steps)
            ref[i] = (cfg, q, steps)
            if i == 2:
                time.sleep(0.05)   # later jobs arrive while the first ones run
        out = {i: _wait(srv, jid) for i, jid in ids.items()}
        st = srv.stats()
    print("server stats", st, flush=True)
    assert st["completed"] == len(JOBS) and st["failed"] == 0
    assert st["peak_active"] >= 2
    for i, (state, r) in out.items():
        cfg, q, steps = ref[i]
        alone = _alone(cfg, q, steps)
        assert r.state == G.JOB_DONE and r.info.step == steps
        assert np.array_equal(state, alone), (i, np.abs(state - alone).max())
    # overlap: some job was admitted before an earlier one had finished
    rs = [out[i][1] for i in range(len(JOBS))]
    assert any(rs[j].admit_s < rs[i].done_s for i in range(len(JOBS)) for j in range(i + 1, len(JOBS)))


def test_admission_respects_hbm_budget():
    print("serving budget", flush=True)
    cfgs = [_job((48, 48, 48), G.WENO5JS, 0.2, 1000.0, k, False) for k in range(5)]
    fp = G.footprint(cfgs[0][0])
    budget = int(2.5 * fp)
    with G.Server(hbm_budget=budget, max_active=8) as srv:
        ids = [srv.submit(c, q, 4) for c, q in cfgs]
        out = [_wait(srv, j) for j in ids]
        st = srv.stats()
    print("server stats", st, flush=True)
    assert st["peak_bytes"] <= budget and st["peak_active"] <= 2
    assert all(r.state == G.JOB_DONE for _, r in out)


def test_graph_replay_matches_launches():
    """Every replay of the captured step, with eager work and synchronisation
    between replays, equals the eager launches bit for bit."""
    cfg, q = _job((32, 32, 32), G.CRWENO5, 0.3, 1000.0, 9, False)
    eager = []
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        for _ in range(5):
            ctx.step(1)
            eager.append(ctx.download())
    with G.Context(cfg) as ctx:
        ctx.set_graph(True)
        ctx.upload(q)
        for k in range(5):
            ctx.step(1)
            st = ctx.stats()           # eager kernels between replays
            assert np.isfinite(st).all()
            assert np.array_equal(ctx.download(), eager[k]), k
        ctx.upload(eager[1])
        ctx.step_async(3)
        ctx.sync()
        assert np.array_equal(ctx.download(), eager[4])


def test_failed_job_does_not_disturb_others():
    good, q = _job((32, 32, 32), G.WENO5JS, 0.2, 1000.0, 3, True)
    bad = G.make_config(n=(18, 16, 16), npr=(4, 1, 1))
    with G.Server(hbm_budget=4 << 30, max_active=4) as srv:
        j1 = srv.submit(good, q, 3)
        j2 = srv.submit(bad, np.ones(5 * 18 * 16 * 16), 3)
        j3 = srv.submit(good, q, 3)
        r1, r2, r3 = _wait(srv, j1), _wait(srv, j2), _wait(srv, j3)
    assert r2[1].state == G.JOB_FAILED
    assert r1[1].state == G.JOB_DONE and r3[1].state == G.JOB_DONE
    assert np.array_equal(r1[0], r3[0])


def test_packed_serving_results_identical_to_alone():
    """With packing on, small compatible jobs that are idle together advance as
    one pack; each result still equals the job run alone."""
    print("serving with packs", flush=True)
    mix = [((32, 32, 32), G.CRWENO5, 0.1, 1600.0, 11, True, 6),
           ((40, 40, 40), G.CRWENO5, 0.3, 900.0, 12, False, 6),
           ((36, 36, 36), G.CRWENO5, 0.2, 2000.0, 13, False, 6),
           ((48, 48, 48), G.CRWENO5, 0.45, 500.0, 14, True, 6),
           ((32, 32, 32), G.WENO5Z, 0.25, 700.0, 15, False, 6),
           ((40, 40, 40), G.WENO5Z, 0.35, 1500.0, 16, True, 6),
           ((160, 160, 160), G.WENO5JS, 0.2, 1000.0, 17, False, 2)]
    ids, ref = {}, {}
    for i, (n, sch, mach, re, seed, tgv, steps) in enumerate(mix):
        cfg, q = _job(n, sch, mach, re, seed, tgv)
        ref[i] = (cfg, q, steps)
    with G.Server(hbm_budget=16 << 30, max_active=8, steps_per_slice=2, pack_cells=1 << 20) as srv:
        # submitted back to back, so that compatible jobs are idle together
        for i, (cfg, q, steps) in ref.items():
            ids[i] = srv.submit(cfg, q, steps)
        out = {i: _wait(srv, jid) for i, jid in ids.items()}
        st = srv.stats()
    print("server stats", st, flush=True)
    assert st["completed"] == len(mix) and st["pack_slices"] > 0
    for i, (state, r) in out.items():
        cfg, q, steps = ref[i]
        assert r.state == G.JOB_DONE and r.info.step == steps
        assert np.array_equal(state, _alone(cfg, q, steps)), i


def test_autotuned_tile_shape_keeps_results():
    """nuc_autotune picks the gradient kernel's tile shape by timing; the state
    after several steps is unchanged by the choice, and the choice is shared by
    contexts of the same shape."""
    cfg, q = _job((48, 40, 64), G.CRWENO5, 0.3, 900.0, 21, False)
    ref = _alone(cfg, q, 3)
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        g = ctx.autotune()
        assert g in (4, 8, 16, 32)
        ctx.step(3)
        assert np.array_equal(ctx.download(), ref)
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        assert ctx.autotune() == g
        ctx.set_graph(True)
        ctx.step(3)
        assert np.array_equal(ctx.download(), ref)
    for planes in (1, 4, 8, 16, 32, 64):
        with G.Context(cfg) as ctx:
            ctx.upload(q)
            ctx.set_tile_shape(planes)
            assert ctx.tile_shape() == planes
            ctx.step(3)
            assert np.array_equal(ctx.download(), ref), planes


def test_jobs_with_device_initial_conditions():
    """Jobs submitted as an initial-condition spec (no host state): the result
    equals the same job started from the device-generated state alone."""
    a = G.make_config(n=(32, 32, 32), scheme=G.CRWENO5, vis=G.CD6, mach=0.2, reynolds=900.0, cfl=0.5)
    b = G.make_config(n=(40, 36, 32), scheme=G.WENO5Z, vis=G.CD6, mach=0.3, reynolds=600.0, cfl=0.5)
    c = G.make_config(n=(32, 32, 8), L=(10.0, 10.0, 2.5), scheme=G.WENO5JS, model=G.EULER3D, cfl=0.5)
    d = G.make_config(n=(16, 16, 16), scheme=G.WENO5JS, mach=0.3, cfl=0.5)
    with G.Server(hbm_budget=4 << 30, max_active=4) as srv:
        ja = srv.submit_ic(a, G.IC_HIT, 4, k0=3.0, u_rms=0.5, seed=12)
        jb = srv.submit_ic(b, G.IC_TGV, 4)
        jc = srv.submit_ic(c, G.IC_IVC, 4)
        jd = srv.submit_ic(d, G.IC_DEFAULT, 4)
        ra, rb, rc, rd = _wait(srv, ja), _wait(srv, jb), _wait(srv, jc), _wait(srv, jd)
    for cfg, res, init in ((a, ra, lambda c: c.init_hit_device(k0=3.0, u_rms=0.5, seed=12)),
                           (b, rb, lambda c: c.init_tgv_device()),
                           (c, rc, lambda c: c.init_ivc_device()),
                           (d, rd, lambda c: c.init_default_device())):
        with G.Context(cfg) as ctx:
            init(ctx)
            ctx.step(4)
            ref = ctx.download()
        assert res[1].state == G.JOB_DONE
        assert np.array_equal(res[0], ref)


def test_serving_stream_slice_identical_alone():
    """Jobs of the benchmark's serving stream (its own generator and seed: the
    boxes of at most 48^3 among its 256 jobs, plus a 64^3 one) through a
    server configured as the benchmark configures it, so that packs form; in
    FP64 and mixed precision every result equals the job run alone."""
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import bench
    jobs, _ = bench.serving_jobs(G, bench.SERVE_JOBS, bench.SERVE_NMIN, bench.SERVE_NMAX, bench.SERVE_RATE,
                                 bench.SERVE_SEED)
    sl = [j for j in jobs if j[4] <= 48][:8] + [j for j in jobs if j[4] == 64][:1]
    assert len(sl) == 9
    steps = 4
    for label, prec in (("fp64", G.FP64), ("mixed", G.MIXED)):
        print("serving stream slice", label, [j[4] for j in sl], flush=True)
        stream = bench.with_precision(sl, prec)
        with G.Server(hbm_budget=bench.SERVE_HBM_GIB << 30, max_active=bench.SERVE_MAX_ACTIVE, steps_per_slice=2,
                      pack_cells=bench.SERVE_PACK_CELLS, autotune=True) as srv:
            ids = [srv.submit_ic(cfg, ic, steps, k0=4.0, u_rms=u_rms, seed=seed)
                   for cfg, ic, u_rms, seed, _ in stream]
            out = [_wait(srv, jid) for jid in ids]
            st = srv.stats()
        print("server stats", st, flush=True)
        assert st["completed"] == len(sl) and st["failed"] == 0
        assert st["pack_slices"] > 0
        for (cfg, ic, u_rms, seed, n), (state, r) in zip(stream, out):
            with G.Context(cfg) as ctx:
                if ic == G.IC_HIT:
                    ctx.init_hit_device(k0=4.0, u_rms=u_rms, seed=seed)
                else:
                    ctx.init_tgv_device()
                ctx.step(steps)
                ref = ctx.download()
            assert r.state == G.JOB_DONE and r.info.step == steps
            assert np.array_equal(state, ref), (label, n)
