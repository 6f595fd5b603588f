"""NUC_FP32: the mixed path's arithmetic (FP32 split fluxes, stencils and
viscous fluxes; FP64 weights of the Z-type schemes, tridiagonal solves, flux
differences and RK3 update) with the conserved state and the RHS stored in
FP32.  Uploads round the FP64 state to nearest; downloads, statistics, spectra
and averages see the FP32 state widened to FP64 exactly.

Parity is stated against the FP64 path, which is itself held to the oracle at
1e-12 (test_gpu_rhs.py, test_gpu_workloads.py):
- exact: the upload rounds to nearest (numpy's float32 conversion), the
  download widens exactly, statistics, Q fields and the stage-0 dt equal those
  of an FP64 context holding the rounded state (spectra to 1e-12: their shell
  sums have no fixed order);
- one RHS: within 1e-6 of the mixed path's RHS on the same rounded state
  (relative to the field's maximum): the only difference is the FP32 storage
  of the two partial sums between the sweeps (2^-24 relative each); against
  FP64 within the mixed path's own error plus that 1e-6;
- after 5 steps of 24^3 turbulence, every scheme the path takes: within 2e-5 of
  FP64 (relative L-infinity per variable; the mixed path's bar is 1e-5, and
  the FP32 state adds one rounding of 2^-24 per stage, 15 stages);
- 20 steps of 48^3 decaying turbulence: kinetic energy and dissipation within
  1e-5 of FP64;
- Shu-Osher (the Mach-3 shock): within 1e-4 of FP64 after 10 steps.
"""
import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu


def _hit_cfg(n, precision, scheme=G.CRWENO5, vis=G.CD8, riemann=G.LF, **kw):
    if isinstance(n, int):
        n = (n, n, n)
    return G.make_config(n=n, scheme=scheme, vis=vis, riemann=riemann, model=G.NAVIERSTOKES3D,
                         viscosity=G.SUTHERLAND, mach=0.3 / np.sqrt(3.0), reynolds=200.0, prandtl=0.72,
                         cfl=0.5, precision=precision, **kw)


def _rounded(q):
    return q.astype(np.float32).astype(np.float64)


def _run(cfg, q, steps):
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        info = ctx.step(steps)
        return ctx.download(), ctx.stats_ex(), info


def _rel_linf(a, b, N, mom_scale=True):
    a, b = a.reshape(5, N), b.reshape(5, N)
    scale = np.abs(b).max(axis=1)
    if mom_scale:
        scale[1:4] = np.abs(b[1:4]).max()
    return np.abs(a - b).max(axis=1) / np.maximum(scale, 1e-300)


def test_fp32_upload_rounds_download_widens_stats_are_fp64_stats():
    print("fp32 upload/download/stats", flush=True)
    n = 32
    c32, c64 = _hit_cfg(n, G.FP32), _hit_cfg(n, G.FP64)
    q = G.init_hit(c64, k0=4.0, u_rms=1.0, seed=31)
    qr = _rounded(q)
    with G.Context(c32) as a, G.Context(c64) as b:
        a.upload(q)
        b.upload(qr)
        assert np.array_equal(a.download(), qr)
        assert np.array_equal(a.stats_ex(), b.stats_ex())
        # the shell sums accumulate in LDS in no fixed order: equal to rounding
        for sa, sb in ((a.spectra_kinetic(), b.spectra_kinetic()),
                       (a.spectrum_scalar(G.SPEC_PRESSURE), b.spectrum_scalar(G.SPEC_PRESSURE))):
            assert np.abs(sa - sb).max() <= 1e-12 * np.abs(sb).max()
        assert np.array_equal(a.q_fields(), b.q_fields())
        # the download did not disturb the state: a second one is the same
        assert np.array_equal(a.download(), qr)


def test_fp32_rhs_is_mixed_rhs_through_fp32_storage():
    print("fp32 rhs", flush=True)
    n = (32, 24, 20)
    for scheme, rie in ((G.CRWENO5, G.LF), (G.WENO5Z, G.AUSM), (G.WENO7JS, G.AUSMP)):
        c32, cmx, c64 = (_hit_cfg(n, p, scheme=scheme, riemann=rie) for p in (G.FP32, G.MIXED, G.FP64))
        q = _rounded(G.init_hit(c64, k0=3.0, u_rms=1.0, seed=32))
        outs = {}
        for lab, c in (("fp32", c32), ("mixed", cmx), ("fp64", c64)):
            with G.Context(c) as ctx:
                ctx.upload(q)
                outs[lab] = ctx.rhs()
        N = n[0] * n[1] * n[2]
        r32, dt32 = outs["fp32"]
        rmx, dtmx = outs["mixed"]
        r64, dt64 = outs["fp64"]
        e_mx = _rel_linf(r32, rmx, N, mom_scale=False)
        e_64 = _rel_linf(r32, r64, N, mom_scale=False)
        m_64 = _rel_linf(rmx, r64, N, mom_scale=False)
        print(scheme, rie, "rhs vs mixed", e_mx, "vs fp64", e_64, "mixed vs fp64", m_64, "dt", dt32, dt64,
              flush=True)
        assert dt32 == dt64 and dtmx == dt64
        assert np.all(e_mx <= 1e-6), e_mx
        # against FP64: the mixed path's own error (FP32 fluxes of order
        # p = 1/(gamma M^2) = 24 differenced to an RHS of order 1) plus that 1e-6
        assert np.all(e_64 <= m_64 + 1e-6), (e_64, m_64)
        assert np.all(m_64 <= 5e-5), m_64


FP32_CASES = [
    (G.WENO5JS, G.CD6, G.LF),
    (G.WENO5Z, G.CD6, G.AUSM),
    (G.WENO7JS, G.CD8, G.LF),
    (G.CRWENO5, G.CD8, G.LF),
    (G.CRWENO5, G.CD8, G.AUSMP),
    (G.WENO3JS, G.CD2, G.AUSMP),
    (G.UPWIND1ST, G.CD6, G.LF),
    (G.HCCS78, G.CD8, G.AUSMP),
]


@pytest.mark.parametrize("case", FP32_CASES, ids=lambda c: f"s{c[0]}-v{c[1]}-r{c[2]}")
def test_fp32_close_to_fp64_every_scheme(case):
    scheme, vis, rie = case
    print("fp32 scheme", case, flush=True)
    n = 24
    c64, c32 = _hit_cfg(n, G.FP64, scheme=scheme, vis=vis, riemann=rie), _hit_cfg(n, G.FP32, scheme=scheme,
                                                                                   vis=vis, riemann=rie)
    q = G.init_hit(c64, k0=3.0, u_rms=1.0, seed=7)
    q64, s64, i64 = _run(c64, q, 5)
    q32, s32, i32 = _run(c32, q, 5)
    err = _rel_linf(q32, q64, n ** 3)
    print(case, "state rel Linf", err, "KE rel", abs(s32[15] - s64[15]) / s64[15], "dt", i32[0], i64[0], flush=True)
    assert np.all(err <= 2e-5), err
    assert abs(s32[15] - s64[15]) <= 2e-6 * s64[15]


def test_fp32_decaying_turbulence_statistics_within_1e5():
    print("fp32 decaying turbulence 48^3", flush=True)
    n, steps = 48, 20
    c64, c32 = _hit_cfg(n, G.FP64), _hit_cfg(n, G.FP32)
    q = G.init_hit(c64, k0=4.0, u_rms=1.0, seed=20261202)
    q64, s64, _ = _run(c64, q, steps)
    q32, s32, _ = _run(c32, q, steps)
    rel = np.abs(s32 - s64) / np.maximum(np.abs(s64), 1e-300)
    print("state rel Linf", _rel_linf(q32, q64, n ** 3), flush=True)
    print("stat rel errors", rel, flush=True)
    eps64, eps32 = s64[25] + s64[26], s32[25] + s32[26]
    assert rel[15] <= 1e-5, rel[15]                              # kinetic energy
    assert abs(eps32 - eps64) <= 1e-5 * eps64, (eps32, eps64)    # dissipation
    assert rel[12] <= 1e-5 and rel[16] <= 1e-5                   # enstrophy, u_rms


def test_fp32_shu_osher_close_to_fp64():
    print("fp32 Shu-Osher 64x12x12", flush=True)
    n = (64, 12, 12)
    kw = dict(n=n, L=(10.0, 10.0 * n[1] / n[0], 10.0 * n[2] / n[0]), scheme=G.WENO5Z, vis=G.CD6, riemann=G.AUSM,
              model=G.NAVIERSTOKES3D, viscosity=G.SUTHERLAND, mach=1.0 / np.sqrt(1.4), reynolds=2000.0,
              prandtl=0.72, cfl=0.4)
    c64, c32 = G.make_config(precision=G.FP64, **kw), G.make_config(precision=G.FP32, **kw)
    q = G.init_shu_osher(c64, pert=0.05, seed=20261602)
    a, _, _ = _run(c64, q, 10)
    b, _, _ = _run(c32, q, 10)
    err = _rel_linf(b, a, n[0] * n[1] * n[2])
    print("Shu-Osher FP32 vs FP64 rel Linf", err, flush=True)
    assert np.all(err <= 1e-4), err


def test_fp32_rk3_stages_and_device_initial_states():
    """nuc_rk3_stage three times equals nuc_step once; the device initial
    conditions are the FP64 ones rounded to nearest."""
    print("fp32 stages and device ICs", flush=True)
    n = 24
    c32, c64 = _hit_cfg(n, G.FP32), _hit_cfg(n, G.FP64)
    q = G.init_hit(c64, k0=3.0, u_rms=1.0, seed=5)
    with G.Context(c32) as a:
        a.upload(q)
        a.step(1)
        one = a.download()
        a.upload(q)
        for s in range(3):
            a.rk3_stage(s)
        assert np.array_equal(a.download(), one)
    with G.Context(c32) as a, G.Context(c64) as b:
        a.init_tgv_device()
        b.init_tgv_device()
        assert np.array_equal(a.download(), _rounded(b.download()))
        a.init_hit_device(k0=3.0, u_rms=1.0, seed=9)
        b.init_hit_device(k0=3.0, u_rms=1.0, seed=9)
        assert np.array_equal(a.download(), _rounded(b.download()))


def test_fp32_not_packable_and_serves_alone():
    """An FP32 context is not packable (the packed kernels take FP64 state);
    the serving loop runs an FP32 job alone, bit-identical to a plain context."""
    print("fp32 pack/serve", flush=True)
    c32 = _hit_cfg(24, G.FP32, scheme=G.WENO5Z, vis=G.CD6)
    q = G.init_hit(c32, k0=3.0, u_rms=1.0, seed=3)
    with G.Context(c32) as a, G.Context(c32) as b:
        assert not G.pack_compatible(a, b)
        with pytest.raises(G.NucError):
            G.Pack([a, b])
        a.upload(q)
        a.step(4)
        alone = a.download()
    import time
    with G.Server(hbm_budget=4 << 30, max_active=4, pack_cells=4 * 40 ** 3) as srv:
        jid = srv.submit(c32, q, 4)
        t0 = time.time()
        r = None
        while r is None and time.time() - t0 < 60:
            r = srv.result(jid)
            time.sleep(0.002)
    assert r is not None
    state, res = r
    assert res.state == G.JOB_DONE
    assert np.array_equal(state, alone)
