"""Solver-level parity: the fused HIP stage (split, reconstruct, viscous,
RHS, TVD-RK3 with on-device periodic wrap) against the oracle's restatement of
NaiverStokesData3d::solve + PDEData3d::solve on the same seeded fields.

Bar: relative L-infinity per conserved variable <= 1e-12 after 10 RK3 steps
(the FP64 bar), measured as max|gpu - oracle| / max|oracle| per variable.

WENO7-Z and hccs: their 7-point smoothness indicators are expanded quadratic
forms with coefficients up to 17246 that cancel to O(h^2), so the reference
algorithm itself moves by ~1e-11 after 10 steps (and ~1e-9 in one RHS) when its
input state is perturbed by one ulp.  No implementation short of bit-identical
arithmetic can meet 1e-12 there.  For those schemes the test measures that
floor with the oracle (same state, perturbed by +-1 ulp) and holds the GPU to
max(1e-12, 3 x floor) per variable.
"""
import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


def _cfgs(n, npr, scheme, vis, riemann, model, mach=0.1, reynolds=1600.0, cfl=0.5):
    kw = dict(n=n, npr=npr, scheme=scheme, vis=vis, riemann=riemann, model=model, mach=mach,
              reynolds=reynolds, cfl=cfl)
    return G.make_config(**kw), O.make_config(**kw)


def _perturbed_tgv(oc, seed, amp=0.2):
    q = O.init_tgv(oc)
    N = oc.nx * oc.ny * oc.nz
    rng = np.random.default_rng(seed)
    q = q.reshape(5, N)
    # smooth-ish perturbation so that w != 0 and the flow is fully 3-D
    q[3] += amp * np.sin(np.linspace(0, 40, N) + rng.random())
    q[0] *= 1.0 + 0.01 * rng.standard_normal(N)
    return q.ravel().copy()


ILL_CONDITIONED = (G.WENO7JS, G.HCCS, G.HCCS78, G.HCCS77)


def _ulp_floor(oc, q, steps):
    """Per-variable change of the oracle's result when q is perturbed by +-1 ulp."""
    N = oc.nx * oc.ny * oc.nz
    rng = np.random.default_rng(99)
    qp = q * (1 + 2.2e-16 * rng.choice([-1.0, 0.0, 1.0], size=q.size))
    outs = []
    for state in (q, qp):
        s = O.Solver(oc)
        s.set_state(state)
        if steps == 0:
            outs.append(s.rhs()[0])
        else:
            s.step(steps)
            outs.append(s.get_state())
    a, b = (o.reshape(5, N) for o in outs)
    return np.abs(a - b).max(axis=1) / np.maximum(np.abs(a).max(axis=1), 1e-300)


def _rel_linf(a, b, N):
    a = a.reshape(5, N)
    b = b.reshape(5, N)
    return np.array([np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300) for v in range(5)])


CASES = [
    # n, rank grid, scheme, vis, riemann, model
    ((16, 16, 16), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((16, 16, 16), (1, 1, 1), G.WENO5Z, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((16, 16, 16), (1, 1, 1), G.WENO7JS, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((16, 16, 16), (1, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((16, 16, 16), (1, 1, 1), G.HCCS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((24, 20, 18), (1, 1, 1), G.WENO5JS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((24, 20, 18), (2, 2, 2), G.CRWENO5, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((32, 32, 32), (2, 1, 2), G.WENO5Z, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((32, 32, 32), (1, 2, 1), G.HCCS, G.CD8, G.AUSM, G.NAVIERSTOKES3D),
    ((32, 24, 16), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.EULER3D),
    ((20, 20, 20), (1, 1, 1), G.WENO7JS, G.CD6, G.AUSM, G.EULER3D),
    # weno3js, upwind1st, hccs78 and cd2 (src/weno3js.cpp, src/upwind1st.cpp,
    # src/hccs78.cpp, src/centraldifference2nd.cpp)
    ((16, 16, 16), (1, 1, 1), G.WENO3JS, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((24, 20, 18), (2, 1, 1), G.UPWIND1ST, G.CD8, G.AUSM, G.NAVIERSTOKES3D),
    ((16, 16, 16), (1, 1, 1), G.HCCS78, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((32, 32, 32), (1, 2, 1), G.HCCS78, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((24, 20, 18), (1, 1, 1), G.WENO5JS, G.CD2, G.LF, G.NAVIERSTOKES3D),
    ((20, 20, 20), (2, 2, 1), G.WENO3JS, G.CD2, G.AUSM, G.NAVIERSTOKES3D),
    ((32, 24, 16), (1, 1, 1), G.UPWIND1ST, G.CD6, G.LF, G.EULER3D),
    # hccs77 (src/hccs77.cpp)
    ((16, 16, 16), (1, 1, 1), G.HCCS77, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((32, 32, 32), (2, 1, 1), G.HCCS77, G.CD8, G.AUSMP, G.NAVIERSTOKES3D),
    # AUSM+ splitting (src/physicsModel.cpp:932-1179)
    ((16, 16, 16), (1, 1, 1), G.WENO5JS, G.CD6, G.AUSMP, G.NAVIERSTOKES3D),
    ((24, 20, 18), (2, 2, 2), G.CRWENO5, G.CD8, G.AUSMP, G.NAVIERSTOKES3D),
    ((20, 20, 20), (1, 1, 1), G.WENO3JS, G.CD2, G.AUSMP, G.EULER3D),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-s{c[2]}-v{c[3]}-r{c[4]}-m{c[5]}")
def test_rhs_matches_oracle(case):
    n, npr, scheme, vis, rie, model = case
    print(f"rhs case {case}", flush=True)
    gc, oc = _cfgs(n, npr, scheme, vis, rie, model)
    q = _perturbed_tgv(oc, 1)
    N = oc.nx * oc.ny * oc.nz
    s = O.Solver(oc)
    s.set_state(q)
    r_ref, dt_ref = s.rhs()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        r_gpu, dt_gpu = ctx.rhs()
    scale = np.abs(r_ref.reshape(5, N)).max(axis=1)
    err = np.abs(r_gpu.reshape(5, N) - r_ref.reshape(5, N)).max(axis=1) / np.maximum(scale, 1e-300)
    print("rhs rel err per variable", err, flush=True)
    tol = np.full(5, 1e-11)
    if scheme in ILL_CONDITIONED:
        tol = np.maximum(tol, 3.0 * _ulp_floor(oc, q, 0))
        print("1-ulp floor x3", tol, flush=True)
    assert np.all(err <= tol), (err, tol)
    assert abs(dt_gpu - dt_ref) <= 1e-14 * dt_ref


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-s{c[2]}-v{c[3]}-r{c[4]}-m{c[5]}")
def test_ten_steps_match_oracle(case):
    n, npr, scheme, vis, rie, model = case
    print(f"10-step case {case}", flush=True)
    gc, oc = _cfgs(n, npr, scheme, vis, rie, model)
    q = _perturbed_tgv(oc, 2)
    N = oc.nx * oc.ny * oc.nz
    s = O.Solver(oc)
    s.set_state(q)
    dt_ref, res_ref, t_ref = s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        dt_gpu, res_gpu, t_gpu = ctx.step(10)
        q_gpu = ctx.download()
    err = _rel_linf(q_gpu, q_ref, N)
    print("10-step rel Linf per variable", err, flush=True)
    tol = np.full(5, 1e-12)
    if scheme in ILL_CONDITIONED:
        tol = np.maximum(tol, 3.0 * _ulp_floor(oc, q, 10))
        print("1-ulp floor x3", tol, flush=True)
    assert np.all(err <= tol), (err, tol)
    assert abs(dt_gpu - dt_ref) <= 1e-13 * dt_ref
    assert abs(t_gpu - t_ref) <= 1e-13 * t_ref
    assert abs(res_gpu - res_ref) <= 1e-9 * abs(res_ref)


def test_tgv64_weno5js_ten_steps():
    """The Taylor-Green configuration: 64^3, WENO5-JS, LF, cd6, Re 1600, M 0.1, CFL 0.5.

    w = 0 initially, so rho*w is normalised by the momentum scale
    max_c |rho u_c| instead of its own (tiny) maximum."""
    gc, oc = _cfgs((64, 64, 64), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.NAVIERSTOKES3D)
    q = O.init_tgv(oc)
    assert np.array_equal(q, G.init_tgv(gc))
    N = 64 ** 3
    oc.nthreads = 8
    s = O.Solver(oc)
    s.set_state(q)
    s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        ctx.step(10)
        q_gpu = ctx.download()
    a = q_gpu.reshape(5, N)
    b = q_ref.reshape(5, N)
    mom = np.abs(b[1:4]).max()
    scale = np.array([np.abs(b[0]).max(), mom, mom, mom, np.abs(b[4]).max()])
    err = np.abs(a - b).max(axis=1) / scale
    print("TGV 64^3 10-step rel Linf", err, flush=True)
    print("TGV 64^3 10-step rel Linf per variable", _rel_linf(q_gpu, q_ref, N), flush=True)
    assert np.all(err <= 1e-12), err


def test_stats_match_oracle():
    gc, oc = _cfgs((32, 32, 32), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.NAVIERSTOKES3D)
    q = _perturbed_tgv(oc, 4)
    s = O.Solver(oc)
    s.set_state(q)
    ref = s.stats()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        got = ctx.stats()
    # means of velocity gradients, vorticity and Q vanish on a periodic box
    # (entries 0-13): compare those against the gradient scale instead
    scale = np.abs(ref)
    scale[:14] = np.maximum(scale[:14], ref[18])
    err = np.abs(got - ref) / scale
    print("stats rel err", err, flush=True)
    # FP64 sums in another order: 5e-14 at worst (the skewness, a ratio of
    # third and second moments of the gradients), 3e-15 elsewhere
    assert np.all(err <= 1e-12), err


def test_state_roundtrip_and_errors():
    gc, _ = _cfgs((16, 16, 16), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.NAVIERSTOKES3D)
    q = np.random.default_rng(0).random(5 * 16 ** 3) + 1.0
    with G.Context(gc) as ctx:
        ctx.upload(q)
        assert np.array_equal(ctx.download(), q)
    bad = G.make_config(n=(18, 16, 16), npr=(4, 1, 1))
    with pytest.raises(G.NucError):
        G.Context(bad)


FORCING_CASES = [
    # n, rank grid, scheme, vis, riemann, model, k_f, helical amplitude
    ((16, 16, 16), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.NAVIERSTOKES3D, 2.0, 0.1),
    ((24, 20, 16), (2, 1, 1), G.CRWENO5, G.CD8, G.AUSM, G.NAVIERSTOKES3D, 1.0, 0.0),
    ((16, 16, 16), (1, 1, 1), G.WENO5Z, G.CD6, G.LF, G.EULER3D, 2.0, 0.3),
    # x and y lines above 512 cells: segmented sweeps, 8 cells per thread in the transforms
    ((1024, 12, 12), (1, 1, 1), G.WENO5Z, G.CD6, G.AUSM, G.NAVIERSTOKES3D, 2.0, 0.1),
    ((12, 768, 12), (1, 2, 1), G.CRWENO5, G.CD8, G.AUSM, G.NAVIERSTOKES3D, 1.0, 0.0),
]


def _forced_cfgs(case):
    n, npr, scheme, vis, rie, model, kf, hel = case
    kw = dict(n=n, npr=npr, scheme=scheme, vis=vis, riemann=rie, model=model, mach=0.1, reynolds=1600.0,
              forcing_k=kf, forcing_amp=0.05, forcing_amp_hel=hel)
    return (G.make_config(forcing=G.FORCING_CONST, **kw), O.make_config(forcing=O.FORCE_CONST, **kw))


@pytest.mark.parametrize("case", FORCING_CASES, ids=lambda c: f"{c[0]}-s{c[2]}-m{c[5]}-k{c[6]}")
def test_forced_rhs_and_ten_steps_match_oracle(case):
    """Const forcing (src/physicsModel.cpp:1368-1677): pruned separable DFT on
    the device against the oracle's direct DFT at the shell modes."""
    print(f"forcing case {case}", flush=True)
    gc, oc = _forced_cfgs(case)
    q = _perturbed_tgv(oc, 3)
    N = oc.nx * oc.ny * oc.nz
    s = O.Solver(oc)
    s.set_state(q)
    r_ref, dt_ref = s.rhs()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        r_gpu, dt_gpu = ctx.rhs()
    err = _rel_linf(r_gpu, r_ref, N)
    print("forced rhs rel err", err, flush=True)
    assert np.all(err <= 1e-11), err
    # the forcing itself is a visible part of the RHS
    oc0 = O.make_config(n=case[0], npr=case[1], scheme=case[2], vis=case[3], riemann=case[4], model=case[5],
                        mach=0.1, reynolds=1600.0)
    s0 = O.Solver(oc0)
    s0.set_state(q)
    r0 = s0.rhs()[0]
    assert np.abs(r_ref - r0).max() > 1e-6 * np.abs(r_ref).max()
    s = O.Solver(oc)
    s.set_state(q)
    s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        ctx.step(10)
        q_gpu = ctx.download()
    err = _rel_linf(q_gpu, q_ref, N)
    print("forced 10-step rel Linf", err, flush=True)
    tol = np.full(5, 1e-12)
    if max(case[0]) > 512:
        # the long-line boxes leave their 12-cell sides under-resolved: the
        # same conditioning allowance as LONG_CASES
        tol = np.maximum(tol, 3.0 * _ulp_floor(oc, q, 10))
        print("max(1e-12, 1-ulp floor x3)", tol, flush=True)
    assert np.all(err <= tol), (err, tol)


# Long lines: every chunk length R (lane h holds R consecutive interfaces,
# R in {2, 3, 4, 5, 6, 8, 12, 16} for n up to 64, 96, 128, 160, 192, 256, 384, 512) in every sweep
# direction, both with n = 32 R (the kernels without tail checks) and with a
# ragged tail, with closure rows inside the line (rank grids).  Thin boxes keep
# the oracle's work small; the short sides are the minimum the halos allow.
LONG_CASES = [
    # n, rank grid, scheme, vis, riemann, model
    ((256, 12, 12), (1, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 256, 12), (1, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 12, 256), (1, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((128, 12, 10), (1, 1, 1), G.WENO5Z, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((10, 12, 160), (1, 1, 1), G.WENO5JS, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((12, 192, 12), (1, 1, 1), G.WENO5Z, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((512, 10, 12), (1, 1, 1), G.CRWENO5, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 10, 480), (1, 1, 1), G.WENO5JS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((100, 12, 12), (2, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 12, 384), (1, 1, 3), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((256, 12, 12), (1, 1, 1), G.HCCS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 160, 12), (1, 1, 1), G.WENO7JS, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 12, 512), (1, 1, 1), G.WENO5Z, G.CD6, G.LF, G.EULER3D),
    ((256, 12, 12), (1, 1, 1), G.WENO3JS, G.CD2, G.LF, G.NAVIERSTOKES3D),
    ((12, 160, 12), (1, 1, 1), G.HCCS78, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 12, 100), (1, 1, 2), G.UPWIND1ST, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 12, 256), (1, 1, 1), G.HCCS77, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    # odd chunk lengths: R = 3 (65-96 cells) and R = 5 (129-160 cells)
    ((96, 12, 12), (1, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 80, 12), (1, 1, 1), G.WENO5Z, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 12, 150), (1, 1, 1), G.HCCS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((136, 12, 12), (2, 1, 1), G.CRWENO5, G.CD8, G.AUSMP, G.NAVIERSTOKES3D),
    ((12, 96, 12), (1, 1, 1), G.WENO7JS, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 12, 144), (1, 1, 1), G.HCCS77, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((72, 12, 12), (1, 1, 1), G.HCCS78, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((160, 12, 12), (1, 1, 1), G.WENO3JS, G.CD2, G.LF, G.EULER3D),
    # 193-224 cells (R = 8 with a tail) and R = 12 (289-384 cells; 384 without tail checks)
    ((224, 12, 12), (1, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((12, 200, 12), (1, 1, 1), G.WENO5Z, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((384, 12, 12), (1, 1, 1), G.WENO5JS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 12, 320), (1, 1, 1), G.WENO7JS, G.CD8, G.AUSM, G.NAVIERSTOKES3D),
    ((300, 12, 12), (3, 1, 1), G.HCCS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 352, 12), (1, 1, 1), G.HCCS77, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
]


@pytest.mark.parametrize("case", LONG_CASES, ids=lambda c: f"{c[0]}-{c[1]}-s{c[2]}-v{c[3]}-r{c[4]}-m{c[5]}")
def test_long_lines_ten_steps_match_oracle(case):
    n, npr, scheme, vis, rie, model = case
    print(f"long-line case {case}", flush=True)
    cfl = 0.5
    if rie == G.LF and n[2] > min(n[0], n[1]):
        # LF's dt omits the z wave speed (the z eigenvalue lands in the eta
        # slot, src/physicsModel.cpp:653-654), so a z line much finer than x/y
        # needs a smaller CFL number to stay stable
        cfl = 0.4 * min(n[0], n[1]) / n[2]
    gc, oc = _cfgs(n, npr, scheme, vis, rie, model, cfl=cfl)
    q = _perturbed_tgv(oc, 5)
    N = oc.nx * oc.ny * oc.nz
    s = O.Solver(oc)
    s.set_state(q)
    dt_ref, res_ref, t_ref = s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        dt_gpu, res_gpu, t_gpu = ctx.step(10)
        q_gpu = ctx.download()
    err = _rel_linf(q_gpu, q_ref, N)
    print("10-step rel Linf per variable", err, flush=True)
    # Thin boxes leave the short directions under-resolved (12 cells per
    # period), where the nonlinear weights sit far from their ideal values and
    # the compact rows lose diagonal dominance: the reference algorithm itself
    # then moves by up to ~2e-11 after 10 steps when its input changes by one
    # ulp (measured here with the oracle).  Every case is held to
    # max(1e-12, 3 x that measured floor); dt and the residual follow the state.
    tol = np.maximum(1e-12, 3.0 * _ulp_floor(oc, q, 10))
    print("max(1e-12, 1-ulp floor x3)", tol, flush=True)
    assert np.all(err <= tol), (err, tol)
    assert abs(dt_gpu - dt_ref) <= max(1e-13, tol.max()) * dt_ref
    assert abs(res_gpu - res_ref) <= max(1e-9, 100 * tol.max()) * abs(res_ref)


# The bar without any conditioning allowance, on the kernel variants the
# 256^3 workload runs (chunks of R = 8 interfaces per lane on 256-cell lines,
# the FULL variant without tail checks), on 384-cell lines (R = 12, FULL) and
# on 512-cell lines (R = 16): the
# workload's own field -- seeded random-phase decaying turbulence,
# E(k) ~ k^4 exp(-2 (k/4)^2), Mach_t 0.3, Re 200 -- on boxes whose short
# sides (32 cells; 16 for the 384- and 512-cell lines) still resolve the spectrum, so
# the nonlinear weights sit near their smooth values and the compact rows stay
# diagonally dominant.  CRWENO5 (src/crweno5.cpp:83-220) with LF and cd8 in
# each sweep direction, at CFL 0.1; WENO5-Z (src/weno5z.cpp) with AUSM on
# 512-cell lines at CFL 0.5.  At CFL 0.5 the CRWENO5 boxes with the long side
# in x or y amplify a 1-ulp change of the initial state to 4.7e-13 and 4.5e-12
# after 10 steps in the oracle itself (the reference algorithm's own floor,
# measured with _ulp_floor), so 1e-12 cannot be a bar there for any
# implementation short of the reference's exact arithmetic; at CFL 0.1 that
# floor is 2e-15 and the bar holds with three orders of margin.  The CFL 0.5
# long lines are in LONG_CASES, held to their measured floor.
STRICT_CASES = [
    ((256, 32, 32), G.CRWENO5, G.CD8, G.LF),
    ((32, 256, 32), G.CRWENO5, G.CD8, G.LF),
    ((32, 32, 256), G.CRWENO5, G.CD8, G.LF),
    ((512, 16, 16), G.WENO5Z, G.CD6, G.AUSM),
    ((16, 512, 16), G.WENO5Z, G.CD6, G.AUSM),
    ((384, 16, 16), G.CRWENO5, G.CD8, G.LF),     # R = 12 without tail checks
]


@pytest.mark.parametrize("case", STRICT_CASES, ids=lambda c: f"{c[0]}-s{c[1]}-v{c[2]}-r{c[3]}")
def test_strict_bar_long_lines_decaying_turbulence(case):
    n, scheme, vis, rie = case
    print(f"strict long-line case {case}", flush=True)
    cfl = 0.1 if scheme == G.CRWENO5 else 0.5
    if rie == G.LF and n[2] > min(n[0], n[1]):
        cfl = 0.4 * min(n[0], n[1]) / n[2]   # LF's dt omits the z wave speed (src/physicsModel.cpp:653-654)
    gc, oc = _cfgs(n, (1, 1, 1), scheme, vis, rie, G.NAVIERSTOKES3D, mach=0.3 / np.sqrt(3.0), reynolds=200.0,
                   cfl=cfl)
    q = G.init_hit(gc, k0=4.0, u_rms=1.0, seed=20261202)
    N = oc.nx * oc.ny * oc.nz
    oc.nthreads = 8
    s = O.Solver(oc)
    s.set_state(q)
    dt_ref, res_ref, _ = s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        dt_gpu, res_gpu, _ = ctx.step(10)
        q_gpu = ctx.download()
    err = _rel_linf(q_gpu, q_ref, N)
    print("10-step rel Linf per variable", err, flush=True)
    assert np.all(err <= 1e-12), err
    assert abs(dt_gpu - dt_ref) <= 1e-13 * dt_ref
    assert abs(res_gpu - res_ref) <= 1e-10 * abs(res_ref)


def test_rk3_stages_one_at_a_time():
    """nuc_rk3_stage: stage 0 gives u1 = u^n - dt RHS(u^n) with dt = cfl *
    dt_local (src/PDEData3d.cpp:70-91, src/TVD-RK3.cpp:59-64); three stages equal
    one nuc_step bit for bit; stages out of order and whole steps in the middle of
    a step are refused."""
    gc, oc = _cfgs((24, 20, 18), (1, 1, 1), G.CRWENO5, G.CD6, G.LF, G.NAVIERSTOKES3D)
    q = _perturbed_tgv(oc, 8)
    with G.Context(gc) as ctx:
        ctx.upload(q)
        ctx.step(1)
        ref = ctx.download()
        ref_info = ctx.sync()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        rhs, dt_local = ctx.rhs()
        with pytest.raises(G.NucError):
            ctx.rk3_stage(1)
        ctx.rk3_stage(0)
        u1 = ctx.download()
        expect = q - gc.cfl * dt_local * rhs
        assert np.abs(u1 - expect).max() <= 1e-14 * np.abs(expect).max()
        with pytest.raises(G.NucError):
            ctx.step(1)
        ctx.rk3_stage(1)
        ctx.rk3_stage(2)
        assert np.array_equal(ctx.download(), ref)
        assert ctx.sync() == ref_info
        ctx.step(1)   # a whole step after a completed one is fine


SOURCE_CASES = [
    # forcing, coriolis, centrifugal, scheme, riemann, model, rank grid
    (G.FORCING_ABC, False, False, G.WENO5JS, G.LF, G.NAVIERSTOKES3D, (1, 1, 1)),
    (G.FORCING_TG, False, False, G.CRWENO5, G.AUSM, G.NAVIERSTOKES3D, (2, 1, 1)),
    (G.FORCING_NONE, True, False, G.WENO5Z, G.LF, G.EULER3D, (1, 1, 1)),
    (G.FORCING_NONE, True, True, G.WENO5JS, G.AUSM, G.NAVIERSTOKES3D, (1, 2, 1)),
    (G.FORCING_CONST, True, True, G.CRWENO5, G.LF, G.NAVIERSTOKES3D, (1, 1, 1)),
]


@pytest.mark.parametrize("case", SOURCE_CASES, ids=lambda c: f"f{c[0]}-cor{int(c[1])}-cen{int(c[2])}-s{c[3]}")
def test_sources_rhs_and_ten_steps_match_oracle(case):
    """ABC/TG forcing and the Coriolis / centrifugal sources
    (src/physicsModel.cpp:1280-1319,1679-1913) against the oracle."""
    _source_case(case, (0.0, 0.0, 0.0))


@pytest.mark.parametrize("case", [SOURCE_CASES[0], SOURCE_CASES[1], SOURCE_CASES[3]],
                         ids=lambda c: f"f{c[0]}-cor{int(c[1])}-cen{int(c[2])}-s{c[3]}")
def test_sources_with_block_origin_match_oracle(case):
    """A nonzero block.in origin is subtracted from every cell centre
    (src/block.cpp:116-118), which shifts the ABC/TG fields and the centrifugal
    well: device and oracle must agree with it."""
    _source_case(case, (0.7, -1.3, 2.1))


def _source_case(case, origin):
    forcing, cor, cen, scheme, rie, model, npr = case
    print(f"source case {case} origin {origin}", flush=True)
    kw = dict(n=(24, 20, 16), npr=npr, scheme=scheme, vis=G.CD6, riemann=rie, model=model, mach=0.2,
              reynolds=800.0, forcing_amp=0.05, forcing_k=1.0 if forcing != G.FORCING_CONST else 2.0,
              forcing_amp_hel=0.0, coriolis=cor, centrifugal=cen, rossby=0.6, x_center=3.0, y_center=3.3,
              well_rad=1.4, origin=origin)
    gc = G.make_config(forcing=forcing, **kw)
    oc = O.make_config(forcing=forcing, **kw)
    q = _perturbed_tgv(oc, 7)
    N = oc.nx * oc.ny * oc.nz
    s = O.Solver(oc)
    s.set_state(q)
    r_ref, _ = s.rhs()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        r_gpu, _ = ctx.rhs()
    err = _rel_linf(r_gpu, r_ref, N)
    print("rhs rel err", err, flush=True)
    assert np.all(err <= 1e-11), err
    s = O.Solver(oc)
    s.set_state(q)
    s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        ctx.step(10)
        q_gpu = ctx.download()
    err = _rel_linf(q_gpu, q_ref, N)
    print("10-step rel Linf", err, flush=True)
    assert np.all(err <= 1e-12), err


# Lines longer than 512 cells: swept in segments (one wave per segment, the
# neighbouring segments' cells as halo, the segment's n+1 interfaces computed
# on both sides of each cut).  Explicit schemes cut at a divisor of n in
# [64, 255], with rank blocks (closure and inner zones) falling inside or
# across the cuts; the compact schemes cut at rank-block edges (segments of
# 128, 256 or 2 x 100 cells; R = 8 and R = 16).  Same policy as LONG_CASES:
# 10 steps against the oracle, max(1e-12, 3 x the measured 1-ulp floor).
SEG_CASES = [
    # n, rank grid, scheme, vis, riemann, model
    ((1024, 10, 10), (1, 1, 1), G.WENO5JS, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((10, 768, 10), (1, 1, 1), G.WENO5Z, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((12, 12, 1536), (1, 1, 1), G.WENO7JS, G.CD8, G.AUSM, G.NAVIERSTOKES3D),
    ((600, 10, 10), (1, 1, 1), G.WENO3JS, G.CD2, G.AUSMP, G.NAVIERSTOKES3D),
    ((10, 1040, 10), (1, 1, 1), G.UPWIND1ST, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((1024, 10, 10), (2, 1, 1), G.HCCS78, G.CD8, G.AUSM, G.NAVIERSTOKES3D),
    ((1000, 10, 10), (5, 1, 1), G.WENO5Z, G.CD6, G.LF, G.NAVIERSTOKES3D),
    ((10, 10, 2048), (1, 1, 1), G.WENO5JS, G.CD6, G.AUSM, G.EULER3D),
    ((1024, 10, 10), (4, 1, 1), G.CRWENO5, G.CD8, G.LF, G.NAVIERSTOKES3D),
    ((10, 1536, 10), (1, 12, 1), G.CRWENO5, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((10, 10, 1200), (1, 1, 12), G.CRWENO5, G.CD8, G.AUSM, G.NAVIERSTOKES3D),
    ((10, 10, 1024), (1, 1, 8), G.HCCS, G.CD6, G.AUSM, G.NAVIERSTOKES3D),
    ((768, 10, 10), (3, 1, 1), G.HCCS77, G.CD8, G.LF, G.NAVIERSTOKES3D),
]


@pytest.mark.parametrize("case", SEG_CASES, ids=lambda c: f"{c[0]}-{c[1]}-s{c[2]}-v{c[3]}-r{c[4]}-m{c[5]}")
def test_segmented_long_lines_ten_steps_match_oracle(case):
    n, npr, scheme, vis, rie, model = case
    print(f"segmented long-line case {case}", flush=True)
    cfl = 0.5
    if rie == G.LF and n[2] > min(n[0], n[1]):
        cfl = 0.4 * min(n[0], n[1]) / n[2]   # LF's dt omits the z wave speed (src/physicsModel.cpp:653-654)
    L = tuple(2 * np.pi * m / max(n) for m in n)   # cubic cells
    kw = dict(n=n, npr=npr, L=L, scheme=scheme, vis=vis, riemann=rie, model=model, mach=0.1, reynolds=1600.0,
              cfl=cfl)
    gc, oc = G.make_config(**kw), O.make_config(**kw)
    oc.nthreads = 8
    q = _perturbed_tgv(oc, 5)
    N = oc.nx * oc.ny * oc.nz
    s = O.Solver(oc)
    s.set_state(q)
    dt_ref, res_ref, t_ref = s.step(10)
    q_ref = s.get_state()
    with G.Context(gc) as ctx:
        ctx.upload(q)
        dt_gpu, res_gpu, t_gpu = ctx.step(10)
        q_gpu = ctx.download()
    err = _rel_linf(q_gpu, q_ref, N)
    print("10-step rel Linf per variable", err, flush=True)
    tol = np.maximum(1e-12, 3.0 * _ulp_floor(oc, q, 10))
    print("max(1e-12, 1-ulp floor x3)", tol, flush=True)
    assert np.all(err <= tol), (err, tol)
    assert abs(dt_gpu - dt_ref) <= max(1e-13, tol.max()) * dt_ref
    assert abs(res_gpu - res_ref) <= max(1e-9, 100 * tol.max()) * abs(res_ref)


@pytest.mark.parametrize("prec", [G.MIXED, G.FP32], ids=["mixed", "fp32"])
def test_segmented_long_lines_reduced_precision(prec):
    """The segmented sweeps in the mixed and FP32 paths: within 2e-5 of FP64
    after 5 steps (relative L-infinity per variable, momentum scaled by its
    largest component)."""
    for n, npr, scheme, rie in (((1024, 12, 12), (1, 1, 1), G.WENO5Z, G.AUSM),
                                ((12, 1024, 12), (1, 4, 1), G.CRWENO5, G.LF)):
        print(f"segmented reduced precision {prec} {n} {scheme}", flush=True)
        L = tuple(2 * np.pi * m / max(n) for m in n)
        kw = dict(n=n, npr=npr, L=L, scheme=scheme, vis=G.CD8, riemann=rie, model=G.NAVIERSTOKES3D,
                  viscosity=G.SUTHERLAND, mach=0.2, reynolds=400.0, cfl=0.5)
        c64, cr = G.make_config(precision=G.FP64, **kw), G.make_config(precision=prec, **kw)
        q = G.init_hit(c64, k0=3.0, u_rms=0.5, seed=23)
        outs = []
        for c in (c64, cr):
            with G.Context(c) as ctx:
                ctx.upload(q)
                ctx.step(5)
                outs.append(ctx.download())
        Nc = n[0] * n[1] * n[2]
        a, b = outs[0].reshape(5, Nc), outs[1].reshape(5, Nc)
        mom = np.abs(a[1:4]).max()
        scale = np.array([np.abs(a[0]).max(), mom, mom, mom, np.abs(a[4]).max()])
        err = np.abs(a - b).max(axis=1) / scale
        print("vs FP64 rel Linf", err, flush=True)
        assert np.all(err <= 2e-5), err
