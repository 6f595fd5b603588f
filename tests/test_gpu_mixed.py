"""The mixed path (NUC_MIXED): split fluxes, reconstruction stencils and the
viscous stencils in FP32; nonlinear weights, the CRWENO5/hccs tridiagonal
solves, interface flux differences, RHS accumulation, the TVD-RK3 update and
every global reduction in FP64.  Its error is stated against the FP64 path,
which is itself held to the oracle at 1e-12 (test_gpu_rhs.py).

Bar: <= 1e-5 relative error in kinetic energy and in dissipation after 20 steps
of decaying turbulence; the other statistics are printed with their errors.
"""
import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu

NAMES = ["ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz", "om_x", "om_y", "om_z", "enstrophy", "Q",
         "helicity", "KE", "u_rms", "M_t", "theta_rms", "mu_mean", "Re_lambda", "lambda", "rho_mean", "tau_t",
         "skewness", "eps_s", "eps_d"]


def _hit_cfg(n, precision, scheme=G.CRWENO5, vis=G.CD8, riemann=G.LF, **kw):
    return G.make_config(n=(n, n, n), scheme=scheme, vis=vis, riemann=riemann, model=G.NAVIERSTOKES3D,
                         viscosity=G.SUTHERLAND, mach=0.3 / np.sqrt(3.0), reynolds=200.0, prandtl=0.72,
                         cfl=0.5, precision=precision, **kw)


def _run(cfg, q, steps):
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        ctx.step(steps)
        return ctx.download(), ctx.stats_ex()


def test_dissipation_statistics_match_closed_form():
    """eps_s = <mu |omega|^2>/Re and eps_d = 4/3 <mu (div u)^2>/Re evaluated
    with numpy (cd6 on the periodic box) from the same state."""
    n = 32
    cfg = _hit_cfg(n, G.FP64, vis=G.CD6)
    q = G.init_hit(cfg, k0=4.0, u_rms=1.0, seed=11)
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        st = ctx.stats_ex()
    Q = q.reshape(5, n, n, n)
    rho = Q[0]
    u = [Q[1 + c] / rho for c in range(3)]
    gam, M = 1.4, 0.3 / np.sqrt(3.0)
    p = np.abs((Q[4] - 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) * (gam - 1))
    T = gam * M * M * p / rho
    h = 2 * np.pi / n

    def d(f, axis):  # cd6, arrays stored (z, y, x)
        ax = {0: 2, 1: 1, 2: 0}[axis]
        r = lambda m: np.roll(f, -m, axis=ax) - np.roll(f, m, axis=ax)
        return (0.75 * r(1) - 0.15 * r(2) + (1 / 60) * r(3)) / h

    g = [[d(u[c], a) for a in range(3)] for c in range(3)]
    om = [g[2][1] - g[1][2], g[0][2] - g[2][0], g[1][0] - g[0][1]]
    S = 110.4 / 288.0
    mu = (1 + S) / (T + S) * T ** 1.5
    eps_s = np.mean(mu * (om[0] ** 2 + om[1] ** 2 + om[2] ** 2)) / 200.0
    div = g[0][0] + g[1][1] + g[2][2]
    eps_d = 4.0 / 3.0 * np.mean(mu * div * div) / 200.0
    print("eps_s", st[25], eps_s, "eps_d", st[26], eps_d, flush=True)
    assert abs(st[25] - eps_s) <= 1e-11 * eps_s
    assert abs(st[26] - eps_d) <= 1e-9 * eps_d + 1e-14 * eps_s


def test_mixed_decaying_turbulence_statistics_within_1e5():
    n, steps = 48, 20
    c64 = _hit_cfg(n, G.FP64)
    cmx = _hit_cfg(n, G.MIXED)
    q = G.init_hit(c64, k0=4.0, u_rms=1.0, seed=20261202)
    q64, s64 = _run(c64, q, steps)
    qmx, smx = _run(cmx, q, steps)
    rel = np.abs(smx - s64) / np.maximum(np.abs(s64), 1e-300)
    for i, nm in enumerate(NAMES):
        print(f"  {nm:10s} fp64 {s64[i]: .12e}  mixed {smx[i]: .12e}  rel {rel[i]:.2e}", flush=True)
    N = n ** 3
    a, b = qmx.reshape(5, N), q64.reshape(5, N)
    state = np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)
    print("state rel Linf per variable", state, flush=True)
    eps64, epsmx = s64[25] + s64[26], smx[25] + smx[26]
    assert rel[15] <= 1e-5, rel[15]                              # kinetic energy
    assert abs(epsmx - eps64) <= 1e-5 * eps64, (epsmx, eps64)    # dissipation
    assert rel[25] <= 1e-5 and rel[12] <= 1e-5 and rel[16] <= 1e-5


MIXED_CASES = [
    (G.WENO5JS, G.CD6, G.LF, G.FORCING_NONE),
    (G.WENO5Z, G.CD6, G.AUSM, G.FORCING_NONE),
    (G.WENO7JS, G.CD8, G.LF, G.FORCING_NONE),
    (G.HCCS, G.CD6, G.AUSM, G.FORCING_NONE),
    (G.CRWENO5, G.CD8, G.LF, G.FORCING_CONST),
    (G.WENO3JS, G.CD2, G.AUSMP, G.FORCING_NONE),
    (G.UPWIND1ST, G.CD6, G.LF, G.FORCING_NONE),
    (G.HCCS78, G.CD8, G.AUSMP, G.FORCING_NONE),
    (G.HCCS77, G.CD6, G.LF, G.FORCING_ABC),
]


@pytest.mark.parametrize("case", MIXED_CASES, ids=lambda c: f"s{c[0]}-v{c[1]}-r{c[2]}-f{c[3]}")
def test_mixed_close_to_fp64_every_scheme(case):
    scheme, vis, rie, forcing = case
    n = 24
    kw = dict(scheme=scheme, vis=vis, riemann=rie, forcing=forcing, forcing_k=2.0, forcing_amp=0.2,
              forcing_amp_hel=0.1)
    c64, cmx = _hit_cfg(n, G.FP64, **kw), _hit_cfg(n, G.MIXED, **kw)
    q = G.init_hit(c64, k0=3.0, u_rms=1.0, seed=7)
    q64, s64 = _run(c64, q, 5)
    qmx, smx = _run(cmx, q, 5)
    N = n ** 3
    a, b = qmx.reshape(5, N), q64.reshape(5, N)
    mom = np.abs(b[1:4]).max()
    scale = np.array([np.abs(b[0]).max(), mom, mom, mom, np.abs(b[4]).max()])
    err = np.abs(a - b).max(axis=1) / scale
    print(case, "state rel Linf", err, "KE rel", abs(smx[15] - s64[15]) / s64[15], flush=True)
    assert np.all(err <= 1e-5), err
    assert abs(smx[15] - s64[15]) <= 1e-6 * s64[15]


@pytest.mark.parametrize("n,scheme,riemann", [((256, 12, 12), G.CRWENO5, G.LF), ((12, 256, 12), G.WENO5JS, G.LF),
                                              ((12, 12, 512), G.WENO5Z, G.AUSM), ((192, 12, 12), G.CRWENO5, G.AUSM),
                                              ((160, 12, 12), G.CRWENO5, G.LF), ((12, 96, 12), G.WENO5Z, G.AUSM),
                                              ((384, 12, 12), G.CRWENO5, G.LF), ((12, 224, 12), G.WENO5JS, G.LF)],
                         ids=lambda v: str(v))
def test_mixed_long_lines_state_error(n, scheme, riemann):
    """The FP32 line buffers and stencils at every long chunk length: the mixed
    state stays within 1e-5 (relative L-infinity per variable) of the FP64 path
    after 5 steps.  Cells are cubic (box lengths proportional to n), so the
    seeded field is resolved in every direction; on under-resolved data the
    compact rows lose diagonal dominance and any rounding, FP64 included, is
    amplified (see the long-line oracle cases)."""
    print(f"mixed long line {n} {scheme} {riemann}", flush=True)
    L = tuple(2 * np.pi * m / max(n) for m in n)
    kw = dict(n=n, L=L, scheme=scheme, vis=G.CD8, riemann=riemann, model=G.NAVIERSTOKES3D,
              viscosity=G.SUTHERLAND, mach=0.2, reynolds=400.0, cfl=0.5)
    c64 = G.make_config(precision=G.FP64, **kw)
    c32 = G.make_config(precision=G.MIXED, **kw)
    q = G.init_hit(c64, k0=3.0, u_rms=0.5, seed=21)
    a, _ = _run(c64, q, 5)
    b, _ = _run(c32, q, 5)
    N = n[0] * n[1] * n[2]
    a = a.reshape(5, N)
    b = b.reshape(5, N)
    mom = np.abs(a[1:4]).max()
    scale = np.array([np.abs(a[0]).max(), mom, mom, mom, np.abs(a[4]).max()])
    err = np.abs(a - b).max(axis=1) / scale
    print("mixed vs FP64 rel Linf", err, flush=True)
    assert np.all(err <= 1e-5), err


def test_mixed_hccs_only_context_is_fp64():
    """hccs and hccs77 keep FP64 stencils in the mixed path; a context whose
    directions all use them keeps its viscous fluxes in FP64 too, so its steps
    equal the FP64 context's bit for bit.  With one FP32-stencil direction the
    viscous fluxes are FP32 again (and the result differs at rounding level)."""
    print("mixed hccs-only context", flush=True)
    n = 24
    for scheme in (G.HCCS, (G.HCCS77, G.HCCS, G.HCCS77)):
        kw = dict(scheme=scheme, vis=G.CD6, riemann=G.AUSM)
        c64, cmx = _hit_cfg(n, G.FP64, **kw), _hit_cfg(n, G.MIXED, **kw)
        q = G.init_hit(c64, k0=3.0, u_rms=1.0, seed=17)
        a, _ = _run(c64, q, 3)
        b, _ = _run(cmx, q, 3)
        assert np.array_equal(a, b), np.abs(a - b).max()
    assert G.footprint(_hit_cfg(n, G.MIXED, scheme=G.HCCS)) == G.footprint(_hit_cfg(n, G.FP64, scheme=G.HCCS))
    kw = dict(scheme=(G.HCCS, G.CRWENO5, G.HCCS), vis=G.CD6, riemann=G.AUSM)
    c64, cmx = _hit_cfg(n, G.FP64, **kw), _hit_cfg(n, G.MIXED, **kw)
    q = G.init_hit(c64, k0=3.0, u_rms=1.0, seed=17)
    a, _ = _run(c64, q, 3)
    b, _ = _run(cmx, q, 3)
    N = n ** 3
    err = np.abs(a - b).reshape(5, N).max(axis=1) / np.abs(a).reshape(5, N).max(axis=1)
    print("mixed hccs/crweno5/hccs vs FP64", err, flush=True)
    assert not np.array_equal(a, b) and np.all(err <= 1e-5), err
