"""Shell spectra on the device (nuc_spectra_kinetic, nuc_spectrum_scalar)
against the numpy restatement of postproc::solveFFT_* / fft::solveShell
(oracle/spectra.py, itself pinned by tests/test_oracle_spectra.py).  The
vorticity of the enstrophy and helicity spectra comes from the oracle's
gradients (the viscous central difference of solveGrad)."""
import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O
import spectra as S

pytestmark = pytest.mark.gpu


def _state(n, seed):
    kw = dict(n=n, scheme=G.WENO5JS, vis=G.CD6, riemann=G.LF, model=G.NAVIERSTOKES3D, mach=0.3, reynolds=500.0)
    gc, oc = G.make_config(**kw), O.make_config(**kw)
    q = G.init_hit(gc, k0=3.0, u_rms=0.5, seed=seed)
    s = O.Solver(oc)
    s.set_state(q)
    s.step(2)                      # density and pressure fluctuations
    return gc, oc, s, s.get_state()


@pytest.mark.parametrize("n", [(32, 32, 32), (32, 24, 16)], ids=str)
def test_spectra_match_restatement(n):
    print(f"spectra {n}", flush=True)
    gc, oc, solver, q = _state(n, 4)
    nx, ny, nz = n
    Q = q.reshape(5, nz, ny, nx)
    rho = Q[0]
    u, v, w = Q[1] / rho, Q[2] / rho, Q[3] / rho
    p = np.abs((Q[4] - 0.5 * rho * (u * u + v * v + w * w)) * (gc.gamma - 1.0))
    P = O.Post(solver, 1)
    P.update()
    om = [P.field(1, 15 + c).reshape(nz, ny, nx) for c in range(3)]
    ens = (om[0] ** 2 + om[1] ** 2 + om[2] ** 2) * 0.5
    hel = om[0] * u + om[1] * v + om[2] * w
    ref_k = S.kinetic(u, v, w)
    ref_s = [S.scalar(f) for f in (rho, p, ens, hel)]
    with G.Context(gc) as ctx:
        ctx.upload(q)
        got_k = ctx.spectra_kinetic()
        got_s = [ctx.spectrum_scalar(f) for f in range(4)]
    for sidx in range(5):
        err = np.abs(got_k[sidx] - ref_k[sidx]).max() / np.abs(ref_k[0]).max()
        print(f"kinetic spectrum {sidx}: rel err {err:.2e}", flush=True)
        assert err <= 1e-12, sidx
    for f in range(4):
        err = np.abs(got_s[f] - ref_s[f]).max() / np.abs(ref_s[f]).max()
        print(f"scalar spectrum {f}: rel err {err:.2e}", flush=True)
        assert err <= 1e-12, f


def test_kinetic_spectrum_of_the_initial_field_is_its_energy():
    """Parseval on the device: a band-limited initial field (modes below kmax)
    has total shell energy 0.5 <|u|^2> = 1.5 u_rms^2, all of it solenoidal."""
    cfg = G.make_config(n=(64, 64, 64), mach=0.2)
    with G.Context(cfg) as ctx:
        ctx.init_hit_device(k0=2.0, u_rms=0.4, seed=3)
        sp = ctx.spectra_kinetic()
    assert abs(sp[0].sum() - 1.5 * 0.4 ** 2) <= 1e-12
    assert sp[1].sum() <= 1e-25 and abs(sp[2].sum() - sp[0].sum()) <= 1e-13


def test_q_fields_against_numpy_and_stats():
    """solveQ's per-point fields (src/postproc.cpp:736-835) on the device:
    gradients against numpy's periodic 6th-order differences, the derived
    fields against their definitions, and their volume means against nuc_stats
    (the global averages of the same fields, computed by its own kernel)."""
    n = 32
    cfg = G.make_config(n=(n, n, n), scheme=G.WENO5JS, vis=G.CD6, mach=0.3)
    q = G.init_hit(cfg, k0=3.0, u_rms=0.6, seed=21)
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        ctx.step(2)                                        # non-uniform density
        state = ctx.download().reshape(5, n, n, n)
        f = ctx.q_fields()
        st = ctx.stats()
    rho = state[0]
    vel = state[1:4] / rho
    h = 2 * np.pi / n

    def d(a, axis):   # cd6, src/centraldifference6th.cpp:42-48
        r = lambda m: np.roll(a, -m, axis=axis) - np.roll(a, m, axis=axis)
        return (0.75 * r(1) - 0.15 * r(2) + r(3) / 60.0) / h

    grads = [d(vel[c], 2 - e) for c in range(3) for e in range(3)]   # (u,v,w) x (x,y,z)
    scale = max(np.abs(g).max() for g in grads)
    for k in range(9):
        assert np.abs(f[k] - grads[k].ravel()).max() <= 1e-12 * scale, k
    ux, uy, uz, vx, vy, vz, wx, wy, wz = [g.ravel() for g in grads]
    u, v, w = (c.ravel() for c in vel)
    om = [wy - vz, uz - wx, vx - uy]
    expect = om + [0.5 * (om[0] ** 2 + om[1] ** 2 + om[2] ** 2),
                   ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy,
                   om[0] * u + om[1] * v + om[2] * w,
                   0.5 * rho.ravel() * (u * u + v * v + w * w)]
    for k in range(7):
        e = expect[k]
        assert np.abs(f[9 + k] - e).max() <= 1e-11 * np.abs(e).max(), 9 + k
    means = f.mean(axis=1)
    for k in range(16):
        assert abs(means[k] - st[k]) <= 1e-12 * max(np.abs(f[k]).max(), 1e-300), (k, means[k], st[k])

