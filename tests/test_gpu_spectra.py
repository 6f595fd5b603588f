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
