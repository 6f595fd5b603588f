"""Pins of the oracle's Const forcing (AddForcing_Const, src/physicsModel.cpp:1368-1677).

FFTW is not vendored with the reference and its tests hold no vectors for this
path, so the forcing is pinned by what the algorithm fixes, with numpy.fft as an
independent DFT:

- the forcing term enters the RHS as -f in the momentum equations and as
  -(f.u - <f.u>) in the energy equation, never in the mass equation
  (src/physicsModel.cpp:1605-1667), so mass, momentum (band k != 0) and total
  energy source all sum to zero;
- f is real, lives only on the shell modes |k - k_f| < 0.5 and is solenoidal;
- with the helical amplitude sigma = 0, adding f to u rescales the solenoidal
  band content so that its energy becomes |A - E_c| (src/physicsModel.cpp:1559),
  where the band sums count k = 0 as |k| = 1 (src/physicsModel.cpp:1483-1484);
- with sigma != 0, f-hat is a real multiple (theta - 1) of s + sigma (i k x s)/|k|.
"""
import numpy as np
import pytest

import oracle as O

N1 = 12


def _state(n=N1, seed=5, mean_u=0.0):
    rng = np.random.default_rng(seed)
    oc = O.make_config(n=(n, n, n), scheme=O.WENO5JS, vis=O.CD6, riemann=O.LF, model=O.NS)
    q = O.init_tgv(oc).reshape(5, n, n, n)
    x = 2 * np.pi * (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    # add low-mode content in every component, incl. compressive parts
    rho = q[0]
    u = q[1] / rho + 0.3 * np.sin(2 * Z.T + 0.4) + 0.2 * np.cos(X.T + Y.T) + mean_u
    v = q[2] / rho + 0.25 * np.cos(2 * X.T - 0.3) + 0.1 * rng.standard_normal(rho.shape) * 0.01
    w = 0.2 * np.sin(Y.T + 2 * X.T) + 0.15 * np.cos(Z.T)
    E = q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / rho
    q[1], q[2], q[3] = rho * u, rho * v, rho * w
    q[4] = E + 0.5 * rho * (u * u + v * v + w * w)
    return q.reshape(-1).copy()


def _rhs(q, n, forcing, k=2.0, amp=0.05, hel=0.0):
    oc = O.make_config(n=(n, n, n), scheme=O.WENO5JS, vis=O.CD6, riemann=O.LF, model=O.NS,
                       forcing=O.FORCE_CONST if forcing else O.FORCE_NONE, forcing_k=k, forcing_amp=amp,
                       forcing_amp_hel=hel)
    s = O.Solver(oc)
    s.set_state(q)
    return s.rhs()[0].reshape(5, n, n, n)


def _forcing(q, n, **kw):
    d = _rhs(q, n, True, **kw) - _rhs(q, n, False, **kw)
    return d


def _prims(q, n):
    q = q.reshape(5, n, n, n)
    return q[1] / q[0], q[2] / q[0], q[3] / q[0]


def _kgrid(n):
    k = np.fft.fftfreq(n, 1.0 / n)
    # arrays are stored (z, y, x): axis 0 is z
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    return KX, KY, KZ


def test_forcing_enters_momentum_and_energy_only_and_conserves():
    q = _state()
    d = _forcing(q, N1)
    assert np.all(d[0] == 0.0)
    f = -d[1:4]
    assert np.abs(f).max() > 1e-6
    for c in range(3):
        assert abs(f[c].sum()) <= 1e-12 * np.abs(f[c]).sum()
    assert abs(d[4].sum()) <= 1e-12 * np.abs(d[4]).sum()
    # energy term is -(f.u - <f.u>)
    u, v, w = _prims(q, N1)
    fu = f[0] * u + f[1] * v + f[2] * w
    np.testing.assert_allclose(d[4], -(fu - fu.mean()), rtol=0, atol=1e-13 * np.abs(fu).max())


def test_forcing_lives_on_shell_and_is_solenoidal():
    q = _state()
    kf = 2.0
    f = -_forcing(q, N1, k=kf)[1:4]
    F = [np.fft.fftn(f[c]) / f[c].size for c in range(3)]
    KX, KY, KZ = _kgrid(N1)
    km = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    shell = np.abs(km - kf) < 0.5
    big = max(np.abs(Fc).max() for Fc in F)
    for Fc in F:
        assert np.abs(Fc[~shell]).max() <= 1e-13 * big
    div = KX * F[0] + KY * F[1] + KZ * F[2]
    assert np.abs(div).max() <= 1e-12 * big * kf


@pytest.mark.parametrize("kf,mean_u", [(2.0, 0.0), (1.0, 0.15)])
def test_band_energy_rescaled_to_target(kf, mean_u):
    q = _state(mean_u=mean_u)
    amp = 0.05
    f = -_forcing(q, N1, k=kf, amp=amp, hel=0.0)[1:4]
    u = _prims(q, N1)
    U = [np.fft.fftn(c) / c.size for c in u]          # forward transform normalised by 1/N
    Fh = [np.fft.fftn(c) / c.size for c in f]         # backward transform is unnormalised
    KX, KY, KZ = _kgrid(N1)
    km = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    kms = np.where(km < 1e-6, 1.0, km)
    band = np.abs(kms - kf) < 0.5
    forced = np.abs(km - kf) < 0.5
    ku = (KX * U[0] + KY * U[1] + KZ * U[2]) / kms
    S = [U[0] - ku * KX / kms, U[1] - ku * KY / kms, U[2] - ku * KZ / kms]
    C = [U[c] - S[c] for c in range(3)]
    e_s = 0.5 * sum(np.abs(S[c]) ** 2 for c in range(3))
    e_c = 0.5 * sum(np.abs(C[c]) ** 2 for c in range(3))
    Es, Ec = e_s[band].sum(), e_c[band].sum()
    theta2 = abs(amp - Ec) / Es
    after = 0.5 * sum(np.abs(S[c] + Fh[c]) ** 2 for c in range(3))
    np.testing.assert_allclose(after[forced].sum(), theta2 * e_s[forced].sum(), rtol=1e-11)
    if kf == 1.0:
        # k = 0 is in the band sums but is never forced
        assert band[0, 0, 0] and not forced[0, 0, 0]
        assert e_s[0, 0, 0] > 1e-4


def test_helical_forcing_is_real_multiple_of_mixed_solenoidal_part():
    q = _state()
    kf, sigma = 2.0, 0.3
    f = -_forcing(q, N1, k=kf, amp=0.05, hel=sigma)[1:4]
    u = _prims(q, N1)
    U = [np.fft.fftn(c) / c.size for c in u]
    Fh = [np.fft.fftn(c) / c.size for c in f]
    KX, KY, KZ = _kgrid(N1)
    km = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    forced = np.abs(km - kf) < 0.5
    kms = np.where(km < 1e-6, 1.0, km)
    ku = (KX * U[0] + KY * U[1] + KZ * U[2]) / kms
    S = np.array([U[0] - ku * KX / kms, U[1] - ku * KY / kms, U[2] - ku * KZ / kms])
    K = np.array([KX, KY, KZ])
    V = S + sigma * 1j * np.cross(K, S, axis=0) / kms
    Fh = np.array(Fh)
    vn = np.sqrt(np.sum(np.abs(V) ** 2, axis=0))
    live = forced & (vn > 1e-8 * vn[forced].max())   # modes the field actually populates
    assert live.sum() >= 6
    ratio = (np.sum(np.conj(V) * Fh, axis=0) / np.maximum(vn, 1e-300) ** 2)[live]
    assert np.abs(ratio.imag).max() <= 1e-10
    assert np.ptp(ratio.real) <= 1e-10 * abs(ratio.real).max()
    resid = Fh[:, forced] - ratio.real.mean() * V[:, forced]
    assert np.abs(resid).max() <= 1e-12 * np.abs(Fh).max()
