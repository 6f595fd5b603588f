"""Pins of the spectra restatement (oracle/spectra.py; postproc::solveFFT_*,
src/postproc.cpp:931-1184, fft::solveShell, src/fft.cpp:384-572): the DFT
definition by brute force, Parseval, the Helmholtz split of solenoidal and
gradient fields, the helical split of a single helical wave, and the shell
rules (k = 0 counted in shell 1 of the kinetic spectra; the scalar spectra
count non-negative indices only)."""
import numpy as np

import spectra as S


def _grid(n):
    x = 2 * np.pi * np.arange(n) / n
    return np.meshgrid(x, x, x, indexing="ij")       # z, y, x


def test_forward_is_the_normalised_dft():
    rng = np.random.default_rng(0)
    f = rng.standard_normal((4, 3, 5))
    nz, ny, nx = f.shape
    ref = np.zeros(f.shape, dtype=complex)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx):
                for z in range(nz):
                    for y in range(ny):
                        for x in range(nx):
                            ph = kx * x / nx + ky * y / ny + kz * z / nz
                            ref[kz, ky, kx] += f[z, y, x] * np.exp(-2j * np.pi * ph)
    assert np.allclose(S.forward(f), ref / f.size, atol=1e-14)


def _low_mode_field(n, seed, solenoidal=None):
    rng = np.random.default_rng(seed)
    z, y, x = _grid(n)
    u = np.zeros((3, n, n, n))
    for _ in range(8):
        k = rng.integers(-3, 4, size=3)
        if not k.any():
            continue
        a = rng.standard_normal(3)
        if solenoidal is True:
            a -= a.dot(k) / k.dot(k) * k
        elif solenoidal is False:
            a = k * rng.standard_normal()            # gradient of a potential: a || k
        ph = k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 2 * np.pi)
        u += a[:, None, None, None] * np.cos(ph)
    return u


def test_parseval_and_helmholtz_split():
    n = 16
    u = _low_mode_field(n, 1)
    sp = S.kinetic(*u)
    assert abs(sp[0].sum() - 0.5 * (u ** 2).sum(0).mean()) <= 1e-13 * sp[0].sum()
    assert np.allclose(sp[1] + sp[2], sp[0], rtol=1e-12, atol=1e-16)   # |c|^2 + |s|^2 = |u|^2 (c perp s)
    assert np.allclose(sp[3] + sp[4], sp[2], rtol=1e-12, atol=1e-16)   # |l|^2 + |r|^2 = |s|^2
    us = _low_mode_field(n, 2, solenoidal=True)
    sp = S.kinetic(*us)
    assert sp[1].max() <= 1e-28 and np.allclose(sp[2], sp[0], rtol=1e-12)
    ug = _low_mode_field(n, 3, solenoidal=False)
    sp = S.kinetic(*ug)
    assert sp[2].max() <= 1e-28 * sp[0].max() + 1e-30 and np.allclose(sp[1], sp[0], rtol=1e-12)


def test_helical_wave_is_all_left_or_all_right():
    n = 16
    z, y, x = _grid(n)
    # u = (cos kz, +-sin kz, 0): circularly polarised waves along z, k = 2, of
    # opposite handedness; each is all in one helical component (with
    # r = (s + i k x s/|k|)/2, the + wave is the left one), the two opposite
    where = []
    for sign in (1.0, -1.0):
        u = np.stack([np.cos(2 * z), sign * np.sin(2 * z), 0 * z])
        sp = S.kinetic(*u)
        assert abs(sp[0][2] - 0.5) <= 1e-13
        left, right = sp[3][2], sp[4][2]
        assert min(left, right) <= 1e-28 and abs(max(left, right) - 0.5) <= 1e-13
        where.append("left" if left > right else "right")
    assert where == ["left", "right"]


def test_shell_rules():
    n = 16
    z, y, x = _grid(n)
    # uniform flow: k = 0 counted as |k| = 1 in the kinetic spectra
    u = np.stack([np.full((n, n, n), 2.0), 0 * x, 0 * x])
    sp = S.kinetic(*u)
    assert abs(sp[0][1] - 2.0) <= 1e-13 and abs(sp[0].sum() - 2.0) <= 1e-13
    # scalar spectra: cos(3x) has modes at +3 and -3; only the non-negative index is counted
    f = np.cos(3 * x)
    s = S.scalar(f)
    assert abs(s[3] - 0.25) <= 1e-14 and abs(s.sum() - 0.25) <= 1e-14
    f = 1.5 + 0 * x
    assert abs(S.scalar(f)[0] - 2.25) <= 1e-14
