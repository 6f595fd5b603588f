"""Shell spectra of postproc::solveFFT_* (src/postproc.cpp:931-1184) and
fft::solveShell (src/fft.cpp:384-572), restated with numpy.

TEST INFRASTRUCTURE ONLY: the checker for the device spectra; the product
never imports it.  numpy.fft stands in for FFTW (absent from the reference
tree, version unpinned in src/makefile:36-38): forward transform e^{-i},
normalised by 1/N (src/fft.cpp:182-300,626-641).
"""
import numpy as np


def forward(f):
    """solveFFT_forward: arrays (nz, ny, nx), result / N."""
    return np.fft.fftn(f) / f.size


def _signed(n):
    # kx0 = i >= n/2 ? i - n : i   (src/postproc.cpp:1015-1017)
    i = np.arange(n, dtype=np.float64)
    return np.where(i >= n / 2.0, i - n, i)


def _shell_index_signed(shape):
    nz, ny, nx = shape
    kz, ky, kx = np.meshgrid(_signed(nz), _signed(ny), _signed(nx), indexing="ij")
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz)
    kmag = np.where(kmag < 1e-6, 1.0, kmag)
    return kx, ky, kz, kmag


def _bin(values, kmag, kmax):
    """solveShell's rounding (src/fft.cpp:430-435): shell = floor(k) or floor(k)+1,
    for kmag < kmax; the output holds shells 0 .. kmax-1."""
    sel = kmag < kmax
    fl = np.floor(kmag[sel])
    ki = np.where((kmag[sel] - fl) - 0.5 <= 0, fl, fl + 1).astype(np.int64)
    out = np.zeros(kmax + 1)
    np.add.at(out, ki, values[sel])
    return out[:kmax]


def kinetic(u, v, w):
    """solveFFT_kinetic (src/postproc.cpp:947-1130): shell sums of the energy of
    the full, compressive, solenoidal, left- and right-helical velocity.
    Returns (5, kmax)."""
    shape = u.shape
    kmax = min(shape) // 2
    ur, vr, wr = forward(u), forward(v), forward(w)
    kx, ky, kz, kmag = _shell_index_signed(shape)
    ku = (kx * ur + ky * vr + kz * wr) / kmag
    s = [ur - ku * kx / kmag, vr - ku * ky / kmag, wr - ku * kz / kmag]
    c = [ur - s[0], vr - s[1], wr - s[2]]
    # i k x s / |k|, as the re/im pairs of src/postproc.cpp:1058-1063
    kxs = [1j * (ky * s[2] - kz * s[1]) / kmag,
           -1j * (kx * s[2] - kz * s[0]) / kmag,
           1j * (kx * s[1] - ky * s[0]) / kmag]
    r = [0.5 * (s[a] + kxs[a]) for a in range(3)]
    l_ = [0.5 * (s[a] - kxs[a]) for a in range(3)]
    e = lambda f: 0.5 * sum(np.abs(x) ** 2 for x in f)
    return np.stack([_bin(e(x), kmag, kmax) for x in ([ur, vr, wr], c, s, l_, r)])


def scalar(f):
    """solveShell of a transformed scalar (src/fft.cpp:480-572): |f^|^2 summed
    over the modes with non-negative indices i, j, k (as written, without the
    wrap to negative wave numbers) and i^2+j^2+k^2 < kmax^2.  Returns (kmax,)."""
    shape = f.shape
    kmax = min(shape) // 2
    fr = forward(f)
    nz, ny, nx = shape
    kz, ky, kx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k0 = kx * kx + ky * ky + kz * kz
    sel = k0 < kmax ** 2
    sq = np.sqrt(k0[sel].astype(np.float64))
    fl = np.floor(sq)
    ki = np.where((sq - fl) - 0.5 <= 0, fl, fl + 1).astype(np.int64)
    out = np.zeros(kmax + 1)
    np.add.at(out, ki, (fr.real ** 2 + fr.imag ** 2)[sel])
    return out[:kmax]
