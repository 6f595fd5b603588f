"""Pins of the oracle's analytic forcings and rotating-frame sources
(src/physicsModel.cpp:1679-1913, applied in solveSource's order, :1280-1319).
Each source is isolated as RHS(with) - RHS(without) on the same state and
checked against a property of the force field, not its formula: the ABC field
is Beltrami (curl f = k f), TG is solenoidal and planar, Coriolis does no work
and is normal to the velocity, the centrifugal well at rest is radial with the
profile 1/(1 + (r/r0)^20); energy terms equal the force times the velocity
(and, as the reference writes it, go to the y-momentum for the centrifugal
force)."""
import numpy as np

import oracle as O

N1 = 32


def _cfg(**kw):
    base = dict(n=(N1, N1, N1), scheme=O.WENO5JS, vis=O.CD6, riemann=O.LF, model=O.NS, mach=0.2,
                reynolds=500.0)
    base.update(kw)
    return O.make_config(**base)


def _state(c, seed, rest=False):
    q = O.init_tgv(c).reshape(5, -1)
    if rest:
        p = (q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0])
        q[1:4] = 0.0
        q[4] = p
    else:
        rng = np.random.default_rng(seed)
        q[3] += 0.1 * np.sin(np.linspace(0, 30, q.shape[1]) + rng.random())
        q[0] *= 1 + 0.02 * rng.standard_normal(q.shape[1])
    return q.ravel().copy()


def _delta(q, **kw):
    s0, s1 = O.Solver(_cfg()), O.Solver(_cfg(**kw))
    s0.set_state(q)
    s1.set_state(q)
    return (s1.rhs()[0] - s0.rhs()[0]).reshape(5, N1, N1, N1)


def _spec_d(f, axis):
    k = np.fft.fftfreq(N1, d=1.0 / N1)
    sh = [1, 1, 1]
    sh[axis] = N1
    return np.real(np.fft.ifft(1j * k.reshape(sh) * np.fft.fft(f, axis=axis), axis=axis))


def _prims(q):
    Q = q.reshape(5, N1, N1, N1)
    return Q[0], Q[1] / Q[0], Q[2] / Q[0], Q[3] / Q[0]


def test_abc_forcing_is_beltrami_and_does_work_f_dot_u():
    q = _state(_cfg(), 1)
    d = _delta(q, forcing=O.FORCE_ABC, forcing_amp=0.3, forcing_k=2.0)
    f = [-d[1], -d[2], -d[3]]                     # RHS -= f
    # arrays are (z, y, x): axis 2 = x, 1 = y, 0 = z
    curl = [_spec_d(f[2], 1) - _spec_d(f[1], 0), _spec_d(f[0], 0) - _spec_d(f[2], 2),
            _spec_d(f[1], 2) - _spec_d(f[0], 1)]
    for a in range(3):
        assert np.abs(curl[a] - 2.0 * f[a]).max() <= 1e-10 * np.abs(f[a]).max()
    _, u, v, w = _prims(q)
    assert np.abs(-d[4] - (f[0] * u + f[1] * v + f[2] * w)).max() <= 1e-13 * np.abs(d[4]).max()
    assert np.all(d[0] == 0)


def test_tg_forcing_is_planar_and_solenoidal():
    q = _state(_cfg(), 2)
    d = _delta(q, forcing=O.FORCE_TG, forcing_amp=0.2, forcing_k=1.0)
    assert np.all(d[3] == 0) and np.all(d[0] == 0)
    div = _spec_d(d[1], 2) + _spec_d(d[2], 1)
    assert np.abs(div).max() <= 1e-11 * np.abs(d[1]).max()
    _, u, v, _ = _prims(q)
    assert np.abs(d[4] - (d[1] * u + d[2] * v)).max() <= 1e-13 * np.abs(d[4]).max()


def test_coriolis_does_no_work():
    q = _state(_cfg(), 3)
    d = _delta(q, coriolis=True, rossby=0.7)
    rho, u, v, w = _prims(q)
    assert np.all(d[3] == 0) and np.all(d[4] == 0) and np.all(d[0] == 0)
    assert np.abs(d[1] * u + d[2] * v).max() <= 1e-14 * (rho * 0.7 * (u * u + v * v)).max()
    mag = (0.7 * rho) ** 2 * (u * u + v * v)
    assert np.abs(d[1] ** 2 + d[2] ** 2 - mag).max() <= 1e-12 * mag.max()


def test_centrifugal_well_at_rest_is_radial():
    c = _cfg()
    q = _state(c, 4, rest=True)
    xc, yc, r0, ro = 3.0, 3.3, 1.2, 0.8
    d = _delta(q, centrifugal=True, rossby=ro, x_center=xc, y_center=yc, well_rad=r0)
    rho = q.reshape(5, N1, N1, N1)[0]
    h = 2 * np.pi / N1
    z, y, x = np.meshgrid(*[(np.arange(N1) + 0.5) * h] * 3, indexing="ij")
    r = np.sqrt((x - xc) ** 2 + (y - yc) ** 2)
    prof = rho * ro * ro / (1.0 + (r / r0) ** 20) / 4.0
    # the difference of two O(1) right-hand sides carries their rounding, so
    # the check is against the largest force, not element by element
    fx, fy = prof * (x - xc), prof * (y - yc)
    assert np.abs(-d[1] - fx).max() <= 1e-13 * np.abs(fx).max()
    assert np.abs(-d[2] - fy).max() <= 1e-13 * np.abs(fy).max()
    assert np.all(d[4] == 0) and np.all(d[3] == 0)   # the energy term goes to y-momentum (zero at rest)


def test_block_origin_shifts_the_source_fields():
    """The block.in origin is subtracted from the cell centres
    (src/block.cpp:116-118).  With an origin of whole cells, (2h, -3h, h), every
    source field on a state at rest is the field without origin rolled by
    (2, -3, 1) cells, whatever the formula; the origin must also change it.
    The well (not periodic) is narrow enough to vanish to 1e-12 at the box
    faces that the roll wraps around (1e-11 of its peak)."""
    h = 2 * np.pi / N1
    kw = dict(forcing=O.FORCE_ABC, forcing_amp=0.05, centrifugal=True, rossby=0.6, x_center=3.0,
              y_center=3.3, well_rad=0.7)
    q = _state(_cfg(), 0, rest=True)
    base = _delta(q, **kw)
    shifted = _delta(q, origin=(2 * h, -3 * h, h), **kw)
    # array axes are (z, y, x); a field f(x - o) sampled at cell i equals f at cell i - o/h
    rolled = np.roll(base, shift=(1, -3, 2), axis=(1, 2, 3))
    np.testing.assert_allclose(shifted, rolled, rtol=0, atol=1e-11 * np.abs(base).max())
    assert np.abs(shifted - base).max() > 1e-3 * np.abs(base).max()
