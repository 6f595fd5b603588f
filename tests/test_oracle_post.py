"""Pins of the oracle's time-averaged statistics and TKE budget
(postproc::solveAveraged0/1/2, src/postproc.cpp:1186-2110).

Each check comes from outside the restated formulas: the primaries from the
definition of the conserved variables, gradients from the analytic derivative
of the Taylor-Green field, the running mean from the arithmetic mean of the
samples, the Favre TKE and production of a two-sample average from the sample
covariance (uniform density), the vanishing of every fluctuation term when one
state is averaged repeatedly, and the zero volume mean of the terms written as
divergences on a periodic box.
"""
import numpy as np
import pytest

import oracle as O

GAMMA, MACH = 1.4, 0.3
TERMS = ["tke", "production", "transport", "pressure_diffusion", "p1", "p2", "p3", "viscous_diffusion",
         "v1", "v2", "v3", "convection"]


def _solver(n=(16, 16, 16), npr=(1, 1, 1), vis=O.CD6):
    c = O.make_config(n=n, npr=npr, scheme=O.CRWENO5, vis=vis, riemann=O.LF, model=O.NS, mach=MACH,
                      reynolds=400.0)
    return O.Solver(c), c


def _hit(c, seed, u0=0.4):
    """Solenoidal random-phase velocity, rho = 1, uniform p (numpy, independent of the product)."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = c.nz, c.ny, c.nx
    N = nx * ny * nz
    z, y, x = np.meshgrid(*[2 * np.pi * np.arange(m) / m for m in (nz, ny, nx)], indexing="ij")
    u = np.zeros((3, nz, ny, nx))
    for _ in range(6):
        k = rng.integers(-2, 3, size=3)
        if not k.any():
            continue
        a = rng.standard_normal(3)
        a -= a.dot(k) / k.dot(k) * k            # amplitude normal to k: divergence free
        ph = k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 2 * np.pi)
        u += a[:, None, None, None] * np.cos(ph)
    u *= u0 / np.sqrt((u ** 2).mean())
    p0 = 1.0 / (GAMMA * MACH ** 2)
    q = np.zeros((5, N))
    q[0] = 1.0
    q[1:4] = u.reshape(3, N)
    q[4] = p0 / (GAMMA - 1) + 0.5 * (u ** 2).sum(0).reshape(N)
    return q.ravel()


def _cd6(f, axis, h):
    r = lambda m: np.roll(f, -m, axis=axis) - np.roll(f, m, axis=axis)
    return (0.75 * r(1) - 0.15 * r(2) + (1.0 / 60.0) * r(3)) / h


def test_primaries_are_the_arithmetic_mean_of_the_samples():
    s, c = _solver()
    N = c.nx * c.ny * c.nz
    P = O.Post(s, 1)
    states = [_hit(c, k) for k in range(3)]
    for q in states:
        s.set_state(q)
        P.update()
    assert P.steps() == 3
    Q = np.stack(states).reshape(3, 5, N)
    rho = Q[:, 0]
    u = Q[:, 1:4] / rho[:, None]
    p = np.abs((Q[:, 4] - 0.5 * rho * (u ** 2).sum(1)) * (GAMMA - 1))
    T = GAMMA * MACH ** 2 * p / rho
    expect = {0: rho, 1: u[:, 0], 2: u[:, 1], 3: u[:, 2], 4: p, 5: T,
              6: p / rho / (GAMMA - 1), 7: np.sqrt(GAMMA * p / rho)}
    for i, f in expect.items():
        got = P.field(0, i)
        assert np.allclose(got, f.mean(0), rtol=1e-13, atol=1e-13 * np.abs(f).max()), i


def test_gradients_of_the_taylor_green_field():
    s, c = _solver(n=(32, 32, 32))
    q = O.init_tgv(c)
    s.set_state(q)
    P = O.Post(s, 1)
    P.update()
    n = 32
    z, y, x = np.meshgrid(*[2 * np.pi * (np.arange(n) + 0.5) / n] * 3, indexing="ij")
    dudx = (np.cos(x) * np.cos(y) * np.cos(z)).ravel()          # u = sin x cos y cos z
    dvdy = (-np.cos(x) * np.cos(y) * np.cos(z)).ravel()         # v = -cos x sin y cos z
    dudz = (-np.sin(x) * np.cos(y) * np.sin(z)).ravel()
    # cd6 truncation at kh = 2 pi / 32: (kh)^6 / 140 ~ 1e-6
    assert np.abs(P.field(1, 0) - dudx).max() < 1e-5
    assert np.abs(P.field(1, 4) - dvdy).max() < 1e-5
    assert np.abs(P.field(1, 2) - dudz).max() < 1e-5
    # omega_2 = dv/dx - du/dy and its square
    o2 = P.field(1, 17)
    assert np.allclose(P.field(1, 20), o2 ** 2, rtol=1e-15, atol=0)
    assert np.allclose(o2, P.field(1, 3) - P.field(1, 1), rtol=0, atol=1e-15)


def test_two_sample_tke_and_production_are_the_sample_covariance():
    s, c = _solver()
    N = c.nx * c.ny * c.nz
    h = 2 * np.pi / c.nx
    qa, qb = _hit(c, 10), _hit(c, 11)
    P = O.Post(s, 3)
    for q in (qa, qb):
        s.set_state(q)
        P.update()
    ua = qa.reshape(5, N)[1:4]
    ub = qb.reshape(5, N)[1:4]
    d = 0.5 * (ua - ub)                                   # fluctuation about the two-sample mean
    tke = 0.5 * (d ** 2).sum(0)
    got = P.field(4, 0)
    assert np.abs(got - tke).max() <= 1e-12 * tke.max()
    # production = sum_ij R_ij d_j <rho u_i>, R_ij = <u_i u_j> - <u_i><u_j> (rho = 1)
    mean = (0.5 * (ua + ub)).reshape(3, c.nz, c.ny, c.nx)
    grad = [[_cd6(mean[i], 2 - j, h).ravel() for j in range(3)] for i in range(3)]
    prod = sum(d[i] * d[j] * grad[i][j] for i in range(3) for j in range(3))
    got = P.field(4, 1)
    assert np.abs(got - prod).max() <= 1e-11 * np.abs(prod).max()


def test_every_fluctuation_term_vanishes_for_a_repeated_state():
    s, c = _solver(n=(20, 16, 12), npr=(2, 1, 1), vis=O.CD8)
    q = _hit(c, 5)
    s.set_state(q)
    s.step(2)                                             # non-uniform density and pressure
    P = O.Post(s, 3)
    for _ in range(3):
        P.update()
    B = P.group(4)
    # scale of the products whose differences form the terms
    scale = {
        "tke": np.abs(P.field(2, 10)).max(),
        "production": np.abs(P.field(2, 10)).max() * np.abs(B[12]).max(),
        "p1": np.abs(P.field(0, 1)).max() * np.abs(P.field(1, 9)).max(),
        "p3": np.abs(P.field(3, 3)).max(),
        "v3": np.abs(P.field(3, 30)).max(),
    }
    for name in ("tke", "production", "p1", "p3", "v3"):
        got = np.abs(B[TERMS.index(name)]).max()
        assert got <= 1e-12 * scale[name], (name, got, scale[name])
    for name in ("p2", "v2", "transport", "convection"):
        got = np.abs(B[TERMS.index(name)]).max()
        assert got <= 1e-9 * scale["p3"] + 1e-14, (name, got)


def test_divergence_terms_have_zero_volume_mean():
    s, c = _solver(n=(16, 16, 16))
    s.set_state(_hit(c, 7))
    P = O.Post(s, 3)
    for _ in range(3):
        P.update()
        s.step(1)
    B = P.group(4)
    for name in ("p2", "v2", "transport", "convection"):
        b = B[TERMS.index(name)]
        assert abs(b.mean()) <= 1e-13 * np.abs(b).max(), name
    # and the non-divergence terms do not vanish
    for name in ("tke", "production", "p3", "v3"):
        assert np.abs(B[TERMS.index(name)]).max() > 0


def test_levels_allocate_their_groups_only():
    s, c = _solver()
    s.set_state(_hit(c, 1))
    P = O.Post(s, 1)
    P.update()
    P.field(1, 20)
    with pytest.raises(ValueError):
        P.field(2, 0)
    with pytest.raises(ValueError):
        O.Post(s, 4)
