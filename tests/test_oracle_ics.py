"""Pins of the oracle's initial_default and initial_ivc (src/block.cpp:285-357)
and the host generators' agreement with them.

The default field is checked against its physical meaning (at rest, T = 1).
The isentropic vortex is checked against the properties that define it rather
than its formulas: p = rho^gamma everywhere, a velocity perturbation tangential
to the circles about (5, 5) with speed b/(2 pi) r exp((1 - r^2)/2), the free
stream (1, 0.5, 0, 1) far from the centre, and the cyclostrophic balance
dp/dr = rho v_theta^2 / r, which holds exactly when the density exponent 2.5
equals 1/(gamma - 1) (gamma = 1.4) -- a wrong exponential or exponent breaks it.
"""
import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O

GAMMA = 1.4


def _prims(q, shape):
    q = q.reshape((5,) + shape)
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    p = (q[4] - 0.5 * rho * (u * u + v * v + w * w)) * (GAMMA - 1.0)
    return rho, u, v, w, p


def _centres(n, L):
    return [(np.arange(m) + 0.5) * (l / m) for m, l in zip(n, L)]


def test_default_is_at_rest_with_unit_temperature():
    c = O.make_config(n=(12, 10, 8), mach=0.3)
    rho, u, v, w, p = _prims(O.init_default(c), (8, 10, 12))
    assert np.all(rho == 1.0) and not u.any() and not v.any() and not w.any()
    T = GAMMA * 0.3 ** 2 * p / rho                       # src/physicsModel.cpp:342
    assert np.allclose(T, 1.0, rtol=1e-14, atol=0)


def test_isentropic_vortex_properties():
    n, L = (320, 320, 2), (10.0, 10.0, 1.0)
    c = O.make_config(n=n, L=L, model=O.EULER)
    shape = (n[2], n[1], n[0])
    rho, u, v, w, p = _prims(O.init_ivc(c), shape)
    xs, ys, _ = _centres(n, L)
    X, Y = np.meshgrid(xs - 5.0, ys - 5.0, indexing="xy")  # (ny, nx)
    r = np.hypot(X, Y)
    b = 0.5
    # isentropic
    assert np.allclose(p, rho ** GAMMA, rtol=1e-13, atol=0)
    # z-independent, no w
    assert np.array_equal(rho[0], rho[1]) and not w.any()
    # tangential perturbation with the closed-form speed
    du, dv = u[0] - 0.5, v[0]
    assert np.abs(du * X + dv * Y).max() <= 1e-15 * r.max()
    speed = b / (2 * np.pi) * r * np.exp(0.5 * (1 - r * r))
    assert np.allclose(np.hypot(du, dv), speed, rtol=1e-12, atol=2e-16)   # u - 0.5 cancels
    # free stream at the corners
    far = r > 4.9
    assert np.allclose(rho[0][far], 1.0, atol=1e-9) and np.allclose(u[0][far], 0.5, atol=1e-9)
    # cyclostrophic balance, x component: dp/dx = rho v_theta^2 / r * X / r, 6th-order differences
    h = L[0] / n[0]
    P = p[0]
    d = (45 * (P[:, 4:-2] - P[:, 2:-4]) - 9 * (P[:, 5:-1] - P[:, 1:-5]) + (P[:, 6:] - P[:, :-6])) / (60 * h)
    R = r[:, 3:-3]
    force = rho[0][:, 3:-3] * speed[:, 3:-3] ** 2 / R ** 2 * X[:, 3:-3]
    assert np.abs(d - force).max() <= 1e-6 * np.abs(force).max()


def test_ivc_cell_centres_follow_the_rank_grid():
    n, L = (40, 40, 4), (10.0, 10.0, 1.0)
    a = O.init_ivc(O.make_config(n=n, L=L, npr=(1, 1, 1)))
    b = O.init_ivc(O.make_config(n=n, L=L, npr=(2, 4, 1)))
    assert np.allclose(a, b, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("npr", [(1, 1, 1), (2, 2, 1)], ids=str)
def test_host_generators_equal_the_oracle(npr):
    kw = dict(n=(24, 20, 4), L=(10.0, 10.0, 2.0), npr=npr, mach=0.25)
    gc, oc = G.make_config(**kw), O.make_config(**kw)
    assert np.array_equal(G.init_default(gc), O.init_default(oc))
    a, b = G.init_ivc(gc), O.init_ivc(oc)
    assert np.allclose(a, b, rtol=2e-15, atol=1e-16)
