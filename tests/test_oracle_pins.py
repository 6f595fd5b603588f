"""Closed-form pins of three oracle functions that the reference's own code
cannot check here:

- the viscous-flux assembly of NaiverStokesData3d::solveViscousFlux
  (src/NaiverStokes3d.cpp:177-303), including the -2/3 div(u) term that the
  reference also puts on the off-diagonal stresses (:267-270);
- the theta <= 0.5 blend of hccs (src/hccs.cpp:146-183), whose source does
  not compile as shipped;
- the derived turbulence scalars of postproc::solveQ (src/postproc.cpp:840-873).

Each expected value comes from the mathematics, not from the oracle's code:
the response of the cd6/cd8 first derivative to a single Fourier mode (its
modified wavenumber), the limit weights of WENO-Z on piecewise-constant data,
exact rational arithmetic of one row, and grid averages of trigonometric
products.  test_planted_mistakes_fail rebuilds the oracle with a known
mistake planted (ORC_PLANT_*) and checks that the pin of that function fails.
"""
import os
import subprocess
import sys
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


# --------------------------------------------------------------- helpers
def kappa(vis, k, h):
    """D[sin(k x)] = kappa * cos(k x) for the cd6 / cd8 first derivative on a
    periodic grid of spacing h (D6 = (3/4 d1 - 3/20 d2 + 1/60 d3)/h,
    D8 = (4/5 d1 - 1/5 d2 + 4/105 d3 - 1/280 d4)/h, d_m = f_{i+m} - f_{i-m},
    and d_m sin(kx) = 2 cos(kx) sin(m k h))."""
    c = {O.CD6: (3 / 4, -3 / 20, 1 / 60), O.CD8: (4 / 5, -1 / 5, 4 / 105, -1 / 280)}[vis]
    return 2.0 / h * sum(cm * np.sin((m + 1) * k * h) for m, cm in enumerate(c))


def _grid(n, L=2 * np.pi):
    h = L / n
    x = (np.arange(n) + 0.5) * h
    # X[d] is the coordinate along direction d on the (z, y, x)-indexed grid
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return h, (X, Y, Z)


def _cons(rho, vel, p, gamma):
    u, v, w = vel
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.concatenate([a.ravel() for a in (rho, rho * u, rho * v, rho * w, E)])


VIS_KW = dict(viscosity=O.CONSTANT, gamma=1.4, mach=0.5, reynolds=10.0, prandtl=0.72)


def _viscous_rhs(q, n, vis, npr=(1, 1, 1)):
    """RHS(Navier-Stokes) - RHS(Euler) of the same state: the viscous part
    -(dFv/dx + dGv/dy + dHv/dz), since the inviscid part is evaluated
    identically by both models."""
    out = []
    for model in (O.NS, O.EULER):
        cfg = O.make_config(n=(n, n, n), npr=npr, scheme=O.WENO5JS, vis=vis, model=model, **VIS_KW)
        s = O.Solver(cfg)
        s.set_state(q)
        out.append(s.rhs()[0])
    return (out[0] - out[1]).reshape(5, n, n, n)


def _close(got, exp, tol=1e-9):
    scale = max(np.abs(exp).max(), 1e-300)
    err = np.abs(got - exp).max() / scale
    assert err < tol, err


# --------------------------------------------------------------- viscous flux
@pytest.mark.parametrize("vis,npr", [(O.CD6, (1, 1, 1)), (O.CD8, (1, 1, 1)), (O.CD6, (2, 2, 1))])
@pytest.mark.parametrize("c,d", [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)])
def test_viscous_shear_mode(vis, npr, c, d):
    """u_c = A sin(x_d), c != d, at uniform rho and p: the only stress is
    tau_cd = mu A k1 cos(x_d) (div u = 0), so the viscous RHS is
    mu A k1^2 sin(x_d) in momentum c and, from the flux u_c tau_cd of energy
    along d, -mu A^2 k1 k2 / 2 cos(2 x_d); mass and the other momenta get 0."""
    n, A, p0 = 16, 0.3, 2.0
    h, X = _grid(n)
    mu = 1.0 / VIS_KW["reynolds"]
    vel = [np.zeros_like(X[0]) for _ in range(3)]
    vel[c] = A * np.sin(X[d])
    q = _cons(np.ones_like(X[0]), vel, np.full_like(X[0], p0), VIS_KW["gamma"])
    got = _viscous_rhs(q, n, vis, npr)
    k1, k2 = kappa(vis, 1, h), kappa(vis, 2, h)
    exp = np.zeros_like(got)
    exp[1 + c] = mu * A * k1 * k1 * np.sin(X[d])
    exp[4] = -mu * A * A * k1 * k2 / 2 * np.cos(2 * X[d])
    _close(got, exp)


@pytest.mark.parametrize("vis", [O.CD6, O.CD8])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_viscous_compression_mode_keeps_offdiagonal_dilatation(vis, d):
    """u_d = A sin(x_d): div u = A k1 cos(x_d), tau_dd = 4/3 mu div u and, as
    the reference codes every stress with -2/3 mu div u
    (src/NaiverStokes3d.cpp:267-270), tau_cd = -2/3 mu div u for c != d too.
    Viscous RHS: 4/3 mu A k1^2 sin(x_d) in momentum d, -2/3 mu A k1^2 sin(x_d)
    in the two other momenta (zero for the textbook stress), and
    -2/3 mu A^2 k1 k2 cos(2 x_d) in energy."""
    n, A, p0 = 16, 0.3, 2.0
    h, X = _grid(n)
    mu = 1.0 / VIS_KW["reynolds"]
    vel = [np.zeros_like(X[0]) for _ in range(3)]
    vel[d] = A * np.sin(X[d])
    q = _cons(np.ones_like(X[0]), vel, np.full_like(X[0], p0), VIS_KW["gamma"])
    got = _viscous_rhs(q, n, vis)
    k1, k2 = kappa(vis, 1, h), kappa(vis, 2, h)
    exp = np.zeros_like(got)
    for cc in range(3):
        exp[1 + cc] = (4 / 3 if cc == d else -2 / 3) * mu * A * k1 * k1 * np.sin(X[d])
    exp[4] = -2 / 3 * mu * A * A * k1 * k2 * np.cos(2 * X[d])
    _close(got, exp)


@pytest.mark.parametrize("vis", [O.CD6, O.CD8])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_viscous_heat_flux_mode(vis, d):
    """Fluid at rest with T = 1 + eps sin(x_d) (rho = 1, p = rho T/(gamma M^2)):
    no stress; the energy flux along d is kappa~ dT/dx_d with
    kappa~ = mu/((gamma-1) M^2 Pr) (src/physicsModel.cpp:234-266), so the
    viscous energy RHS is kappa~ eps k1^2 sin(x_d) and everything else is 0."""
    n, eps = 16, 0.05
    g, M, Pr = VIS_KW["gamma"], VIS_KW["mach"], VIS_KW["prandtl"]
    h, X = _grid(n)
    mu = 1.0 / VIS_KW["reynolds"]
    kap = mu / ((g - 1) * M * M * Pr)
    T = 1.0 + eps * np.sin(X[d])
    rho = np.ones_like(T)
    zero = np.zeros_like(T)
    q = _cons(rho, (zero, zero, zero), rho * T / (g * M * M), g)
    got = _viscous_rhs(q, n, vis)
    k1 = kappa(vis, 1, h)
    exp = np.zeros_like(got)
    exp[4] = kap * eps * k1 * k1 * np.sin(X[d])
    _close(got, exp)


# --------------------------------------------------------------- hccs theta <= 0.5
def _hccs_rows_limit(f):
    """(lo, di, hi, rhs) of an interior hccs row on piecewise-constant data with
    one jump in its 5-point stencil, from the limit of the WENO-Z weights: a
    sub-stencil with the jump has beta > 0 = beta of the smooth ones, whose
    alpha_k ~ c_k (tau/1e-40)^2 dominate, so omega = the CRWENO ideal weights
    (0.2, 0.5, 0.3) renormalised over the smooth sub-stencils; sum(alpha) ~
    1e80, theta ~ 1e-160 <= 0.5, and the row is (1-theta)/theta times the
    CRWENO row of those weights plus the upwind-compact row, i.e. the CRWENO row
    itself to 1e-160 relative.  Without a jump theta = 1: upwind compact."""
    smooth = [f[k] == f[k + 1] == f[k + 2] for k in range(3)]
    if all(smooth):
        q7 = (-f[0] + 19 * f[1] + 239 * f[2] + 159 * f[3] + 4 * f[4]) / 420
        return 2 / 7, 4 / 7, 1 / 7, q7
    c = np.array([0.2, 0.5, 0.3]) * np.array(smooth)
    w0, w1, w2 = c / c.sum()
    lo = 2 / 3 * w0 + 1 / 3 * w1
    di = 1 / 3 * w0 + 2 / 3 * (w1 + w2)
    hi = 1 / 3 * w2
    rhs = (w0 * f[1] + (5 * (w0 + w1) + w2) * f[2] + (w1 + 5 * w2) * f[3]) / 6
    return lo, di, hi, rhs


@pytest.mark.parametrize("jumps", [((20, 1.0),), ((14, 2.0), (27, -1.5)), ((16, -1.0), (23, 3.0))])
def test_hccs_blend_branch_on_steps(jumps):
    """Steps at least 8 cells from the line ends: the explicit closure and edge
    interfaces see constant data; the interior rows around each jump take the
    theta <= 0.5 branch.  The line solution must equal the dense solve of the
    limit rows above, closed by the constant values."""
    n, bf, ts = 40, 4, 4
    cells = np.arange(-bf, n + bf)
    f = np.zeros(n + 2 * bf)
    for at, dv in jumps:
        f[cells >= at] += dv
    F = lambda k: f[bf + k]
    out = O.line_reconstruct(O.HCCS, 1, f, n, ts, bf)
    M = n - 2 * ts + 1
    A = np.zeros((M, M))
    b = np.zeros(M)
    nblend = 0
    for r in range(M):
        m = ts + r
        if r in (0, M - 1):
            A[r, r] = 1.0
            b[r] = F(m - 1)          # constant data: every candidate value is the cell value
            continue
        st = [F(m - 3 + z) for z in range(5)]
        lo, di, hi, rhs = _hccs_rows_limit(st)
        nblend += not (lo == 2 / 7 and hi == 1 / 7)
        A[r, r - 1], A[r, r], A[r, r + 1] = lo, di, hi
        b[r] = rhs
    assert nblend >= 4 * len(jumps)
    x = np.linalg.solve(A, b)
    np.testing.assert_allclose(out[ts:n - ts + 1], x, rtol=0, atol=1e-13)
    np.testing.assert_allclose(out[:ts], f[bf - 1:bf + ts - 1], atol=1e-15)


def _row_exact(f):
    """One interior hccs row in exact rational arithmetic from the reference's
    definitions (src/hccs.cpp:138-183): smoothness indicators with 13/12 and 1/4,
    tau = |is2 - is0|, alpha_k = c_k (1 + (tau/is_k)^2) (the 1e-40 is below the
    rounding of these indicators), theta = 1/(1 + (sum alpha - 1)^2)."""
    f = [Fraction(v) for v in f]
    g0, g1 = Fraction(13, 12), Fraction(1, 4)
    is0 = g0 * (f[0] - 2 * f[1] + f[2]) ** 2 + g1 * (f[0] - 4 * f[1] + 3 * f[2]) ** 2
    is1 = g0 * (f[1] - 2 * f[2] + f[3]) ** 2 + g1 * (f[1] - f[3]) ** 2
    is2 = g0 * (f[2] - 2 * f[3] + f[4]) ** 2 + g1 * (3 * f[2] - 4 * f[3] + f[4]) ** 2
    tau = abs(is2 - is0)
    c = (Fraction(1, 5), Fraction(1, 2), Fraction(3, 10))
    al = [ck * (1 + (tau / isk) ** 2) for ck, isk in zip(c, (is0, is1, is2))]
    S = sum(al)
    w = [a / S for a in al]
    theta = 1 / (1 + (S - 1) ** 2)
    q7 = (-f[0] + 19 * f[1] + 239 * f[2] + 159 * f[3] + 4 * f[4]) / 420
    if theta > Fraction(1, 2):
        return theta, (Fraction(2, 7), Fraction(4, 7), Fraction(1, 7), q7)
    k = (1 - theta) / theta
    lo = (Fraction(2, 3) * w[0] + Fraction(1, 3) * w[1]) * k + Fraction(2, 7)
    di = (Fraction(1, 3) * w[0] + Fraction(2, 3) * (w[1] + w[2])) * k + Fraction(4, 7)
    hi = Fraction(1, 3) * w[2] * k + Fraction(1, 7)
    rhs = (w[0] * f[1] + (5 * (w[0] + w[1]) + w[2]) * f[2] + (w[1] + 5 * w[2]) * f[3]) / 6 * k + q7
    return theta, (lo, di, hi, rhs)


# stencils with theta in (0, 0.5] at moderate values, and one above 0.5
ROW_KATS = [(0, 2, 4, 4, 1), (1, 0, 3, 2, 4), (2, 2, 1, 2, 3), (1, 1, 3, 0, 1), (4, 2, 0, 0, 1),
            (2, 4, 1, 4, 0), (3, 2, 4, 1, 2)]


def test_hccs_row_exact_rational_kats():
    """Moderate theta, where both halves of the blend matter: catches 1/theta
    or theta/(1-theta) in place of (1-theta)/theta, a swapped weight or a
    dropped upwind-compact term."""
    seen = set()
    for st in ROW_KATS:
        theta, exact = _row_exact(st)
        seen.add("blend" if theta <= Fraction(1, 2) else "compact")
        if theta <= Fraction(1, 2):
            assert Fraction(1, 20) < theta, st       # genuinely mixed
        got = O.hccs_row(np.array(st, dtype=float))
        np.testing.assert_allclose(got, [float(e) for e in exact], rtol=1e-13, atol=1e-15)
    assert seen == {"blend", "compact"}


def test_hccs_line_translation_invariance_mixed_branches():
    """Every row, blended or not, reproduces constants (the CRWENO and the
    upwind-compact rows each sum to their rhs on constant data), so the line
    solution commutes with adding a constant: x(f + c) = x(f) + c.  Random data
    puts most rows on the blend branch with moderate theta."""
    rng = np.random.default_rng(7)
    n, bf, ts = 48, 4, 4
    f = rng.standard_normal(n + 2 * bf)
    nblend = sum(_row_exact(np.round(f[k:k + 5], 6))[0] <= Fraction(1, 2) for k in range(bf, n))
    assert nblend > n // 4
    base = O.line_reconstruct(O.HCCS, 1, f, n, ts, bf)
    for c in (0.75, -3.0, 10.0):
        sh = O.line_reconstruct(O.HCCS, 1, f + c, n, ts, bf)
        np.testing.assert_allclose(sh, base + c, rtol=0, atol=1e-11 * (1 + abs(c)))


# --------------------------------------------------------------- derived scalars
@pytest.mark.parametrize("vis", [O.CD6, O.CD8])
def test_solveq_derived_scalars_single_modes(vis):
    """u = A sin x + B sin 2x, v = C sin x, w = 0 at uniform rho = 1 and p0.
    Grid averages over full periods: <sin^2> = 1/2, <cos^2 x cos 2x> = 1/4 and
    the odd products vanish, with ux = a cos x + b cos 2x (a = A k1, b = B k2):
      u_rms^2 = <|u|^2>/3 = (A^2+B^2+C^2)/6;   div = <ux^2>/3 = (a^2+b^2)/6;
      skew = (<ux^3>/3) / div^1.5, <ux^3> = 3 a^2 b / 4;
      lambda = sqrt(u_rms^2/div); Re_lambda = u_rms lambda rho/mu Re;
      tau_t = lambda/u_rms; M_t = sqrt(3 u_rms^2)/c, c = sqrt(gamma p0);
      mu_mean = (1+S)/(T+S) T^1.5 with S = T_ref/T_inf (src/postproc.cpp:840-873)."""
    n, A, B, C = 16, 0.4, 0.15, 0.25
    g, M, Re = 1.4, 0.3, 250.0
    T0 = 1.3
    p0 = T0 / (g * M * M)
    h, X = _grid(n)
    x = X[0]
    u = A * np.sin(x) + B * np.sin(2 * x)
    v = C * np.sin(x)
    w = np.zeros_like(x)
    q = _cons(np.ones_like(x), (u, v, w), np.full_like(x, p0), g)
    cfg = O.make_config(n=(n, n, n), scheme=O.WENO5JS, vis=vis, model=O.NS, gamma=g, mach=M, reynolds=Re)
    s = O.Solver(cfg)
    s.set_state(q)
    st = s.stats()
    k1, k2 = kappa(vis, 1, h), kappa(vis, 2, h)
    a, b = A * k1, B * k2
    urms2 = (A * A + B * B + C * C) / 6
    div = (a * a + b * b) / 6
    S = 110.4 / 288.0
    mu = (1 + S) / (T0 + S) * T0 ** 1.5
    lam = np.sqrt(urms2 / div)
    exp = {16: np.sqrt(urms2), 17: np.sqrt(3 * urms2) / np.sqrt(g * p0), 18: np.sqrt(div), 19: mu,
           20: np.sqrt(urms2) * lam * 1.0 / mu * Re, 21: lam, 22: 1.0, 23: lam / np.sqrt(urms2),
           24: (3 * a * a * b / 4 / 3) / div ** 1.5}
    for i, e in exp.items():
        assert abs(st[i] - e) <= 1e-12 * abs(e), (i, st[i], e)
    # the 16 means of the same fields: KE = <rho |u|^2>/2, omega_z = vx - uy = C k1 cos x
    assert abs(st[15] - (A * A + B * B + C * C) / 4) <= 1e-13
    assert abs(st[12] - (C * k1) ** 2 / 4) <= 1e-13


# --------------------------------------------------------------- planted mistakes
PLANTS = [
    ("ORC_PLANT_TAUXY", "test_viscous_compression_mode_keeps_offdiagonal_dilatation"),
    ("ORC_PLANT_HEAT", "test_viscous_heat_flux_mode"),
    ("ORC_PLANT_EFLUX", "test_viscous_shear_mode"),
    ("ORC_PLANT_HCCS_SWAP", "test_hccs_blend_branch_on_steps"),
    ("ORC_PLANT_HCCS_SWAP", "test_hccs_row_exact_rational_kats"),
    ("ORC_PLANT_HCCS_INV", "test_hccs_row_exact_rational_kats"),
    ("ORC_PLANT_LAMBDA", "test_solveq_derived_scalars_single_modes"),
    ("ORC_PLANT_SKEW", "test_solveq_derived_scalars_single_modes"),
]


@pytest.mark.parametrize("macro,test", PLANTS)
def test_planted_mistakes_fail(macro, test, tmp_path):
    """Build the oracle with one mistake planted and run the pin in a fresh
    process against that build: it must fail."""
    so = tmp_path / "liboracle_planted.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-std=c99", "-ffp-contract=off", "-fopenmp", "-shared",
                           "-D" + macro, "-o", str(so), os.path.join(ROOT, "oracle", "nuc_oracle.c"), "-lm"])
    env = dict(os.environ, NUC_ORACLE_LIB=str(so))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_oracle_pins.py") + "::" + test],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout
