"""The mixed path evaluates WENO7's smoothness indicators (src/weno7js.cpp:80-99,
also hccs's closures and the hccs77/hccs78 rows) as sums of squares of
differences instead of the reference's expanded quadratic forms, so that FP32
does not cancel terms up to 17246 f^2.  The factorisation must be exact:
checked here in rational arithmetic on random integer stencils against the
expanded forms as the reference writes them, and the device source
(csrc/nuc_device.h, w7_sos) must hold the same coefficients.
"""
import os
import re
from fractions import Fraction as F

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def expanded(f):
    """The four indicators in the reference's expanded form on f[0..6]."""
    is0 = (f[0] * (547 * f[0] - 3882 * f[1] + 4642 * f[2] - 1854 * f[3])
           + f[1] * (7043 * f[1] - 17246 * f[2] + 7042 * f[3])
           + f[2] * (11003 * f[2] - 9402 * f[3]) + 2107 * f[3] * f[3])
    is1 = (f[1] * (267 * f[1] - 1642 * f[2] + 1602 * f[3] - 494 * f[4])
           + f[2] * (2843 * f[2] - 5966 * f[3] + 1922 * f[4])
           + f[3] * (3443 * f[3] - 2522 * f[4]) + 547 * f[4] * f[4])
    is2 = (f[2] * (547 * f[2] - 2522 * f[3] + 1922 * f[4] - 494 * f[5])
           + f[3] * (3443 * f[3] - 5966 * f[4] + 1602 * f[5])
           + f[4] * (2843 * f[4] - 1642 * f[5]) + 267 * f[5] * f[5])
    is3 = (f[3] * (2107 * f[3] - 9402 * f[4] + 7042 * f[5] - 1854 * f[6])
           + f[4] * (11003 * f[4] - 17246 * f[5] + 4642 * f[6])
           + f[5] * (7043 * f[5] - 3882 * f[6]) + 547 * f[6] * f[6])
    return is0, is1, is2, is3


# (e0 coefficients, e1 coefficients, k0, k1, k2) of the two forms
OUTER = ((547, -1394, 927), (2126, -3167), F(1, 547), F(85040, 547 * 2126 ** 2), F(203060, 1063))
INNER = ((267, -554, 247), (2126, -1345), F(1, 267), F(85040, 267 * 2126 ** 2), F(203060, 1063))


def sos(a, b, c, d, form):
    (p0, p1, p2), (r1, r2), k0, k1, k2 = form
    g0, g1, g2 = a - b, b - c, c - d
    e0 = p0 * g0 + p1 * g1 + p2 * g2
    e1 = r1 * g1 + r2 * g2
    return k0 * e0 * e0 + k1 * e1 * e1 + k2 * g2 * g2


def test_sum_of_squares_equals_expanded_forms_exactly():
    rng = np.random.default_rng(5)
    for trial in range(200):
        f = [F(int(v)) for v in rng.integers(-10 ** 6, 10 ** 6, size=7)]
        if trial % 3 == 0:   # also fractional data
            f = [v / F(int(rng.integers(1, 997))) for v in f]
        e = expanded(f)
        s = (sos(f[0], f[1], f[2], f[3], OUTER), sos(f[1], f[2], f[3], f[4], INNER),
             sos(f[5], f[4], f[3], f[2], INNER), sos(f[6], f[5], f[4], f[3], OUTER))
        assert e == s, (trial, e, s)


def test_sum_of_squares_terms_are_non_negative_weights():
    for form in (OUTER, INNER):
        assert all(k > 0 for k in form[2:])


def test_device_source_holds_the_same_coefficients():
    src = open(os.path.join(ROOT, "nuc3d-cartisian_amd", "csrc", "nuc_device.h")).read()
    body = src[src.index("__device__ __forceinline__ T w7_sos("):]
    body = body[:body.index("\n}\n")]
    for text in ("T(547) * g0 - T(1394) * g1 + T(927) * g2", "T(267) * g0 - T(554) * g1 + T(247) * g2",
                 "T(2126) * g1 - (outer ? T(3167) : T(1345)) * g2", "T(1.0 / 547.0)", "T(1.0 / 267.0)",
                 "T(85040.0 / (547.0 * 2126.0 * 2126.0))", "T(85040.0 / (267.0 * 2126.0 * 2126.0))",
                 "T(203060.0 / 1063.0) * g2 * g2"):
        assert text in body, text
    # the FP32 stencil path calls the forms on the right (mirrored) stencils
    w = src[src.index("if constexpr (sizeof(T) == 4) {"):]
    w = w[:w.index("} else {")]
    calls = re.findall(r"is(\d) = w7_sos\(f\[(\d)\], f\[(\d)\], f\[(\d)\], f\[(\d)\], (true|false)\);", w)
    assert calls == [("0", "0", "1", "2", "3", "true"), ("1", "1", "2", "3", "4", "false"),
                     ("2", "5", "4", "3", "2", "false"), ("3", "6", "5", "4", "3", "true")], calls


def _sos32(a, b, c, d, form):
    """w7_sos in FP32 arithmetic"""
    (p0, p1, p2), (r1, r2), k0, k1, k2 = form
    f = np.float32
    g0, g1, g2 = f(a) - f(b), f(b) - f(c), f(c) - f(d)
    e0 = f(p0) * g0 + f(p1) * g1 + f(p2) * g2
    e1 = f(r1) * g1 + f(r2) * g2
    return float(f(float(k0)) * e0 * e0 + f(float(k1)) * e1 * e1 + f(float(k2)) * g2 * g2)


def test_fp32_sum_of_squares_beats_fp32_expanded_on_smooth_data():
    """On smooth data with an offset the FP32 expanded forms lose the
    indicators to cancellation (4.7e-3 relative here); the sums of squares in
    FP32 keep them to about FP32's rounding of the input differences."""
    x = np.linspace(0.0, 0.6, 7)
    f64 = 3.0 + np.sin(x)
    exact = [float(v) for v in expanded([F(float(v)) for v in f64])]
    f32 = f64.astype(np.float32)
    ex32 = [float(v) for v in expanded([np.float32(v) for v in f32])]
    s32 = [_sos32(f32[0], f32[1], f32[2], f32[3], OUTER), _sos32(f32[1], f32[2], f32[3], f32[4], INNER),
           _sos32(f32[5], f32[4], f32[3], f32[2], INNER), _sos32(f32[6], f32[5], f32[4], f32[3], OUTER)]
    err_ex = max(abs(a - b) / b for a, b in zip(ex32, exact))
    err_sos = max(abs(a - b) / b for a, b in zip(s32, exact))
    print("FP32 expanded", err_ex, "FP32 sums of squares", err_sos)
    assert err_ex > 1e-4
    assert err_sos < 5e-6
    assert err_ex > 100 * err_sos
