"""Pin the C restatement (oracle/nuc_oracle.c) against the reference's own
scheme objects, compiled unmodified from src/ into oracle/_ref/libnucref.so.

Bit-exact: both are compiled without FMA contraction, and the oracle keeps the
reference's evaluation order and its libm pow() calls.
"""
import numpy as np
import pytest

import oracle as O



@pytest.fixture(autouse=True)
def _reference_objects():
    """Loaded when a test here runs, not at collection: a GPU-only run
    (-m gpu) deselects these tests and never maps the reference's objects."""
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")

SHAPES = [(16, 3, 2), (24, 2, 3), (40, 2, 1), (64, 1, 2), (13, 2, 2)]


def _data(rng, n1, n2, n3, bf, kind):
    if kind == "smooth":
        x = np.linspace(0, 2 * np.pi, n1 + 2 * bf, endpoint=False)
        base = np.sin(x) + 0.3 * np.cos(3 * x) + 2.0
        line = np.broadcast_to(base, (n3, n2, n1 + 2 * bf)) + 0.01 * rng.standard_normal((n3, n2, n1 + 2 * bf))
    elif kind == "step":
        line = np.where(rng.random((n3, n2, n1 + 2 * bf)) < 0.5, 1.0, 10.0)
    else:
        line = rng.standard_normal((n3, n2, n1 + 2 * bf)) * 3.0 + 1.0
    line = np.ascontiguousarray(line)
    return (np.ascontiguousarray(line[:, :, bf:bf + n1]), np.ascontiguousarray(line[:, :, :bf]),
            np.ascontiguousarray(line[:, :, bf + n1:]))


@pytest.mark.parametrize("scheme", [O.WENO5JS, O.WENO5Z, O.WENO7JS, O.CRWENO5, O.WENO3JS, O.UPWIND1ST, O.HCCS78,
                                    O.HCCS77])
@pytest.mark.parametrize("upwind", [1, -1])
@pytest.mark.parametrize("kind", ["smooth", "random", "step"])
def test_reconstruct_bitexact(scheme, upwind, kind):
    rng = np.random.default_rng(100 + scheme * 7 + upwind + 3 * len(kind))
    ts = O.halo(scheme)
    for (n1, n2, n3) in SHAPES:
        bf = 4
        fin, gL, gR = _data(rng, n1, n2, n3, bf, kind)
        a = O.field_reconstruct(scheme, upwind, ts, fin, gL, gR)
        b = O.field_reconstruct(scheme, upwind, ts, fin, gL, gR, use_ref=True)
        assert np.array_equal(a, b), (scheme, upwind, n1, np.abs(a - b).max())


@pytest.mark.parametrize("scheme,ts", [(O.CD6, 3), (O.CD6, 4), (O.CD8, 4), (O.CD2, 1), (O.CD2, 2), (O.CD2, 4)])
def test_diff_bitexact(scheme, ts):
    rng = np.random.default_rng(7 + ts + 10 * scheme)
    for (n1, n2, n3) in SHAPES:
        bf = 4
        fin, gL, gR = _data(rng, n1, n2, n3, bf, "random")
        a = O.field_diff(scheme, ts, fin, 0.173, gL, gR)
        b = O.field_diff(scheme, ts, fin, 0.173, gL, gR, use_ref=True)
        assert np.array_equal(a, b)


def test_cd8_with_3_wide_halo_reads_outside_stencil():
    """cd8 handed a tilesize of 3 (crweno5/weno5 halo, src/fieldOperator.cpp:520-526)
    reads in[i-4] for i=3 -- the previous line's last cell, or before the array --
    and in[i+4] for i=n-4, the next line's first cell.  The oracle widens the halo
    to 4 instead; this documents the divergence."""
    rng = np.random.default_rng(3)
    fin, gL, gR = _data(rng, 20, 3, 2, 4, "random")
    a = O.field_diff(O.CD8, 3, fin, 0.1, gL, gR)
    b = O.field_diff(O.CD8, 3, fin, 0.1, gL, gR, use_ref=True)
    # identical away from cells 3 and n-4 of each line, which read across the line ends
    mask = np.ones_like(a, dtype=bool)
    mask[:, :, 3] = False
    mask[:, :, 20 - 4] = False
    assert np.array_equal(a[mask], b[mask])
    assert not np.array_equal(a[:, :, 3], b[:, :, 3])


@pytest.mark.parametrize("stage", [0, 1, 2])
def test_rk3_bitexact(stage):
    rng = np.random.default_rng(stage)
    r, un, ui = (rng.standard_normal((5, 37)) for _ in range(3))
    a = O.rk3(stage, r, un, ui, 0.0137)
    b = O.rk3(stage, r, un, ui, 0.0137, use_ref=True)
    assert np.array_equal(a, b)
