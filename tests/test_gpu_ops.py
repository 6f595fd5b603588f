"""Operator-level parity: the gfx950 operator hooks (same device code as the
fused sweeps) against the oracle, on the reference's Field calling convention.

Tolerance: the GPU rearranges each formula (x*x for pow, one normalising
division, partition + PCR tridiagonal solve instead of Thomas), so results
agree to rounding: |gpu - oracle| <= 1e-12 * max(1, |oracle|).
"""
import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12

SHAPES = [(16, 3, 2), (24, 2, 3), (40, 2, 1), (63, 2, 1), (64, 1, 2), (100, 2, 1), (128, 1, 1), (191, 1, 1),
          (255, 1, 1), (300, 1, 1), (511, 1, 1)]


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


def _close(a, b):
    err = np.abs(a - b) / np.maximum(1.0, np.abs(b))
    return err.max()


@pytest.mark.parametrize("scheme", [G.WENO5JS, G.WENO5Z, G.WENO7JS, G.CRWENO5, G.HCCS, G.WENO3JS, G.UPWIND1ST,
                                    G.HCCS78, G.HCCS77])
@pytest.mark.parametrize("upwind", [1, -1])
@pytest.mark.parametrize("kind", ["smooth", "random"])
def test_reconstruct_matches_oracle(scheme, upwind, kind):
    rng = np.random.default_rng(11 + scheme + upwind)
    ts = O.halo(scheme)
    for (n1, n2, n3) in SHAPES:
        print(f"reconstruct scheme={scheme} upwind={upwind} kind={kind} n1={n1}", flush=True)
        fin, gL, gR = _data(rng, n1, n2, n3, 4, kind)
        ref = O.field_reconstruct(scheme, upwind, ts, fin, gL, gR)
        got = G.op_reconstruct(scheme, upwind, ts, fin, gL, gR)
        assert _close(got, ref) <= TOL, (n1, _close(got, ref))


@pytest.mark.parametrize("scheme", [G.WENO5JS, G.CRWENO5, G.HCCS, G.WENO3JS, G.HCCS78, G.HCCS77])
def test_reconstruct_discontinuous(scheme):
    """Step data drives the WENO weights and the hccs theta switch far from ideal."""
    rng = np.random.default_rng(5 + scheme)
    ts = O.halo(scheme)
    for (n1, n2, n3) in [(32, 4, 2), (128, 2, 1), (256, 1, 1)]:
        fin, gL, gR = _data(rng, n1, n2, n3, 4, "step")
        for up in (1, -1):
            ref = O.field_reconstruct(scheme, up, ts, fin, gL, gR)
            got = G.op_reconstruct(scheme, up, ts, fin, gL, gR)
            assert _close(got, ref) <= 1e-11, (n1, up, _close(got, ref))


@pytest.mark.parametrize("scheme", [G.WENO5JS, G.CRWENO5, G.HCCS78])
def test_reconstruct_zones(scheme):
    """interpolationInner writes only out[ts..n-ts]; BoundaryL/R only the edges."""
    rng = np.random.default_rng(2)
    ts = O.halo(scheme)
    fin, gL, gR = _data(rng, 32, 2, 2, 4, "smooth")
    ref = O.field_reconstruct(scheme, 1, ts, fin, gL, gR)
    sentinel = -7.0
    inner = G.op_reconstruct(scheme, 1, ts, fin, zones=1, out=np.full((2, 2, 33), sentinel))
    assert np.all(inner[:, :, :ts] == sentinel) and np.all(inner[:, :, 32 - ts + 1:] == sentinel)
    assert _close(inner[:, :, ts:32 - ts + 1], ref[:, :, ts:32 - ts + 1]) <= TOL
    left = G.op_reconstruct(scheme, 1, ts, fin, gL=gL, zones=2, out=np.full((2, 2, 33), sentinel))
    assert _close(left[:, :, :ts], ref[:, :, :ts]) <= TOL and np.all(left[:, :, ts:] == sentinel)


@pytest.mark.parametrize("scheme,ts", [(G.CD6, 3), (G.CD6, 4), (G.CD8, 4), (G.CD2, 1), (G.CD2, 2), (G.CD2, 4)])
def test_diff_matches_oracle(scheme, ts):
    rng = np.random.default_rng(9)
    for (n1, n2, n3) in SHAPES[:6]:
        fin, gL, gR = _data(rng, n1, n2, n3, 4, "random")
        ref = O.field_diff(scheme, ts, fin, 0.173, gL, gR)
        got = G.op_diff(scheme, ts, fin, 0.173, gL, gR)
        assert _close(got, ref) <= TOL


def test_cd8_tilesize3_rejected():
    fin = np.ones((1, 1, 20))
    with pytest.raises(G.NucError):
        G.op_diff(G.CD8, 3, fin, 0.1, np.ones((1, 1, 4)), np.ones((1, 1, 4)))


@pytest.mark.parametrize("stage", [0, 1, 2])
def test_rk3_matches_oracle(stage):
    rng = np.random.default_rng(stage)
    r, un, ui = (rng.standard_normal(1000) for _ in range(3))
    ref = O.rk3(stage, r, un, ui, 0.0137)
    got = G.op_rk3(stage, r, un, ui, 0.0137)
    assert _close(got, ref) <= 1e-15
