"""The oracle against the committed vectors of the reference's own scheme
objects (tests/golden/ref_vectors.npz, written by tests/golden/make_ref_vectors.py
from oracle/_ref, the reference's sources compiled unmodified).  Runs without
the reference tree; bit-exact like test_oracle_vs_reference.py."""
import os

import numpy as np

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


def _load():
    return np.load(GOLD, allow_pickle=False)


def test_reconstruction_vectors_bitexact():
    z = _load()
    keys = sorted({k.split("_")[1] for k in z.files if k.startswith("rec_")}, key=int)
    assert len(keys) == 96
    for k in keys:
        scheme, upwind, ts = (int(v) for v in z[f"rec_{k}_meta"])
        got = O.field_reconstruct(scheme, upwind, ts, z[f"rec_{k}_in"], z[f"rec_{k}_gL"], z[f"rec_{k}_gR"])
        assert np.array_equal(got, z[f"rec_{k}_out"]), (k, scheme, upwind)


def test_difference_vectors_bitexact():
    z = _load()
    keys = sorted({k.split("_")[1] for k in z.files if k.startswith("dif_")}, key=int)
    assert len(keys) == 12
    for k in keys:
        scheme, ts = (int(v) for v in z[f"dif_{k}_meta"])
        got = O.field_diff(scheme, ts, z[f"dif_{k}_in"], 0.173, z[f"dif_{k}_gL"], z[f"dif_{k}_gR"])
        assert np.array_equal(got, z[f"dif_{k}_out"]), (k, scheme, ts)


def test_rk3_vectors_bitexact():
    z = _load()
    for stage in range(3):
        r, un, ui = z[f"rk3_{stage}_in"]
        assert np.array_equal(O.rk3(stage, r, un, ui, 0.0137), z[f"rk3_{stage}_out"])
