"""The reference-shaped host classes (nuc3d_host.hpp: fieldOperator3d and its
scheme classes, tvdrk3rd, PDEData3d, NaiverStokesData3d / EulerData3D), driven
from C++ the way the reference's call sites drive them (src/block.cpp:191-215,
src/fieldOperator.cpp:387-639) and checked against the oracle: the test
program tests/cpp/test_host_classes.cpp, built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_host_classes")


def test_host_classes_against_oracle():
    assert os.path.exists(BIN), "tests/cpp/test_host_classes is not built (__graft_entry__.build())"
    print("host-class test program", flush=True)
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden", "inp_operators")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "PASS" in r.stdout
    # the buffer sizes of hccs78 / weno3js / upwind1st (src/fieldOperator.cpp:128-220)
    assert "bufferSize 4 2 1" in r.stdout
