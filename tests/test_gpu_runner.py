"""The reference-shaped executable (nuc3d_run: src/NUC3d.cpp:3-14,
src/singleBlock.cpp:31-71) end to end: inp/*.in in, per-rank restart and
Tecplot files out every saveStep, and a restart from an intermediate step that
reproduces the uninterrupted run bit for bit; the final state equals the same
steps taken through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "nuc3d-cartisian_amd", "nuc3d_run")


def _write_inp(d, start, end):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "block.in"), "w") as f:
        f.write("1 2 1\n24 24 24\n6.283185307179586 6.283185307179586 6.283185307179586\n")
    with open(os.path.join(d, "Method.in"), "w") as f:
        f.write("scheme_time RK3\nscheme_x_ivs crweno5\nscheme_y_ivs crweno5\nscheme_z_ivs crweno5\n"
                "scheme_x_vis cd6\nscheme_y_vis cd6\nscheme_z_vis cd6\n")
    with open(os.path.join(d, "PhysModel.in"), "w") as f:
        f.write("Neqs 5\nEosName IdealGas\nRiemannSolver LF\nPDEModel NavierStokes3d\nViscousModel Sutherland\n"
                "Forcing None\nReynolds 400.0\nMach 0.2\n")
    with open(os.path.join(d, "IOController.in"), "w") as f:
        f.write(f"startStep {start}\nendStep {end}\nsaveStep 2\npostStep 100\npostSaveStep 100\nBenchmark 2\n"
                "cfl 0.5\ntl 100.0\ndt 0.01\nBinary yes\nTecplot yes\nPostProc no\n")


def _run(cwd, inp):
    print(f"nuc3d_run {inp}", flush=True)
    r = subprocess.run([RUNNER, inp], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_runner_restart_reproduces_the_uninterrupted_run(tmp_path):
    if not os.path.exists(RUNNER):
        pytest.fail("nuc3d_run is not built (make -C nuc3d-cartisian_amd)")
    flow = tmp_path / "flowData"
    flow.mkdir()
    _write_inp(tmp_path / "inp", 0, 4)
    out = _run(tmp_path, str(tmp_path / "inp"))
    assert out.count("========step =") == 4
    names = sorted(os.listdir(flow))
    for step in (2, 4):
        for r in range(2):
            assert f"NUC3d_step_{step}_ID_{r}.bin" in names and f"NUC3d_step_{step}_id_{r}.dat" in names
    cfg, _, _ = G.config_from_inp(str(tmp_path / "inp"))
    t4, q4 = G.read_restart(cfg, flow, 4)
    # the same 4 steps through the C ABI
    with G.Context(cfg) as ctx:
        ctx.upload(G.init_tgv(cfg))
        ctx.step(4)
        ref = ctx.download()
    assert np.array_equal(q4, ref)
    # restart at step 2 (inputQ_binary, src/block.cpp:640-671) and run to 4 again
    keep = {r: (flow / f"NUC3d_step_4_ID_{r}.bin").read_bytes() for r in range(2)}
    for r in range(2):
        os.remove(flow / f"NUC3d_step_4_ID_{r}.bin")
    _write_inp(tmp_path / "inp2", 2, 4)
    out = _run(tmp_path, str(tmp_path / "inp2"))
    assert out.count("========step =") == 2
    for r in range(2):
        assert (flow / f"NUC3d_step_4_ID_{r}.bin").read_bytes() == keep[r]
