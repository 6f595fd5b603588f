"""The reference-shaped executable (nuc3d_run: src/NUC3d.cpp:3-14,
src/singleBlock.cpp:31-71) end to end: inp/*.in in; per-rank restart, Tecplot
and averaged-statistics files out every saveStep; enstrophy.dat,
turb_para.dat and the shell spectra files every postStep; a restart (state and running average) from an
intermediate step that reproduces the uninterrupted run bit for bit; the final
state, statistics and averages equal the same steps taken through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "nuc3d-cartisian_amd", "nuc3d_run")


def _write_inp(d, start, end, average_start=0):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "PostProc.in"), "w") as f:
        f.write(f"AverageType 1\nAverageStartStep {average_start}\nSimutaneousData 1\nAveragedSaveType 1\n"
                "SimutaneousSaveType 1\n")
    with open(os.path.join(d, "block.in"), "w") as f:
        f.write("1 2 1\n24 24 24\n6.283185307179586 6.283185307179586 6.283185307179586\n")
    with open(os.path.join(d, "Method.in"), "w") as f:
        f.write("scheme_time RK3\nscheme_x_ivs crweno5\nscheme_y_ivs crweno5\nscheme_z_ivs crweno5\n"
                "scheme_x_vis cd6\nscheme_y_vis cd6\nscheme_z_vis cd6\n")
    with open(os.path.join(d, "PhysModel.in"), "w") as f:
        f.write("Neqs 5\nEosName IdealGas\nRiemannSolver LF\nPDEModel NavierStokes3d\nViscousModel Sutherland\n"
                "Forcing None\nReynolds 400.0\nMach 0.2\n")
    with open(os.path.join(d, "IOController.in"), "w") as f:
        f.write(f"startStep {start}\nendStep {end}\nsaveStep 2\npostStep 1\npostSaveStep 100\nBenchmark 2\n"
                "cfl 0.5\ntl 100.0\ndt 0.01\nBinary yes\nTecplot yes\nPostProc yes\n")


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
    aver = tmp_path / "AveragedFlowData"
    aver.mkdir()
    (tmp_path / "flowDataTemp").mkdir()
    _write_inp(tmp_path / "inp", 0, 4)
    out = _run(tmp_path, str(tmp_path / "inp"))
    assert out.count("========step =") == 4
    names = sorted(os.listdir(flow))
    for step in (2, 4):
        for r in range(2):
            assert f"NUC3d_step_{step}_ID_{r}.bin" in names and f"NUC3d_step_{step}_id_{r}.dat" in names
    assert sorted(os.listdir(aver)) == sorted([f"Averaged_step_{a}_id_{r}.bin" for a in (2, 4) for r in range(2)]
                                              + [f"Averaged_tec_step_{a}_id_{r}.dat" for a in (2, 4) for r in range(2)])
    cfg, _, _ = G.config_from_inp(str(tmp_path / "inp"))
    t4, q4 = G.read_restart(cfg, flow, 4)
    # the same 4 steps through the C ABI, with the statistics and averages of each
    stats = []
    with G.Context(cfg) as ctx:
        ctx.upload(G.init_tgv(cfg))
        ctx.average_begin(1)
        for _ in range(4):
            ctx.step(1)
            ctx.average_update()
            stats.append(ctx.stats())
        ref = ctx.download()
        rho_avg = ctx.average_field(0, 0)
        kin = ctx.spectra_kinetic()
        ens_sp = ctx.spectrum_scalar(G.SPEC_ENSTROPHY)
    assert np.array_equal(q4, ref)
    ens = [list(map(float, l.split())) for l in open(tmp_path / "enstrophy.dat")]
    tur = [list(map(float, l.split())) for l in open(tmp_path / "turb_para.dat")]
    assert [int(r[0]) for r in ens] == [1, 2, 3, 4] and all(len(r) == 18 for r in ens)
    assert all(len(r) == 11 for r in tur)
    for k in range(4):   # 12 significant digits, as written
        assert np.allclose(ens[k][2:], stats[k][:16], rtol=1e-11, atol=1e-300)
        assert np.allclose(tur[k][2:], stats[k][16:25], rtol=1e-11, atol=1e-300)
    # shell spectra of every post step (src/fft.cpp:449-470): "k E E/sum"
    temp = tmp_path / "flowDataTemp"
    for step in range(1, 5):
        for name in ("kin", "kinc", "kins", "kinl", "kinr", "hel", "ens", "den", "pre"):
            assert (temp / f"power_shell_{step}_{name}.dat").exists(), (step, name)
    for name, e in (("kin", kin[0]), ("kins", kin[2]), ("kinr", kin[4]), ("ens", ens_sp)):
        rows = np.array([list(map(float, l.split())) for l in open(temp / f"power_shell_4_{name}.dat") if l.strip()])
        assert rows.shape == (12, 3) and np.array_equal(rows[:, 0], np.arange(12))
        assert np.allclose(rows[:, 1], e, rtol=1e-11, atol=1e-300), name
        assert np.allclose(rows[:, 2], e / e.sum(), rtol=1e-11, atol=1e-300), name
    # solveQ's per-point fields as Tecplot (postproc::OutputTemporal) at every save
    for step in (2, 4):
        for r in range(2):
            head = open(temp / f"Q_tec_step_{step}_id_{r}.dat").readline()
            assert head.startswith("title=time_")
    # the averaged file at step 4: [int averStep][10 primaries][21 gradients] per rank block
    raw = np.fromfile(aver / "Averaged_step_4_id_1.bin", dtype=np.uint8)
    assert int(raw[:4].view(np.int32)[0]) == 4
    blk = raw[4:].view(np.float64).reshape(31, 24, 12, 24)[0]           # rho of rank 1 (cy = 1)
    assert np.array_equal(blk, rho_avg.reshape(24, 24, 24)[:, 12:, :])
    keep_avg = {r: (aver / f"Averaged_step_4_id_{r}.bin").read_bytes() for r in range(2)}
    # restart at step 2 (inputQ_binary, src/block.cpp:640-671) and run to 4 again
    keep = {r: (flow / f"NUC3d_step_4_ID_{r}.bin").read_bytes() for r in range(2)}
    for r in range(2):
        os.remove(flow / f"NUC3d_step_4_ID_{r}.bin")
        os.remove(aver / f"Averaged_step_4_id_{r}.bin")
    _write_inp(tmp_path / "inp2", 2, 4, average_start=2)
    out = _run(tmp_path, str(tmp_path / "inp2"))
    assert out.count("========step =") == 2
    for r in range(2):
        assert (flow / f"NUC3d_step_4_ID_{r}.bin").read_bytes() == keep[r]
        assert (aver / f"Averaged_step_4_id_{r}.bin").read_bytes() == keep_avg[r]
    lines = open(tmp_path / "enstrophy.dat").read().splitlines()
    assert lines[4:] == lines[2:4]   # steps 3 and 4 again, the same text


def test_runner_precision_option(tmp_path):
    """--precision selects the context's precision (the inp files carry none;
    FP64 is the default): after 4 steps the mixed and FP32 runs' restart files
    hold states within 1e-5 of the FP64 run's (relative L-infinity per variable,
    momentum against its largest component), and an unknown name is refused."""
    if not os.path.exists(RUNNER):
        pytest.fail("nuc3d_run is not built (make -C nuc3d-cartisian_amd)")
    states = {}
    for prec in ("fp64", "mixed", "fp32"):
        d = tmp_path / prec
        for sub in ("flowData", "AveragedFlowData", "flowDataTemp"):
            (d / sub).mkdir(parents=True)
        _write_inp(d / "inp", 0, 4)
        print(f"nuc3d_run --precision {prec}", flush=True)
        r = subprocess.run([RUNNER, str(d / "inp"), "--precision", prec], cwd=d, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        cfg, _, _ = G.config_from_inp(str(d / "inp"))
        states[prec] = G.read_restart(cfg, d / "flowData", 4)[1]
    N = 24 ** 3
    a = states["fp64"].reshape(5, N)
    # the Taylor-Green field starts with w = 0: momentum errors are measured
    # against the largest momentum component
    scale = np.abs(a).max(axis=1)
    scale[1:4] = np.abs(a[1:4]).max()
    for prec in ("mixed", "fp32"):
        b = states[prec].reshape(5, N)
        err = np.abs(a - b).max(axis=1) / scale
        print(prec, "vs fp64", err, flush=True)
        assert np.all(err <= 1e-5) and not np.array_equal(a, b), err
    r = subprocess.run([RUNNER, str(tmp_path / "fp64" / "inp"), "--precision", "fp16"], cwd=tmp_path / "fp64",
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "--precision" in r.stderr
