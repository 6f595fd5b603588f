"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol its headers declare, the inp/*.in readers follow the reference's parsing
rules, footprints, and the serving loop's handling of a job that cannot start.

The inp fixtures under tests/golden/ are written in the reference's formats:
block.in (src/block.cpp:49-61), Method.in (src/fieldOperator.cpp:21-125),
PhysModel.in (six positional pairs, then key/value; src/physicsModel.cpp:72-156)
and IOController.in (src/IOcontroller.cpp:12-110).
"""
import ctypes
import os
import re
import time

import numpy as np
import pytest

import nuc3d_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared():
    names = set()
    for h in ("nuc3d_gpu.h", "nuc3d_serve.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^(?:const\s+)?(?:int|long long|double|char|void)\s*\*?\s*(nuc_\w+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    declared = _declared()
    assert len(declared) >= 30
    assert declared == set(G.EXPORTS), declared.symmetric_difference(G.EXPORTS)
    L = G.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_inp_tgv64():
    c, s, e = G.config_from_inp(os.path.join(GOLD, "inp_tgv64"))
    assert (c.npx, c.npy, c.npz, c.nx, c.ny, c.nz) == (1, 1, 1, 64, 64, 64)
    assert abs(c.lx - 2 * np.pi) < 1e-15
    assert list(c.scheme) == [G.WENO5JS] * 3 and list(c.vis_scheme) == [G.CD6] * 3
    assert (c.riemann, c.model, c.viscosity, c.forcing) == (G.LF, G.NAVIERSTOKES3D, G.SUTHERLAND, G.FORCING_NONE)
    assert (c.reynolds, c.mach, c.prandtl, c.gamma) == (1600.0, 0.1, 0.72, 1.4)
    assert (s, e, c.cfl) == (0, 200, 0.5)
    assert c.precision == G.FP64


def test_inp_sixth_positional_pair_is_the_forcing_slot():
    """PhysModel.in's first six lines are read positionally; a file whose sixth
    pair is 'Reynolds 1600.0' (no Forcing line) loses that Reynolds number: the
    value goes to the forcing slot (unknown name -> None) and Reynolds keeps its
    default 1000 (src/physicsModel.cpp:22-43,72-77,104-110)."""
    c, s, e = G.config_from_inp(os.path.join(GOLD, "inp_reynolds_as_sixth_pair"))
    assert c.forcing == G.FORCING_NONE
    assert c.reynolds == 1000.0
    assert (c.mach, c.prandtl) == (0.5, 0.7)
    assert (c.riemann, c.model) == (G.AUSM, G.EULER3D)
    assert (c.npx, c.npy, c.npz, c.nx, c.ny, c.nz) == (2, 1, 2, 32, 48, 64)
    assert (c.lx, c.ly, c.lz) == (6.0, 7.0, 8.0)
    assert list(c.scheme) == [G.WENO5Z, G.CRWENO5, G.WENO7JS]
    # scheme_y_vis absent -> default cd6 (src/fieldOperator.cpp:22-33)
    assert list(c.vis_scheme) == [G.CD8, G.CD6, G.CD8]
    assert (s, e, c.cfl) == (5, 1000, 0.6)


def test_inp_forced_hccs():
    c, s, e = G.config_from_inp(os.path.join(GOLD, "inp_forced_hccs"))
    assert list(c.scheme) == [G.HCCS] * 3
    assert (c.riemann, c.viscosity, c.forcing) == (G.AUSM, G.CONSTANT, G.FORCING_CONST)
    assert (c.forcing_amp, c.forcing_amp_hel, c.forcing_k) == (0.5, 0.0, 2.0)
    assert c.reynolds == 800.0
    assert (c.nx, c.npx) == (512, 4)


def test_inp_remaining_operators():
    """hccs78, weno3js, upwind1st, cd2 and AUSMp as Method.in / PhysModel.in
    name them (src/fieldOperator.cpp:128-345, src/physicsModel.cpp:16-20)."""
    c, s, e = G.config_from_inp(os.path.join(GOLD, "inp_operators"))
    assert list(c.scheme) == [G.HCCS78, G.WENO3JS, G.UPWIND1ST]
    assert list(c.vis_scheme) == [G.CD2, G.CD8, G.CD2]
    assert c.riemann == G.AUSMP
    assert (c.nx, c.ny, c.nz, c.npy, c.reynolds, c.mach) == (48, 32, 40, 2, 900.0, 0.4)


def test_inp_missing_directory_is_an_error():
    with pytest.raises(G.NucError):
        G.config_from_inp(os.path.join(GOLD, "does_not_exist"))


def test_inp_postproc_average_type_is_checked(tmp_path):
    """PostProc.in's first value is TimeAverageType 0..3; any other is the
    reference's "No such postproc" error (src/postproc.cpp:85-154).  Without
    the file no post-processing is configured."""
    import shutil
    d = tmp_path / "inp"
    shutil.copytree(os.path.join(GOLD, "inp_tgv64"), d)
    G.config_from_inp(str(d))                                   # no PostProc.in
    (d / "PostProc.in").write_text("AverageType 3\nAverageStartStep 0\nSimutaneousData 1\n")
    G.config_from_inp(str(d))
    (d / "PostProc.in").write_text("AverageType 4\nAverageStartStep 0\n")
    with pytest.raises(G.NucError):
        G.config_from_inp(str(d))


def test_footprint_scales_with_cells_and_options():
    a = G.footprint(G.make_config(n=(64, 64, 64)))
    b = G.footprint(G.make_config(n=(128, 64, 64)))
    assert 1.9 < b / a < 2.1
    # 29 FP64 fields per cell for Navier-Stokes
    assert abs(a / 64 ** 3 / 8 - 29) < 1.5
    e = G.footprint(G.make_config(n=(64, 64, 64), model=G.EULER3D))
    assert e < a
    f = G.footprint(G.make_config(n=(64, 64, 64), forcing=G.FORCING_CONST, forcing_k=2.0))
    assert f > a


def test_server_reports_a_job_that_cannot_start_and_keeps_serving():
    """A job whose configuration nuc_create rejects (grid not divisible by the
    rank grid) is reported FAILED; the loop stays up and can be destroyed."""
    bad = G.make_config(n=(18, 16, 16), npr=(4, 1, 1))
    q = np.ones(5 * 18 * 16 * 16)
    with G.Server(hbm_budget=1 << 30, max_active=2) as srv:
        jid = srv.submit(bad, q, 3)
        t0 = time.time()
        res = None
        while res is None and time.time() - t0 < 30:
            res = srv.result(jid)
            time.sleep(0.01)
        assert res is not None
        state, r = res
        assert state is None and r.state == G.JOB_FAILED and r.error != 0
        st = srv.stats()
        assert st["failed"] == 1 and st["active"] == 0 and st["bytes_in_use"] == 0


def test_server_rejects_job_above_budget():
    cfg = G.make_config(n=(64, 64, 64))
    with G.Server(hbm_budget=G.footprint(cfg) - 1, max_active=2) as srv:
        with pytest.raises(G.NucError):
            srv.submit(cfg, np.ones(5 * 64 ** 3), 1)


def test_shu_osher_initial_condition():
    """nuc_init_shu_osher: the two states on either side of x = -4 (box centred
    on 0), a divergence-free perturbation of the requested rms, and the seed
    fixing the field."""
    nx, ny, nz = 64, 16, 16
    cfg = G.make_config(n=(nx, ny, nz), L=(10.0, 2.5, 2.5), mach=1.0 / np.sqrt(1.4))
    q0 = G.init_shu_osher(cfg, pert=0.0).reshape(5, nz, ny, nx)
    x = (np.arange(nx) + 0.5) * 10.0 / nx - 5.0
    left = x < -4.0
    assert np.allclose(q0[0][..., left], 3.857143) and np.allclose(q0[1][..., left], 3.857143 * 2.629369)
    assert np.allclose(q0[0][..., ~left], 1.0 + 0.2 * np.sin(5.0 * x[~left]))
    assert np.all(q0[1][..., ~left] == 0) and np.all(q0[2] == 0) and np.all(q0[3] == 0)
    assert np.allclose(q0[4][..., ~left], 1.0 / 0.4)
    assert np.allclose(q0[4][..., left], 10.33333 / 0.4 + 0.5 * 3.857143 * 2.629369 ** 2)
    q = G.init_shu_osher(cfg, pert=0.05, seed=7).reshape(5, nz, ny, nx)
    assert np.array_equal(q, G.init_shu_osher(cfg, pert=0.05, seed=7).reshape(5, nz, ny, nx))
    assert not np.array_equal(q, G.init_shu_osher(cfg, pert=0.05, seed=8).reshape(5, nz, ny, nx))
    assert np.array_equal(q[0], q0[0])
    du = [q[1 + c] / q[0] - q0[1 + c] / q0[0] for c in range(3)]
    rms = np.sqrt(np.mean(du[0] ** 2 + du[1] ** 2 + du[2] ** 2) / 3.0)
    assert abs(rms - 0.05 * 2.629369) < 1e-12
    # spectral divergence of the perturbation vanishes (arrays stored z, y, x)
    L = (2.5, 2.5, 10.0)
    k = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(n, d=Ld / n) for n, Ld in zip((nz, ny, nx), L)], indexing="ij")
    div = k[2] * np.fft.fftn(du[0]) + k[1] * np.fft.fftn(du[1]) + k[0] * np.fft.fftn(du[2])
    mag = np.sqrt(sum(np.abs(k[a] * np.fft.fftn(du[2 - a])) ** 2 for a in range(3)))
    assert np.abs(div).max() <= 1e-10 * mag.max()
    with pytest.raises(G.NucError):
        G.init_shu_osher(cfg, pert=-1.0)


def test_inp_abc_forcing_and_rotating_frame_sources():
    """Forcing ABC and the Coriolis/Centrifugal switches with their parameters
    (src/physicsModel.cpp:22-56,133-156); any value but "no" switches a source on."""
    c, s, e = G.config_from_inp(os.path.join(GOLD, "inp_abc_rotating"))
    assert c.forcing == G.FORCING_ABC and (c.forcing_amp, c.forcing_k) == (0.05, 1.0)
    assert (c.coriolis, c.centrifugal) == (1, 1)
    assert (c.rossby, c.x_center, c.y_center, c.well_rad) == (0.5, 3.0, 3.2, 1.5)
    assert (s, e, c.cfl) == (0, 50, 0.4)
    c2, _, _ = G.config_from_inp(os.path.join(GOLD, "inp_tgv64"))
    assert (c2.coriolis, c2.centrifugal, c2.rossby, c2.well_rad) == (0, 0, 0.0, 1.0)


def test_create_failure_message_without_a_context():
    """nuc_create reports why it refused a configuration through
    nuc_last_error(NULL) on the calling thread (validation runs before any
    device call), and a successful-looking empty message otherwise."""
    L = G.lib()
    bad = G.make_config(n=(18, 16, 16), npr=(4, 1, 1))
    h = ctypes.c_void_p()
    rc = L.nuc_create(ctypes.byref(bad), ctypes.byref(h))
    assert rc == 1 and not h.value
    assert "divide evenly by the rank grid" in L.nuc_last_error(None).decode()
    bad2 = G.make_config(n=(16, 16, 16), scheme=42)
    assert L.nuc_create(ctypes.byref(bad2), ctypes.byref(h)) == 1
    assert "unknown inviscid scheme" in L.nuc_last_error(None).decode()
    with pytest.raises(G.NucError, match="divide evenly"):
        G.Context(bad)


def test_fp32_rejects_what_it_does_not_take():
    """NUC_FP32 refuses hccs/hccs77 (their theta blend needs FP64 stencils),
    forcing and rotating-frame sources, and the caller-buffer RHS entry point
    (whose buffers are FP64), with NUC_EINVAL before any device call; the
    footprint counts FP32 state/RHS/viscous fields and the FP64 view."""
    L = G.lib()
    h = ctypes.c_void_p()
    bad = [dict(scheme=G.HCCS), dict(scheme=(G.WENO5JS, G.HCCS77, G.WENO5JS)),
           dict(forcing=G.FORCING_CONST, forcing_k=2.0), dict(forcing=G.FORCING_ABC),
           dict(forcing=G.FORCING_TG), dict(coriolis=True, rossby=1.0)]
    for kw in bad:
        cfg = G.make_config(n=(16, 16, 16), precision=G.FP32, **kw)
        assert L.nuc_create(ctypes.byref(cfg), ctypes.byref(h)) == 1, kw
        assert "NUC_FP32" in L.nuc_last_error(None).decode(), kw
        assert not h.value
    ok = G.make_config(n=(16, 16, 16), precision=G.FP32)
    buf = np.zeros(5 * 16 ** 3)
    rc = L.nuc_rhs_noalloc(ctypes.byref(ok), buf.ctypes.data_as(ctypes.c_void_p),
                           buf.ctypes.data_as(ctypes.c_void_p), None, None)
    assert rc == 1 and "NUC_FP32" in L.nuc_last_error(None).decode()
    n = 64 ** 3
    f64, fmx, f32 = (G.footprint(G.make_config(n=(64, 64, 64), precision=p)) for p in (G.FP64, G.MIXED, G.FP32))
    # per cell: FP64 20 x 8 + 9 x 8, mixed 20 x 8 + 9 x 4, FP32 20 x 4 + 9 x 4 + 5 x 8 (+ small tables)
    assert abs((f64 - fmx) / n - 36) < 0.01 and abs((fmx - f32) / n - 40) < 0.01


def test_long_line_rules():
    """Lines above 512 cells are swept in segments: explicit schemes need a
    divisor of n in [64, 255], compact schemes rank blocks of at most 511
    cells; a box holds at most 2^29 cells (refused with NUC_EINVAL before any
    device call)."""
    L = G.lib()
    h = ctypes.c_void_p()
    bad = [(dict(n=(1031, 12, 12)), "divisor in [64, 255]"),
           (dict(n=(1024, 12, 12), scheme=G.CRWENO5), "rank blocks of at most 511"),
           (dict(n=(1024, 12, 12), npr=(2, 1, 1), scheme=G.HCCS), "rank blocks of at most 511"),
           (dict(n=(1024, 1024, 1024)), "2^29 cells")]
    for kw, msg in bad:
        cfg = G.make_config(**kw)
        assert L.nuc_create(ctypes.byref(cfg), ctypes.byref(h)) == 1, kw
        assert msg in L.nuc_last_error(None).decode(), (kw, L.nuc_last_error(None))
        assert not h.value
