"""The reference's file formats, per rank of its rank grid (host code only):
binary restart files (block::outputQ_binary / inputQ_binary,
src/block.cpp:609-671) and Tecplot output (block::outputQ_tecplot,
src/block.cpp:553-607).

The bytes are checked against the format read independently with numpy: a
rank's file is [double time][5 fields of its block, x fastest], rank
(cx, cy, cz) = (cz*npy + cy)*npx + cx (MPI's row-major numbering of the
reference's Cartesian communicator with dims (npz, npy, npx),
src/MPICommunicator.cpp:26-33).
"""
import os

import numpy as np
import pytest

import nuc3d_amd as G


def _cfg(n, npr):
    return G.make_config(n=n, npr=npr, L=(6.0, 5.0, 4.0), mach=0.3)


def _state(cfg, seed):
    rng = np.random.default_rng(seed)
    N = cfg.nx * cfg.ny * cfg.nz
    q = np.empty((5, N))
    q[0] = 1.0 + 0.1 * rng.random(N)
    q[1:4] = 0.3 * rng.standard_normal((3, N))
    q[4] = 20.0 + rng.random(N)
    return q.ravel()


@pytest.mark.parametrize("npr", [(1, 1, 1), (2, 1, 1), (2, 3, 2)], ids=str)
def test_restart_files_per_rank(tmp_path, npr):
    n = (12, 12, 8)
    cfg = _cfg(n, npr)
    q = _state(cfg, 1)
    G.write_restart(cfg, tmp_path, 40, 1.25, q)
    nranks = npr[0] * npr[1] * npr[2]
    assert sorted(os.listdir(tmp_path)) == sorted(f"NUC3d_step_40_ID_{r}.bin" for r in range(nranks))
    Q = q.reshape(5, n[2], n[1], n[0])
    nl = [n[d] // npr[d] for d in range(3)]
    for r in range(nranks):
        cx, cy, cz = r % npr[0], (r // npr[0]) % npr[1], r // (npr[0] * npr[1])
        raw = np.fromfile(tmp_path / f"NUC3d_step_40_ID_{r}.bin", dtype=np.float64)
        assert raw.size == 1 + 5 * nl[0] * nl[1] * nl[2]
        assert raw[0] == 1.25
        blk = Q[:, cz * nl[2]:(cz + 1) * nl[2], cy * nl[1]:(cy + 1) * nl[1], cx * nl[0]:(cx + 1) * nl[0]]
        assert np.array_equal(raw[1:], blk.ravel())
    t, back = G.read_restart(cfg, tmp_path, 40)
    assert t == 1.25 and np.array_equal(back, q)


def test_single_rank_restart_is_time_then_state(tmp_path):
    cfg = _cfg((10, 8, 6), (1, 1, 1))
    q = _state(cfg, 2)
    G.write_restart(cfg, tmp_path, 0, 0.5, q)
    raw = (tmp_path / "NUC3d_step_0_ID_0.bin").read_bytes()
    assert raw == np.float64(0.5).tobytes() + q.tobytes()


def test_restart_read_errors(tmp_path):
    cfg = _cfg((12, 12, 8), (2, 1, 1))
    q = _state(cfg, 3)
    G.write_restart(cfg, tmp_path, 5, 0.0, q)
    with pytest.raises(G.NucError):
        G.read_restart(cfg, tmp_path, 6)                          # no such step
    with pytest.raises(G.NucError):
        G.read_restart(_cfg((12, 12, 10), (2, 1, 1)), tmp_path, 5)  # files too short for this grid
    os.remove(tmp_path / "NUC3d_step_5_ID_1.bin")
    with pytest.raises(G.NucError):
        G.read_restart(cfg, tmp_path, 5)                          # a rank's file missing


def _read_tecplot(path):
    lines = open(path).read().split("\n")
    return lines[:4], np.array([float(x) for x in lines[4:] if x.strip()])


@pytest.mark.parametrize("npr", [(1, 1, 1), (2, 2, 1)], ids=str)
def test_tecplot_files(tmp_path, npr):
    n = (8, 6, 4)
    cfg = _cfg(n, npr)
    q = _state(cfg, 4)
    origin = np.array([0.5, -1.0, 0.25])
    G.write_tecplot(cfg, tmp_path, 7, q, origin=origin)
    nl = [n[d] // npr[d] for d in range(3)]
    Q = q.reshape(5, n[2], n[1], n[0])
    L = (6.0, 5.0, 4.0)
    for r in range(npr[0] * npr[1] * npr[2]):
        cx, cy, cz = r % npr[0], (r // npr[0]) % npr[1], r // (npr[0] * npr[1])
        head, vals = _read_tecplot(tmp_path / f"NUC3d_step_7_id_{r}.dat")
        assert head[0] == "title=NUC3d"
        assert head[1] == "variables=x,y,z,rho,u,v,w,pressure,temperature,e,sound_speed"
        assert head[2] == f" Zone I = {nl[0] + 1}, J= {nl[1] + 1}, K={nl[2] + 1}"
        assert head[3] == " DATAPACKING=BLOCK, VARLOCATION=([4-11]=CELLCENTERED)"
        nn = (nl[0] + 1) * (nl[1] + 1) * (nl[2] + 1)
        nc = nl[0] * nl[1] * nl[2]
        assert vals.size == 3 * nn + 8 * nc
        # node coordinates, x fastest
        k, j, i = np.meshgrid(np.arange(nl[2] + 1), np.arange(nl[1] + 1), np.arange(nl[0] + 1), indexing="ij")
        for d, (idx, c) in enumerate(((i, cx), (j, cy), (k, cz))):
            x = L[d] / npr[d] * c + L[d] / n[d] * idx - origin[d]
            assert np.allclose(vals[d * nn:(d + 1) * nn], x.ravel(), rtol=1e-11, atol=1e-11)
        # cell-centred primitives from the conserved block
        blk = Q[:, cz * nl[2]:(cz + 1) * nl[2], cy * nl[1]:(cy + 1) * nl[1], cx * nl[0]:(cx + 1) * nl[0]]
        blk = blk.reshape(5, -1)
        rho = blk[0]
        u, v, w = blk[1] / rho, blk[2] / rho, blk[3] / rho
        p = np.abs((blk[4] - 0.5 * rho * (u * u + v * v + w * w)) * 0.4)
        expect = [rho, u, v, w, p, 1.4 * 0.09 * p / rho, p / rho / 0.4, np.sqrt(1.4 * p / rho)]
        got = vals[3 * nn:].reshape(8, nc)
        for f in range(8):
            assert np.allclose(got[f], expect[f], rtol=1e-11, atol=1e-12), f   # 12 significant digits

