"""Initial conditions generated on the device (nuc_init_tgv_device,
nuc_init_hit_device) against the host generators: the same fields up to
rounding of the transcendental functions and of the synthesis sums, and the
same solution after stepping."""
import time

import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu


def _download_after(cfg, init):
    with G.Context(cfg) as ctx:
        init(ctx)
        return ctx.download()


@pytest.mark.parametrize("n,npr", [((32, 24, 16), (2, 2, 1)), ((20, 20, 20), (1, 1, 1))], ids=str)
def test_tgv_device_equals_host(n, npr):
    cfg = G.make_config(n=n, npr=npr, mach=0.1)
    host = G.init_tgv(cfg)
    dev = _download_after(cfg, lambda c: c.init_tgv_device())
    N = n[0] * n[1] * n[2]
    a, b = dev.reshape(5, N), host.reshape(5, N)
    scale = np.maximum(np.abs(b).max(axis=1), np.abs(b[1:4]).max())
    err = np.abs(a - b).max(axis=1) / scale
    print("TGV device vs host", err, flush=True)
    assert np.all(err <= 1e-14), err


@pytest.mark.parametrize("n,L,k0", [((32, 32, 32), None, 3.0), ((48, 24, 16), (6.0, 3.0, 2.0), 4.0),
                                    ((96, 96, 96), None, 4.0)], ids=str)
def test_hit_device_equals_host(n, L, k0):
    cfg = G.make_config(n=n, L=L, mach=0.2)
    host = G.init_hit(cfg, k0=k0, u_rms=0.7, seed=77)
    dev = _download_after(cfg, lambda c: c.init_hit_device(k0=k0, u_rms=0.7, seed=77))
    N = n[0] * n[1] * n[2]
    a, b = dev.reshape(5, N), host.reshape(5, N)
    scale = np.maximum(np.abs(b).max(axis=1), np.abs(b[1:4]).max())
    err = np.abs(a - b).max(axis=1) / scale
    print(f"HIT {n} device vs host", err, flush=True)
    assert np.all(err <= 1e-13), err


def test_stepping_from_device_initial_state():
    cfg = G.make_config(n=(32, 32, 32), scheme=G.CRWENO5, vis=G.CD8, mach=0.2, reynolds=500.0)
    q = G.init_hit(cfg, k0=3.0, u_rms=0.5, seed=8)
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        ctx.step(3)
        ref = ctx.download()
    with G.Context(cfg) as ctx:
        ctx.init_hit_device(k0=3.0, u_rms=0.5, seed=8)
        ctx.step(3)
        got = ctx.download()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("3 steps from device vs host initial state", err, flush=True)
    assert err <= 1e-12


def test_large_box_device_initialisation_time():
    cfg = G.make_config(n=(256, 256, 256), mach=0.3 / np.sqrt(3.0))
    with G.Context(cfg) as ctx:
        t0 = time.perf_counter()
        ctx.init_hit_device(k0=4.0, u_rms=1.0, seed=20261202)
        dt = time.perf_counter() - t0
        st = ctx.stats()
    print(f"256^3 decaying-turbulence field on the device: {dt * 1e3:.1f} ms, u_rms {st[16]:.6f}", flush=True)
    assert abs(st[16] - 1.0) < 1e-12
