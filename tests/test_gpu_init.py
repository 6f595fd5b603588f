"""Initial conditions generated on the device (nuc_init_tgv_device,
nuc_init_hit_device, nuc_init_ivc_device, nuc_init_default_device) against the
host generators and the oracle: the same fields up to rounding of the
transcendental functions and of the synthesis sums, and the same solution after
stepping."""
import time

import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


def _download_after(cfg, init):
    with G.Context(cfg) as ctx:
        init(ctx)
        return ctx.download()


@pytest.mark.parametrize("n,npr,origin", [((32, 24, 16), (2, 2, 1), (0.0, 0.0, 0.0)),
                                          ((20, 20, 20), (1, 1, 1), (0.0, 0.0, 0.0)),
                                          ((20, 16, 12), (2, 1, 1), (0.4, -1.1, 2.5))], ids=str)
def test_tgv_device_equals_host(n, npr, origin):
    cfg = G.make_config(n=n, npr=npr, mach=0.1, origin=origin)
    host = G.init_tgv(cfg)
    if any(origin):
        oc = O.make_config(n=n, npr=npr, mach=0.1, origin=origin)
        np.testing.assert_array_equal(host, O.init_tgv(oc))
        assert not np.array_equal(host, G.init_tgv(G.make_config(n=n, npr=npr, mach=0.1)))
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


@pytest.mark.parametrize("npr,origin", [((1, 1, 1), (0.0, 0.0, 0.0)), ((2, 2, 1), (0.0, 0.0, 0.0)),
                                        ((2, 1, 1), (-1.5, 0.75, 0.0))], ids=str)
def test_ivc_and_default_device_equal_host_and_oracle(npr, origin):
    kw = dict(n=(40, 32, 8), L=(10.0, 10.0, 2.5), npr=npr, mach=0.3, origin=origin)
    cfg, oc = G.make_config(**kw), O.make_config(**kw)
    N = 40 * 32 * 8
    for name, host, orc, init in (("ivc", G.init_ivc, O.init_ivc, lambda c: c.init_ivc_device()),
                                  ("default", G.init_default, O.init_default, lambda c: c.init_default_device())):
        dev = _download_after(cfg, init).reshape(5, N)
        for ref in (host(cfg).reshape(5, N), orc(oc).reshape(5, N)):
            mom = np.abs(ref[1:4]).max() or np.abs(ref).max()   # the default field's momenta are 0
            scale = np.maximum(np.abs(ref).max(axis=1), mom)
            err = np.abs(dev - ref).max(axis=1) / scale
            print(f"{name} {npr} device vs reference", err, flush=True)
            assert np.all(err <= 1e-14), (name, err)


def test_isentropic_vortex_ten_steps_match_oracle():
    """The reference's Benchmark 1 case (Euler, isentropic vortex on a 10 x 10
    box) from the device initial state, against the oracle from its own."""
    kw = dict(n=(32, 32, 8), L=(10.0, 10.0, 2.5), npr=(2, 1, 1), scheme=G.WENO5JS, vis=G.CD6,
              riemann=G.AUSM, model=G.EULER3D, cfl=0.5)
    cfg, oc = G.make_config(**kw), O.make_config(**kw)
    s = O.Solver(oc)
    s.set_state(O.init_ivc(oc))
    s.step(10)
    ref = s.get_state().reshape(5, -1)
    with G.Context(cfg) as ctx:
        ctx.init_ivc_device()
        ctx.step(10)
        got = ctx.download().reshape(5, -1)
    scale = np.maximum(np.abs(ref).max(axis=1), np.abs(ref[1:4]).max())
    err = np.abs(got - ref).max(axis=1) / scale
    print("isentropic vortex, 10 steps, device vs oracle", err, flush=True)
    assert np.all(err <= 1e-12), err


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
