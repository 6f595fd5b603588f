"""nuc_rhs_noalloc: one RHS evaluation on caller-owned device buffers with the
library's fixed scratch array instead of any device allocation.  Its result
must equal nuc_rhs (the context entry point, NaiverStokesData3d::solve,
src/NaiverStokes3d.cpp:37-73) on the same state bit for bit: both run the same
kernels on the same inputs.  Two calls in a row, the second on a smaller grid,
check that nothing of the first call's scratch contents leaks into the second.
"""
import ctypes

import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


def _state(cfg_kw, seed):
    oc = O.make_config(**cfg_kw)
    q = O.init_tgv(oc)
    N = oc.nx * oc.ny * oc.nz
    rng = np.random.default_rng(seed)
    q = q.reshape(5, N)
    q[3] += 0.2 * np.sin(np.linspace(0, 40, N) + rng.random())
    q[0] *= 1.0 + 0.01 * rng.standard_normal(N)
    return q.ravel().copy()


CASES = [
    # larger grid first, then a smaller one (FP64 CRWENO5/LF/cd8, then mixed
    # WENO5-Z/AUSM/cd6 with the ABC forcing's source kernel)
    dict(n=(32, 24, 20), scheme=G.CRWENO5, vis=G.CD8, riemann=G.LF, model=G.NAVIERSTOKES3D),
    dict(n=(16, 20, 12), scheme=G.WENO5Z, vis=G.CD6, riemann=G.AUSM, model=G.NAVIERSTOKES3D,
         precision=G.MIXED, forcing=G.FORCING_ABC, forcing_amp=0.05, forcing_k=1.0),
]


def _hip():
    """The HIP runtime the product library links (its soname is already loaded
    in this process), for the caller-owned device buffers of the test."""
    G.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipStreamDestroy.argtypes = [vp]
    return h


H2D, D2H = 1, 2


def _noalloc(h, cfg, q, stream):
    L = G.lib()
    N = cfg.nx * cfg.ny * cfg.nz
    qd, rd, dd = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for p, nb in ((qd, 5 * N * 8), (rd, 5 * N * 8), (dd, 8)):
        assert h.hipMalloc(ctypes.byref(p), nb) == 0
    try:
        assert h.hipMemcpy(qd, q.ctypes.data, 5 * N * 8, H2D) == 0
        nan = np.full(5 * N, np.nan)
        assert h.hipMemcpy(rd, nan.ctypes.data, 5 * N * 8, H2D) == 0
        rc = L.nuc_rhs_noalloc(ctypes.byref(cfg), qd, rd, dd, stream)
        assert rc == 0, L.nuc_last_error(None).decode()
        assert h.hipStreamSynchronize(stream) == 0
        r = np.empty(5 * N)
        dt = np.empty(1)
        assert h.hipMemcpy(r.ctypes.data, rd, 5 * N * 8, D2H) == 0
        assert h.hipMemcpy(dt.ctypes.data, dd, 8, D2H) == 0
    finally:
        for p in (qd, rd, dd):
            h.hipFree(p)
    return r, float(dt[0])


def test_noalloc_rhs_equals_context_rhs_twice_shrinking():
    h = _hip()
    st = ctypes.c_void_p()
    assert h.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        # the first call on a stream of the caller's, the second on the null stream
        for i, (kw, stream) in enumerate(zip(CASES, (st, None))):
            print(f"no-alloc case {i}: {kw}", flush=True)
            cfg = G.make_config(**kw)
            q = _state({k: v for k, v in kw.items() if k not in ("precision", "forcing", "forcing_amp", "forcing_k")},
                       10 + i)
            with G.Context(cfg) as ctx:
                ctx.upload(q)
                r_ref, dt_ref = ctx.rhs()
            r, dt = _noalloc(h, cfg, q, stream)
            assert np.isfinite(r).all()
            assert np.array_equal(r, r_ref), np.abs(r - r_ref).max()
            assert dt == dt_ref
    finally:
        h.hipStreamDestroy(st)


def test_noalloc_rejects_oversize_and_const_forcing():
    L = G.lib()
    big = G.make_config(n=(128, 128, 128))
    assert 128 ** 3 > 1048576
    x = np.zeros(8)
    # refused before any device work: host pointers never reach a kernel
    rc = L.nuc_rhs_noalloc(ctypes.byref(big), x.ctypes.data, x.ctypes.data, None, None)
    assert rc == 1 and "scratch capacity" in L.nuc_last_error(None).decode()
    frc = G.make_config(n=(16, 16, 16), forcing=G.FORCING_CONST, forcing_amp=0.1, forcing_k=2.0)
    rc = L.nuc_rhs_noalloc(ctypes.byref(frc), x.ctypes.data, x.ctypes.data, None, None)
    assert rc == 1 and "Const forcing" in L.nuc_last_error(None).decode()


def test_noalloc_segmented_long_line_equals_context_rhs():
    """A line above 512 cells (swept in segments) through the no-allocation
    entry point: bit-identical to the context's RHS."""
    h = _hip()
    kw = dict(n=(1024, 12, 12), L=(2 * np.pi, 2 * np.pi * 12 / 1024, 2 * np.pi * 12 / 1024), scheme=G.WENO5Z,
              vis=G.CD8, riemann=G.AUSM, model=G.NAVIERSTOKES3D)
    print(f"no-alloc segmented case: {kw}", flush=True)
    cfg = G.make_config(**kw)
    q = _state(kw, 31)
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        r_ref, dt_ref = ctx.rhs()
    r, dt = _noalloc(h, cfg, q, None)
    assert np.isfinite(r).all()
    assert np.array_equal(r, r_ref), np.abs(r - r_ref).max()
    assert dt == dt_ref
