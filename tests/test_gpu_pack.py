"""Packed launches (nuc_pack_*): several small boxes advanced by one launch per
kernel.  Every member's state, dt, residual and time must be bit-identical to
the same box stepped alone with nuc_step, with and without graph replay, for
members of different sizes and physics parameters sharing one kernel
configuration; incompatible or unpackable contexts are refused.
"""
import numpy as np
import pytest

import nuc3d_amd as G

pytestmark = pytest.mark.gpu


def _cfg(n, scheme=G.CRWENO5, riemann=G.LF, vis=G.CD6, model=G.NAVIERSTOKES3D, precision=G.FP64, mach=0.2,
         re=1000.0, **kw):
    return G.make_config(n=n, scheme=scheme, vis=vis, riemann=riemann, model=model, mach=mach, reynolds=re,
                         cfl=0.5, precision=precision, **kw)


def _alone(cfg, q, steps):
    with G.Context(cfg) as ctx:
        ctx.upload(q)
        info = ctx.step(steps)
        return ctx.download(), info


def _packed(cfgs, qs, slices, use_graph):
    ctxs = [G.Context(c) for c in cfgs]
    try:
        for c, q in zip(ctxs, qs):
            c.upload(q)
        with G.Pack(ctxs, use_graph=use_graph) as p:
            for k in slices:
                p.step_async(k)
            infos = [c.sync() for c in ctxs]
            outs = [c.download() for c in ctxs]
        return outs, infos
    finally:
        for c in ctxs:
            c.close()


MEMBER_SETS = {
    # line-length class R = 2 (33..64 cells, tails in every direction)
    "crweno5-lf-fp64": [
        (_cfg((32, 32, 32), mach=0.1, re=1600.0), "tgv", 0),
        (_cfg((48, 40, 36), mach=0.3, re=800.0), "hit", 1),
        (_cfg((40, 56, 44), mach=0.25, re=2400.0), "hit", 2),
    ],
    # mixed precision, AUSM, WENO5-Z, cd8
    "weno5z-ausm-mixed": [
        (_cfg((24, 24, 24), scheme=G.WENO5Z, riemann=G.AUSM, vis=G.CD8, precision=G.MIXED, mach=0.4), "hit", 3),
        (_cfg((36, 30, 28), scheme=G.WENO5Z, riemann=G.AUSM, vis=G.CD8, precision=G.MIXED, mach=0.15), "tgv", 4),
    ],
    # R = 4 (97..128 cells), WENO5-JS, plus a rank grid (closure placement per member)
    "weno5js-r4": [
        (_cfg((112, 100, 120), scheme=G.WENO5JS), "hit", 5),
        (_cfg((100, 104, 98), scheme=G.WENO5JS, npr=(2, 1, 1), mach=0.35), "hit", 6),
    ],
    # Euler, hccs (7-point), lines of exactly 64 cells (no tail checks)
    "hccs-euler-full": [
        (_cfg((64, 64, 64), scheme=G.HCCS, model=G.EULER3D, riemann=G.AUSM), "tgv", 7),
        (_cfg((64, 64, 64), scheme=G.HCCS, model=G.EULER3D, riemann=G.AUSM, mach=0.4), "hit", 8),
    ],
    # the remaining operators: hccs77 with AUSM+ and cd2; WENO3-JS; hccs78
    "hccs77-ausmp-cd2": [
        (_cfg((40, 36, 32), scheme=G.HCCS77, riemann=G.AUSMP, vis=G.CD2), "hit", 9),
        (_cfg((48, 48, 40), scheme=G.HCCS77, riemann=G.AUSMP, vis=G.CD2, mach=0.3), "tgv", 10),
    ],
    "weno3js-upwind-mixed": [
        (_cfg((32, 32, 32), scheme=G.WENO3JS, precision=G.MIXED), "hit", 11),
        (_cfg((40, 34, 36), scheme=G.WENO3JS, precision=G.MIXED, mach=0.35), "hit", 12),
    ],
    # R = 3 (65..96 cells, the odd chunk length)
    "hccs78-r3": [
        (_cfg((72, 66, 80), scheme=G.HCCS78, vis=G.CD8), "hit", 13),
        (_cfg((96, 70, 68), scheme=G.HCCS78, vis=G.CD8, npr=(1, 2, 1)), "hit", 14),
    ],
    "crweno5-r3-mixed": [
        (_cfg((96, 96, 96), precision=G.MIXED), "hit", 15),
        (_cfg((80, 90, 70), precision=G.MIXED, mach=0.3), "tgv", 16),
    ],
}


@pytest.mark.parametrize("use_graph", [False, True], ids=["launches", "graph"])
@pytest.mark.parametrize("name", sorted(MEMBER_SETS))
def test_pack_members_bit_identical_to_alone(name, use_graph):
    members = MEMBER_SETS[name]
    print(f"pack {name} graph={use_graph}", flush=True)
    cfgs = [m[0] for m in members]
    qs = [G.init_tgv(c) if kind == "tgv" else G.init_hit(c, k0=3.0, u_rms=0.5, seed=seed)
          for c, kind, seed in members]
    steps = 5
    outs, infos = _packed(cfgs, qs, [2, 3], use_graph)
    for i, (c, q) in enumerate(zip(cfgs, qs)):
        ref, info = _alone(c, q, steps)
        assert np.array_equal(outs[i], ref), (name, i, np.abs(outs[i] - ref).max())
        assert infos[i] == info, (name, i, infos[i], info)


def test_pack_compatibility_rules():
    a = G.Context(_cfg((32, 32, 32)))
    b = G.Context(_cfg((48, 48, 48)))
    c = G.Context(_cfg((32, 32, 32), scheme=G.WENO5JS))
    d = G.Context(_cfg((64, 64, 64)))                                  # n = 64: tail-free kernels
    e = G.Context(_cfg((32, 32, 32), forcing=G.FORCING_CONST, forcing_k=2.0))
    f = G.Context(_cfg((160, 32, 32)))                                 # line longer than 128
    try:
        assert G.pack_compatible(a, b)
        assert not G.pack_compatible(a, c)
        assert not G.pack_compatible(a, d)
        assert not G.pack_compatible(a, e)
        assert not G.pack_compatible(f, f)
        for bad in ([a, c], [a, e], [f], [a, a]):
            with pytest.raises(G.NucError):
                G.Pack(bad)
        with G.Pack([a, b]) as p:
            p.step_async(0)
    finally:
        for x in (a, b, c, d, e, f):
            x.close()
