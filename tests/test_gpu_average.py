"""Time-averaged statistics and TKE budget on the device (nuc_average_*) against
the oracle's restatement of postproc::solveAveraged0/1/2
(src/postproc.cpp:1186-2110).

The device and the oracle average the same sequence of states (the oracle's
states, uploaded to the device), so the comparison isolates the averaging and
budget arithmetic.  Budget terms are differences of nearly equal averaged
products, so each field is held to max(1e-12 x its magnitude, 10 x the oracle's
own change when its input states move by one ulp).
"""
import numpy as np
import pytest

import nuc3d_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

COUNTS = (10, 21, 28, 49, 45)
GROUPS = ("primary", "grads", "derived", "tke", "budget")


def _cfgs(n, npr, scheme, vis, rie):
    kw = dict(n=n, npr=npr, scheme=scheme, vis=vis, riemann=rie, model=G.NAVIERSTOKES3D, mach=0.3,
              reynolds=400.0, cfl=0.5)
    return G.make_config(**kw), O.make_config(**kw)


def _oracle_average(oc, states, avg_type):
    s = O.Solver(oc)
    P = O.Post(s, avg_type)
    for q in states:
        s.set_state(q)
        P.update()
    return [P.group(g) for g in range(avg_type + 1 if avg_type < 3 else 5)]


def _states(gc, oc, seed, n):
    q = G.init_hit(gc, k0=3.0, u_rms=0.5, seed=seed)
    s = O.Solver(oc)
    s.set_state(q)
    out = []
    for _ in range(n):
        out.append(s.get_state())
        s.step(1)
    return out


CASES = [((16, 16, 16), (1, 1, 1), G.CRWENO5, G.CD6, G.LF),
         ((24, 20, 16), (2, 1, 1), G.WENO5JS, G.CD8, G.AUSM)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-s{c[2]}-v{c[3]}")
def test_average_and_budget_match_oracle(case):
    n, npr, scheme, vis, rie = case
    print(f"average case {case}", flush=True)
    gc, oc = _cfgs(n, npr, scheme, vis, rie)
    states = _states(gc, oc, 31, 3)
    ref = _oracle_average(oc, states, 3)
    rng = np.random.default_rng(5)
    pert = [q * (1 + 2.2e-16 * rng.choice([-1.0, 0.0, 1.0], size=q.size)) for q in states]
    ref_p = _oracle_average(oc, pert, 3)
    with G.Context(gc) as ctx:
        ctx.average_begin(3)
        for q in states:
            ctx.upload(q)
            ctx.average_update()
        assert ctx.average_steps() == 3
        got = [ctx.average_group(g) for g in range(5)]
        means = [ctx.average_means(g) for g in range(5)]
    for g in range(5):
        worst = 0.0
        for f in range(COUNTS[g]):
            o, d = ref[g][f], got[g][f]
            floor = np.abs(ref_p[g][f] - o).max()
            tol = max(1e-12 * np.abs(o).max(), 10 * floor, 1e-300)
            err = np.abs(d - o).max()
            worst = max(worst, err / tol)
            assert err <= tol, (GROUPS[g], f, err, tol)
            # volume means: fixed-order device reduction of the same field
            assert abs(means[g][f] - d.mean()) <= 1e-13 * np.abs(d).max() + 1e-300, (GROUPS[g], f)
        print(f"{GROUPS[g]}: worst error / tolerance {worst:.3f}", flush=True)


def test_average_api_rules():
    gc, oc = _cfgs((16, 16, 16), (1, 1, 1), G.WENO5JS, G.CD6, G.LF)
    q = G.init_hit(gc, k0=3.0, u_rms=0.5, seed=2)
    with G.Context(gc) as ctx:
        ctx.upload(q)
        with pytest.raises(G.NucError):
            ctx.average_update()                   # nothing begun
        with pytest.raises(G.NucError):
            ctx.average_begin(4)
        ctx.average_begin(1)
        ctx.average_update()
        ctx.average_update()
        assert ctx.average_steps() == 2
        rho = ctx.average_field(0, 0)
        assert np.allclose(rho, q[:16 ** 3])       # the same state twice: its own value
        with pytest.raises(G.NucError):
            ctx.average_field(2, 0)                # derived fields need type 2
        ctx.average_begin(2)                       # restart: zero fields, step 0
        assert ctx.average_steps() == 0
        ctx.average_update()
        ctx.average_field(2, 27)
        ctx.average_end()
        with pytest.raises(G.NucError):
            ctx.average_update()


def test_mixed_path_budget_error():
    """The mixed path's error in the volume-averaged TKE budget terms against
    the FP64 path, decaying turbulence 48^3, averages over 10 steps."""
    terms = ["tke", "production", "transport", "pressure_diffusion", "p1", "p2", "p3", "viscous_diffusion",
             "v1", "v2", "v3", "convection"]
    out = {}
    for prec in (G.FP64, G.MIXED):
        cfg = G.make_config(n=(48, 48, 48), scheme=G.CRWENO5, vis=G.CD8, riemann=G.LF, model=G.NAVIERSTOKES3D,
                            mach=0.3 / np.sqrt(3.0), reynolds=200.0, cfl=0.5, precision=prec)
        q = G.init_hit(cfg, k0=4.0, u_rms=1.0, seed=20261202)
        with G.Context(cfg) as ctx:
            ctx.upload(q)
            ctx.average_begin(3)
            for _ in range(10):
                ctx.step(1)
                ctx.average_update()
            out[prec] = (ctx.average_means(4)[:12], ctx.average_means(0))
    b64, b32 = out[G.FP64][0], out[G.MIXED][0]
    scale = np.abs(b64[[0, 1, 3, 7]]).max()          # the non-divergence terms set the budget's scale
    for i, t in enumerate(terms):
        e = abs(b32[i] - b64[i]) / max(abs(b64[i]), scale)
        print(f"budget {t:20s} fp64 {b64[i]: .6e} mixed {b32[i]: .6e} rel {e:.2e}", flush=True)
        assert e <= 1e-5, t
    assert abs(b32[0] - b64[0]) <= 1e-5 * abs(b64[0])   # tke itself


def _check_average_tecplot(path, ctx, n, npr, rank):
    """postproc::OutputAveraged_tec's layout (src/postproc.cpp:523-720)."""
    lines = open(path).read().split("\n")
    assert lines[0] == "title=time_0.125"
    names = lines[1].split(",")
    assert names[:3] == ["variables=x", "y", "z"] and len(names) == 3 + 153
    assert names[3] == "rho" and names[3 + 31] == "rhou" and names[3 + 108] == "tke" and names[-1] == "dw_tkedz"
    nl = [n[d] // npr[d] for d in range(3)]
    assert lines[2] == f" Zone I = {nl[0] + 1}, J= {nl[1] + 1}, K={nl[2] + 1}"
    assert lines[3] == " DATAPACKING=BLOCK, VARLOCATION=([4-156]=CELLCENTERED)"
    vals = np.array([float(x) for x in lines[4:] if x.strip()])
    nn = (nl[0] + 1) * (nl[1] + 1) * (nl[2] + 1)
    nc = nl[0] * nl[1] * nl[2]
    assert vals.size == 3 * nn + 153 * nc
    cx = rank % npr[0]
    for g, i, k in ((0, 0, 0), (1, 15, 10 + 15), (4, 0, 108)):   # rho, omega0, tke
        f = ctx.average_field(g, i).reshape(n[2], n[1], n[0])[:, :, cx * nl[0]:(cx + 1) * nl[0]].ravel()
        got = vals[3 * nn + k * nc:3 * nn + (k + 1) * nc]
        assert np.allclose(got, f, rtol=1e-11, atol=1e-12 * np.abs(f).max()), (g, i)


def test_averaged_files_round_trip(tmp_path):
    """postproc's averaged-statistics files (src/postproc.cpp:157-256,412-522):
    write per rank, read back into a new context, and the continued average
    equals the uninterrupted one bit for bit."""
    gc, _ = _cfgs((16, 16, 16), (2, 1, 1), G.WENO5JS, G.CD6, G.LF)
    q = G.init_hit(gc, k0=3.0, u_rms=0.5, seed=4)
    with G.Context(gc) as a:
        a.upload(q)
        a.average_begin(3)
        for _ in range(3):
            a.step(1)
            a.average_update()
        a.average_write(3, tmp_path)
        tec_dir = tmp_path / "tec"
        tec_dir.mkdir()
        a.average_write_tecplot(3, tec_dir, 0.125)
        _check_average_tecplot(tec_dir / "Averaged_tec_step_3_id_1.dat", a, (16, 16, 16), (2, 1, 1), 1)
        names = sorted(p.name for p in tmp_path.iterdir() if p.is_file())
        assert names == ["Averaged_step_3_id_0.bin", "Averaged_step_3_id_1.bin"]
        raw = np.fromfile(tmp_path / names[1], dtype=np.uint8)
        assert raw.size == 4 + 153 * 8 * 16 * 16 * 8                  # int averStep + 153 fields of 8x16x16
        assert int(raw[:4].view(np.int32)[0]) == 3
        state = a.download()
        a.step(1)
        a.average_update()
        ref = [a.average_group(g) for g in range(5)]
    # the reader's name (_ID_) as well as the writer's (_id_)
    (tmp_path / names[0]).rename(tmp_path / "Averaged_step_3_ID_0.bin")
    with G.Context(gc) as b:
        b.upload(state)
        b.average_read(3, tmp_path, 3)
        assert b.average_steps() == 3
        b.step(1)
        b.average_update()
        got = [b.average_group(g) for g in range(5)]
    for g in range(5):
        assert np.array_equal(got[g], ref[g]), g
    with G.Context(gc) as c:
        with pytest.raises(G.NucError):
            c.average_read(3, tmp_path, 4)                             # no such averaging step

