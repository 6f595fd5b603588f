"""Write tests/golden/ref_vectors.npz: input/output vectors of the reference's
own scheme objects (weno5js, weno5z, weno7js, crweno5, weno3js, upwind1st,
hccs78, hccs77, centraldifference2nd/6th/8th, tvdrk3rd), computed by oracle/_ref/libnucref.so -- the reference's sources
compiled unmodified from /root/reference/src by oracle/Makefile.

The fixture lets the oracle be checked against the reference where oracle/_ref
cannot be built (no reference tree).  Regenerate with
    make -f oracle/Makefile all && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def lines(rng, n1, n2, n3, bf, kind):
    if kind == "smooth":
        x = np.linspace(0, 2 * np.pi, n1 + 2 * bf, endpoint=False)
        base = np.sin(x) + 0.3 * np.cos(3 * x) + 2.0
        f = np.broadcast_to(base, (n3, n2, n1 + 2 * bf)) + 0.01 * rng.standard_normal((n3, n2, n1 + 2 * bf))
    elif kind == "step":
        f = np.where(rng.random((n3, n2, n1 + 2 * bf)) < 0.5, 1.0, 10.0)
    else:
        f = rng.standard_normal((n3, n2, n1 + 2 * bf)) * 3.0 + 1.0
    f = np.ascontiguousarray(f)
    return (np.ascontiguousarray(f[:, :, bf:bf + n1]), np.ascontiguousarray(f[:, :, :bf]),
            np.ascontiguousarray(f[:, :, bf + n1:]))


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libnucref.so is not built")
    out = {}
    rng = np.random.default_rng(20261202)
    k = 0
    for scheme in (O.WENO5JS, O.WENO5Z, O.WENO7JS, O.CRWENO5):
        for upwind in (1, -1):
            for kind in ("smooth", "random", "step"):
                for (n1, n2, n3) in ((16, 2, 1), (27, 1, 2)):
                    fin, gL, gR = lines(rng, n1, n2, n3, 4, kind)
                    ref = O.field_reconstruct(scheme, upwind, O.halo(scheme), fin, gL, gR, use_ref=True)
                    out[f"rec_{k}_meta"] = np.array([scheme, upwind, O.halo(scheme)])
                    out[f"rec_{k}_in"], out[f"rec_{k}_gL"], out[f"rec_{k}_gR"] = fin, gL, gR
                    out[f"rec_{k}_out"] = ref
                    k += 1
    j = 0
    for scheme, ts in ((O.CD6, 3), (O.CD6, 4), (O.CD8, 4)):
        for (n1, n2, n3) in ((16, 2, 1), (29, 1, 2)):
            fin, gL, gR = lines(rng, n1, n2, n3, 4, "random")
            out[f"dif_{j}_meta"] = np.array([scheme, ts])
            out[f"dif_{j}_in"], out[f"dif_{j}_gL"], out[f"dif_{j}_gR"] = fin, gL, gR
            out[f"dif_{j}_out"] = O.field_diff(scheme, ts, fin, 0.173, gL, gR, use_ref=True)
            j += 1
    for stage in range(3):
        r, un, ui = (rng.standard_normal((5, 23)) for _ in range(3))
        out[f"rk3_{stage}_in"] = np.stack([r, un, ui])
        out[f"rk3_{stage}_out"] = O.rk3(stage, r, un, ui, 0.0137, use_ref=True)
    # schemes added later draw after the first set, which keeps its vectors
    for scheme in (O.WENO3JS, O.UPWIND1ST, O.HCCS78):
        for upwind in (1, -1):
            for kind in ("smooth", "random", "step"):
                for (n1, n2, n3) in ((16, 2, 1), (27, 1, 2)):
                    fin, gL, gR = lines(rng, n1, n2, n3, 4, kind)
                    ref = O.field_reconstruct(scheme, upwind, O.halo(scheme), fin, gL, gR, use_ref=True)
                    out[f"rec_{k}_meta"] = np.array([scheme, upwind, O.halo(scheme)])
                    out[f"rec_{k}_in"], out[f"rec_{k}_gL"], out[f"rec_{k}_gR"] = fin, gL, gR
                    out[f"rec_{k}_out"] = ref
                    k += 1
    for scheme, ts in ((O.CD2, 1), (O.CD2, 2), (O.CD2, 4)):
        for (n1, n2, n3) in ((16, 2, 1), (29, 1, 2)):
            fin, gL, gR = lines(rng, n1, n2, n3, 4, "random")
            out[f"dif_{j}_meta"] = np.array([scheme, ts])
            out[f"dif_{j}_in"], out[f"dif_{j}_gL"], out[f"dif_{j}_gR"] = fin, gL, gR
            out[f"dif_{j}_out"] = O.field_diff(scheme, ts, fin, 0.173, gL, gR, use_ref=True)
            j += 1
    for scheme in (O.HCCS77,):
        for upwind in (1, -1):
            for kind in ("smooth", "random", "step"):
                for (n1, n2, n3) in ((16, 2, 1), (27, 1, 2)):
                    fin, gL, gR = lines(rng, n1, n2, n3, 4, kind)
                    ref = O.field_reconstruct(scheme, upwind, O.halo(scheme), fin, gL, gR, use_ref=True)
                    out[f"rec_{k}_meta"] = np.array([scheme, upwind, O.halo(scheme)])
                    out[f"rec_{k}_in"], out[f"rec_{k}_gL"], out[f"rec_{k}_gR"] = fin, gL, gR
                    out[f"rec_{k}_out"] = ref
                    k += 1
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print(f"{k} reconstruction, {j} difference, 3 rk3 vectors")


if __name__ == "__main__":
    main()
