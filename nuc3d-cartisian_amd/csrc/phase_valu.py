"""Per-phase instruction counts of one direction sweep (k_sweep, nuc_sweep.h),
from the gfx950 assembly of an instrumented copy of this directory's sources.

The copy (in a temporary directory; the sources here are only read) differs
from the product kernel in three ways, none of which changes what a phase
computes:
- the variable loop is unrolled and the run-time flags of a middle sweep
  (direction y, rhs read and written, no RK3 update, cd8 viscous terms, LF
  splitting) are constants, so that every block is executed exactly once per
  wave and the static count of each phase is its dynamic count per wave;
- an empty asm statement with a scheduling barrier marks each phase boundary;
- inside a CRWENO5 row the indicator, weight and row values pass through empty
  asm statements, so that the compiler cannot move their arithmetic across the
  boundaries.

The total is checked against SQ_INSTS_VALU / SQ_WAVES of the product kernel
(a measured figure, passed with --measured).

    python3 nuc3d-cartisian_amd/csrc/phase_valu.py [--measured VALU_PER_WAVE]
"""
import argparse
import collections
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PH = ('#define PH(x) do { __builtin_amdgcn_sched_barrier(0); asm volatile(";PHASE " #x); '
      '__builtin_amdgcn_sched_barrier(0); } while (0)\n')

# (anchor, replacement) pairs; the flag substitutions (FLAGS) replace every
# occurrence, the phase marks (the edit lists) must match exactly once
FLAGS = [
    ("const int n = A.n, d = A.dir;", "const int n = A.n, d = 1;"),
    ("A.riemann == RIE_AUSMP", "false"),
    ("A.epilogue", "false"),
    ("(!A.first || A.accum)", "(true)"),
    ("A.fv_f32 ?", "false ?"),
    ("A.vis != VIS_NONE", "true"),
]
SWEEP_EDITS = [
    ("(innermask >> k) & 1u, A.vis, T(A.inv_h))", "(innermask >> k) & 1u, VIS_CD8, T(A.inv_h))"),
    ("if (A.fmean && v == 4)", "if (false)"),
    ("#pragma unroll 1\n    for (int v = 0; v < 5; v++) {",
     "#pragma unroll\n    for (int v = 0; v < 5; v++) {\n        PH(split);"),
    ("    // split parameters of the own cells", "    PH(prologue);\n    // split parameters of the own cells"),
    ("        fetch(v + 1, qnx, xnx);", "        PH(fetch);\n        fetch(v + 1, qnx, xnx);"),
    ("        sync_lines(d);\n        double out[R];", "        PH(barrier);\n        sync_lines(d);\n        double out[R];"),
    ("        sync_wave();\n        if (wv) {", "        PH(write_interfaces);\n        sync_wave();\n        if (wv) {"),
    ("        // the final phase's own global reads", "        PH(final);\n        // the final phase's own global reads"),
]
LINES_EDITS = [
    ("    T win[W];\n#pragma unroll\n    for (int t = 0; t < W; t++) win[t]",
     "    PH(window);\n    T win[W];\n#pragma unroll\n    for (int t = 0; t < W; t++) win[t]"),
    ("            T f4 = win[r + 4];\n            T d4 = f4 - f3;\n            T e3 = T(13.0 / 12.0) * sq(d4 - d3);\n"
     "            T b0, b1, b2, q0, q1, q2, h2;\n            w5_diff<T>(f2, f3, d1, d2, d3, d4, e1, e2, e3, b0, b1, b2, q0, q1, q2, h2);\n"
     "            TW P0, P1, P2;",
     "            PH(indicators);\n            T f4 = win[r + 4];\n            T d4 = f4 - f3;\n            T e3 = T(13.0 / 12.0) * sq(d4 - d3);\n"
     "            T b0, b1, b2, q0, q1, q2, h2;\n            w5_diff<T>(f2, f3, d1, d2, d3, d4, e1, e2, e3, b0, b1, b2, q0, q1, q2, h2);\n"
     "            TW P0, P1, P2;"),
    ("            TW P0, P1, P2;   // weights with unit ideal weights\n"
     "            w5_weights<TW>(TW(b0), TW(b1), TW(b2), TW(1), TW(1), TW(1), P0, P1, P2);\n"
     "            // Every row is scaled",
     "            asm volatile(\"\" : \"+v\"(b0), \"+v\"(b1), \"+v\"(b2), \"+v\"(q0), \"+v\"(q1), \"+v\"(q2));\n"
     "            PH(weights);\n"
     "            TW P0, P1, P2;   // weights with unit ideal weights\n"
     "            w5_weights<TW>(TW(b0), TW(b1), TW(b2), TW(1), TW(1), TW(1), P0, P1, P2);\n"
     "            asm volatile(\"\" : \"+v\"(P0), \"+v\"(P1), \"+v\"(P2));\n"
     "            PH(rows);\n"
     "            // Every row is scaled"),
    ("                              \"+v\"(a), \"+v\"(b), \"+v\"(c), \"+v\"(d));\n            ch.row(r, a, b, c, d);",
     "                              \"+v\"(a), \"+v\"(b), \"+v\"(c), \"+v\"(d));\n            PH(elimination);\n"
     "            ch.row(r, a, b, c, d);"),
    ("        ch.solve(h);\n#pragma unroll\n        for (int r = 0; r < R; r++) out[r] = ch.Df[r];\n    } else if constexpr (SCH == HCCS77)",
     "        PH(back_substitution);\n        ch.solve(h);\n#pragma unroll\n        for (int r = 0; r < R; r++) out[r] = ch.Df[r];\n"
     "    } else if constexpr (SCH == HCCS77)"),
    ("        pcr32(al, ga, de, h);\n        T X1 = de;", "        PH(pcr);\n        pcr32(al, ga, de, h);\n        PH(fold_back);\n        T X1 = de;"),
]
ORDER = ["entry", "prologue", "split", "fetch", "barrier", "window", "indicators", "weights", "rows",
         "elimination", "back_substitution", "pcr", "fold_back", "write_interfaces", "final"]


def edit(text, edits, what, every=False):
    for a, b in edits:
        if every:
            if a not in text:
                raise SystemExit(f"{what}: flag not found: {a!r}")
            text = text.replace(a, b)
            continue
        if text.count(a) != 1:
            raise SystemExit(f"{what}: anchor found {text.count(a)} times: {a[:70]!r}")
        text = text.replace(a, b)
    return text


def build_tu(tmp, inst):
    for h in os.listdir(HERE):
        if h.endswith(".h"):
            shutil.copy(os.path.join(HERE, h), tmp)
    sweep = open(os.path.join(HERE, "nuc_sweep.h")).read()
    sweep = edit(sweep, FLAGS, "nuc_sweep.h", every=True)
    sweep = edit(sweep, SWEEP_EDITS, "nuc_sweep.h")
    open(os.path.join(tmp, "nuc_sweep.h"), "w").write(PH + sweep)
    lines = open(os.path.join(tmp, "nuc_lines.h")).read()
    open(os.path.join(tmp, "nuc_lines.h"), "w").write(PH + edit(lines, LINES_EDITS, "nuc_lines.h"))
    tu = ('#include "nuc_sweep.h"\n'
          f"template __global__ void nucsweep::k_sweep<{inst}>(nucsweep::SweepArgs);\n")
    path = os.path.join(tmp, "phase.hip")
    open(path, "w").write(tu)
    return path


def count(asm):
    L = asm.split("\n")
    start = next(i for i, l in enumerate(L) if re.match(r"^_ZN8nucsweep7k_sweep\S*:", l))
    end = next(i for i in range(start, len(L)) if "s_endpgm" in L[i])
    cnt = collections.defaultdict(collections.Counter)
    cur = "entry"
    for l in L[start:end + 1]:
        t = l.strip()
        if t.startswith(";PHASE"):
            cur = t.split()[1]
            continue
        if not t or t.startswith((";", ".")) or t.endswith(":"):
            continue
        op = t.split()[0]
        c = cnt[cur]
        if op.startswith("v_"):
            c["valu"] += 1
            if "_f64" in op:
                c["f64"] += 1
        elif op.startswith("ds_"):
            c["lds"] += 1
        elif op.startswith(("global_", "buffer_")):
            c["vmem"] += 1
        elif op.startswith("s_") and op not in ("s_waitcnt", "s_nop", "s_barrier"):
            c["salu"] += 1
    vg = re.search(r"\.num_vgpr, (\d+)", asm)
    return cnt, int(vg.group(1)) if vg else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--inst", default="3, 0, 8, 4, double, true, 0",
                    help="k_sweep template arguments (default: CRWENO5, LF, R=8, 4 lines, FP64, FULL, no RK3 update)")
    ap.add_argument("--measured", type=float, default=None, help="measured VALU wave-instructions per wave")
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as tmp:
        tu = build_tu(tmp, a.inst)
        out = os.path.join(tmp, "phase.s")
        r = subprocess.run([HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                            "-Xclang", "-packed-fp32-ops", "-I" + INC, "-I" + tmp, "--offload-device-only", "-S",
                            tu, "-o", out], capture_output=True, text=True)
        if r.returncode != 0:
            sys.exit(r.stderr)
        cnt, vgpr = count(open(out).read())
    tot = collections.Counter()
    for c in cnt.values():
        tot.update(c)
    print(f"k_sweep<{a.inst}> (instrumented copy: {vgpr} VGPRs), per wave over the 5 variables\n")
    print("| phase | VALU | of which FP64 | share of VALU | LDS | VMEM | SALU |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for k in ORDER + sorted(set(cnt) - set(ORDER)):
        c = cnt.get(k)
        if not c:
            continue
        print(f"| {k} | {c['valu']} | {c['f64']} | {c['valu'] / tot['valu']:.1%} | {c['lds']} | {c['vmem']} | {c['salu']} |")
    print(f"| total | {tot['valu']} | {tot['f64']} | 100% | {tot['lds']} | {tot['vmem']} | {tot['salu']} |")
    if a.measured:
        print(f"\nmeasured VALU per wave {a.measured:.0f}; instrumented static total / measured = "
              f"{tot['valu'] / a.measured:.3f}")


if __name__ == "__main__":
    main()
