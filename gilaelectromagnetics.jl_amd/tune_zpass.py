"""Offline tuner of the pass-3 shape (elements per thread, columns per
workgroup) for every compiled transform length, both element types.

    python tune_zpass.py build   --out DIR            # on a machine with hipcc
    python tune_zpass.py measure --out DIR > res.jsonl  # on the GPU
    python tune_zpass.py table   res.jsonl            # rewrites csrc/gla_tune.hpp

`build` compiles one library per candidate pair (a ComplexF32 shape and a
ComplexF64 shape forced for all lengths); `measure` times pass 3 of a self
operator of N^3 cells for N <= 128 (64 x 64 x N beyond) for each length N with
each library (the library's own per-pass profile entry, HIP events around
repeated launches of the pass); `table` keeps, per element type and length,
the fastest shape when it beats the default by more than --margin (5%) and
writes the ones that differ from the default into the tables of
csrc/gla_tune.hpp.  (A 64 x 64 x N grid for every length understated the
gain of a shape at 96^3 by a factor of three: 8% there, 27% on the cube.)
"""
import argparse
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
# candidate pairs: ((E, PX) for ComplexF32, (E, PX) for ComplexF64); the first
# pair is the default shape
CANDIDATES = [((8, 16), (4, 8)), ((4, 8), (2, 8)), ((4, 16), (4, 4)), ((8, 8), (8, 4))]
LENGTHS = [16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512]
DEFAULT = {"c64": (8, 16), "c128": (4, 8)}


def build(out):
    os.makedirs(out, exist_ok=True)
    procs = []
    for i, (f, d) in enumerate(CANDIDATES):
        extra = f"-DGLA_ZT_E_F={f[0]} -DGLA_ZT_PX_F={f[1]} -DGLA_ZT_E_D={d[0]} -DGLA_ZT_PX_D={d[1]}"
        cmd = ["make", "-s", "-j3", "-C", CSRC, f"BUILD={out}/obj{i}", f"LIB={out}/lib{i}.so", f"EXTRA={extra}"]
        procs.append((i, subprocess.Popen(cmd)))
        if len(procs) == 2:
            for j, p in procs:
                if p.wait():
                    sys.exit(f"build {j} failed")
            procs = []
    for j, p in procs:
        if p.wait():
            sys.exit(f"build {j} failed")


def measure_one(lib_index, out, lengths):
    """Runs in a child process with GILA_AMD_LIB set to one candidate."""
    import ctypes
    from fractions import Fraction
    import numpy as np
    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as ge
    gila = ge.load_package()
    from gila_amd import _capi, inputs
    f, d = CANDIDATES[lib_index]
    for n in lengths:
        for typ, shape in ((np.complex64, f), (np.complex128, d)):
            N = (n, n, n) if n <= 128 else (64, 64, n)
            mem = gila.GlaOprMem(gila.GlaKerOpt(True), gila.GlaVol(N, (Fraction(1, 32),) * 3), setTyp=typ)
            x = inputs.complex_gaussian(N + (3,), seed=2).astype(typ)
            xd = torch.from_numpy(x.ravel(order="F").copy()).cuda()
            yd = torch.empty_like(xd)
            ms = (ctypes.c_double * 5)()
            by = (ctypes.c_int64 * 5)()
            st = torch.cuda.current_stream().cuda_stream
            _capi.check(gila.lib().gla_opr_profile(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                                   ctypes.c_void_p(yd.data_ptr()), yd.numel(), 1,
                                                   ctypes.c_void_p(st), 10, ms, by))
            torch.cuda.synchronize()
            print(json.dumps({"dtype": "c64" if typ == np.complex64 else "c128", "n": n, "grid": list(N), "e": shape[0],
                              "px": shape[1], "zpass_us": ms[2] * 1000.0}), flush=True)
            del mem, xd, yd


def measure(out, lengths):
    for i in range(len(CANDIDATES)):
        env = dict(os.environ, GILA_AMD_LIB=os.path.join(out, f"lib{i}.so"))
        r = subprocess.run([sys.executable, __file__, "_one", "--out", out, "--lib", str(i),
                            "--lengths", ",".join(map(str, lengths))], env=env)
        if r.returncode:
            sys.exit(f"measurement with candidate {i} failed ({r.returncode})")


def table(results, margin):
    """Per (element type, length) the fastest shape, kept only when it beats
    the default shape by more than `margin` (the timing grid is not every
    workload; small differences stay with the default)."""
    best, dflt = {}, {}
    for line in open(results):
        r = json.loads(line)
        key = (r["dtype"], r["n"])
        shape = (r["e"], r["px"])
        if shape == DEFAULT[r["dtype"]]:
            dflt[key] = r["zpass_us"]
        if key not in best or r["zpass_us"] < best[key][0]:
            best[key] = (r["zpass_us"], shape)
    for key in list(best):
        if key in dflt and best[key][0] > (1.0 - margin) * dflt[key]:
            best[key] = (dflt[key], DEFAULT[key[0]])
    path = os.path.join(CSRC, "gla_tune.hpp")
    src = open(path).read()
    for dt, tag in (("c64", "C64"), ("c128", "C128")):
        rows = [f"            case {n}: return {{{s[0]}, {s[1]}}};"
                for (d, n), (_, s) in sorted(best.items()) if d == dt and s != DEFAULT[dt]]
        body = "".join(r + "\n" for r in rows)
        src = re.sub(rf"(// BEGIN TUNED {tag}\n)(.*?)([ \t]*// END TUNED {tag})",
                     lambda m: m.group(1) + body + m.group(3), src, flags=re.S)
    open(path, "w").write(src)
    for (d, n), (us, s) in sorted(best.items()):
        print(f"{d} {n}: {s} {us:.1f} us")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["build", "measure", "table", "_one"])
    ap.add_argument("results", nargs="?")
    ap.add_argument("--out", default=os.path.join(HERE, "_tune"))
    ap.add_argument("--lib", type=int, default=0)
    ap.add_argument("--lengths", default=",".join(map(str, LENGTHS)))
    ap.add_argument("--margin", type=float, default=0.05)
    a = ap.parse_args()
    lengths = [int(v) for v in a.lengths.split(",")]
    if a.cmd == "build":
        build(a.out)
    elif a.cmd == "measure":
        measure(a.out, lengths)
    elif a.cmd == "_one":
        measure_one(a.lib, a.out, lengths)
    else:
        table(a.results, a.margin)


if __name__ == "__main__":
    main()
