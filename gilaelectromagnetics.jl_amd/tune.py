"""Offline autotuner of the kernel launch shapes, keyed on transform length,
element type and grid shape; writes the tables of csrc/gla_tune.hpp.

    python tune.py build   --out DIR              # on a machine with hipcc
    python tune.py measure --out DIR > res.jsonl  # on the GPU
    python tune.py table   res.jsonl              # rewrites csrc/gla_tune.hpp

Seven parts, each a set of candidate libraries that force one choice for every
length (compile-time overrides of gla_tune.hpp):

  zpass  pass 3: (elements per thread, columns per workgroup), timed on a
         large cross-section (128 x 128 x N: 256 x 256 columns) and a small
         one (64 x 64 x N: 128 x 128 columns, below ZSMALL_COLS) -- the two
         grid classes of ztune
  rows   passes 1, 2, 4, 5: elements per thread of the axis-1 rows and the
         axis-2 columns and the axis-2 bundle width, timed on N^3 (N <= 128;
         N x N x 128 beyond: both row axes of length N)
  plan   the radix plan policy (every thread busy in every stage, against the
         fewest LDS exchanges) for all passes of a length, timed as pass 3 on
         the large cross-section plus the row passes on the rows grid
  xk     passes 1 and 5: the two parities of a row through the stages
         together or one after the other, timed on the rows grid
  gg     pass 3: the first point's G^ loads issued before the forward
         transform or not, timed on the large cross-section
  yk     passes 2 and 4: the two parities of a bundle together or in turn,
         timed on the rows grid
  rl     pass 3: the input column kept, read again or split into both
         parities first, timed on the large cross-section

Operators come from all-zero Green data (the timing does not depend on the
values), so no setup runs.  Each (part, element type, length[, class]) keeps
its current table entry (or default) unless a candidate beats it by more than
--margin (5%): a timing grid is not every workload, and differences within
the run-to-run noise stay with the incumbent.  (A 64 x 64 x N grid for every
length understated the pass-3 gain of a shape at 96^3 threefold against the
cube: the cross-section matters, hence the classes.)
"""
import argparse
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LENGTHS = [16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512]

# pass 3: ((E, PX) ComplexF32, (E, PX) ComplexF64); the fallback defaults of
# ztune are (8, 16) and (4, 8)
ZCAND = [((8, 16), (4, 8)), ((4, 8), (2, 8)), ((4, 16), (4, 4)), ((8, 8), (8, 4)), ((4, 4), (2, 4)),
         ((2, 8), (4, 16))]
ZDEFAULT = {"c64": (8, 16), "c128": (4, 8)}
# rows: ((ex, ey, pxy) ComplexF32, (ex, ey, pxy) ComplexF64)
RCAND = [((4, 4, 16), (4, 4, 8)), ((8, 8, 16), (8, 8, 8)), ((4, 8, 8), (4, 8, 4)), ((8, 4, 32), (8, 4, 16))]


def rdefault(dt, n):
    if dt == "c64":
        return (4, 4, 16) if n <= 160 else (8, 8, 16)
    return (4, 4, 8) if n <= 128 else (8, 8, 8)


def candidates():
    """(name, extra compile flags, description) of every candidate library."""
    out = []
    for i, (f, d) in enumerate(ZCAND):
        out.append((f"z{i}", f"-DGLA_ZT_E_F={f[0]} -DGLA_ZT_PX_F={f[1]} -DGLA_ZT_E_D={d[0]} -DGLA_ZT_PX_D={d[1]}",
                    {"part": "zpass", "c64": f, "c128": d}))
    for i, (f, d) in enumerate(RCAND):
        out.append((f"r{i}", f"-DGLA_RT_EX_F={f[0]} -DGLA_RT_EY_F={f[1]} -DGLA_RT_PXY_F={f[2]} "
                             f"-DGLA_RT_EX_D={d[0]} -DGLA_RT_EY_D={d[1]} -DGLA_RT_PXY_D={d[2]}",
                    {"part": "rows", "c64": f, "c128": d}))
    out.append(("p1", "-DGLA_PT_F=1 -DGLA_PT_D=1", {"part": "plan", "c64": 1, "c128": 1}))
    out.append(("p0", "-DGLA_PT_F=0 -DGLA_PT_D=0", {"part": "plan", "c64": 0, "c128": 0}))
    out.append(("x1", "-DGLA_XK_F=1 -DGLA_XK_D=1", {"part": "xk", "c64": 1, "c128": 1}))
    out.append(("x2", "-DGLA_XK_F=2 -DGLA_XK_D=2", {"part": "xk", "c64": 2, "c128": 2}))
    out.append(("g3", "-DGLA_ZGG_F=3 -DGLA_ZGG_D=3", {"part": "gg", "c64": 3, "c128": 3}))
    out.append(("g7", "-DGLA_ZGG_F=7 -DGLA_ZGG_D=7", {"part": "gg", "c64": 7, "c128": 7}))
    out.append(("y1", "-DGLA_YK_F=1 -DGLA_YK_D=1", {"part": "yk", "c64": 1, "c128": 1}))
    out.append(("y2", "-DGLA_YK_F=2 -DGLA_YK_D=2", {"part": "yk", "c64": 2, "c128": 2}))
    for v in (0, 1, 2):
        out.append((f"l{v}", f"-DGLA_ZRELOAD_F={v} -DGLA_ZRELOAD_D={v}", {"part": "rl", "c64": v, "c128": v}))
    return out


def build(out, jobs, names=()):
    os.makedirs(out, exist_ok=True)
    for name, extra, _ in candidates():
        if names and name not in names:
            continue
        cmd = ["make", "-s", f"-j{jobs}", "-C", CSRC, f"BUILD={out}/obj_{name}", f"LIB={out}/lib_{name}.so",
               f"EXTRA={extra}"]
        if subprocess.run(cmd).returncode:
            sys.exit(f"build {name} failed")
        print(f"built {name}", flush=True)


def measure_one(name, lengths):
    """Runs in a child process with GILA_AMD_LIB set to one candidate."""
    import ctypes
    from fractions import Fraction
    import numpy as np
    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as ge
    gila = ge.load_package()
    from gila_amd import _capi
    desc = {n: d for n, _, d in candidates()}[name]
    part = desc["part"]

    def timed(N, typ):
        vol = gila.GlaVol(N, (Fraction(1, 32),) * 3)
        dims = _capi.gla_dims()
        _capi.check(gila.lib().gla_opr_dims_for(ctypes.byref(vol._c()), None, ctypes.byref(dims)))
        U = tuple(int(u) for u in dims.fur_ext)
        fur = [np.zeros(U + (6,), dtype=typ) for _ in range(8)]
        mem = gila.GlaOprMem(gila.GlaKerOpt(True), vol, egoFur=fur, setTyp=typ)
        n = 3 * int(np.prod(N))
        xd = torch.ones(n, dtype=torch.complex64 if typ == np.complex64 else torch.complex128, device="cuda")
        yd = torch.empty_like(xd)
        ms = (ctypes.c_double * 5)()
        by = (ctypes.c_int64 * 5)()
        st = torch.cuda.current_stream().cuda_stream
        _capi.check(gila.lib().gla_opr_profile(mem._h, ctypes.c_void_p(xd.data_ptr()), n,
                                               ctypes.c_void_p(yd.data_ptr()), n, 1, ctypes.c_void_p(st), 10, ms,
                                               by))
        torch.cuda.synchronize()
        del mem, xd, yd
        return [ms[i] * 1000.0 for i in range(5)]

    for n in lengths:
        for typ, dt in ((np.complex64, "c64"), (np.complex128, "c128")):
            rec = {"cand": name, "part": part, "dtype": dt, "n": n, "choice": desc[dt]}
            if part in ("zpass", "plan", "gg", "rl"):
                rec["large_us"] = timed((128, 128, n), typ)
            if part == "zpass":
                rec["small_us"] = timed((64, 64, n), typ)
            if part in ("rows", "plan", "xk", "yk"):
                rec["rows_us"] = timed((n, n, min(n, 128)), typ)
            print(json.dumps(rec), flush=True)


def measure(out, lengths, names):
    for name, _, _ in candidates():
        if names and name not in names:
            continue
        env = dict(os.environ, GILA_AMD_LIB=os.path.join(out, f"lib_{name}.so"))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "_one", "--cand", name,
                            "--lengths", ",".join(map(str, lengths))], env=env)
        if r.returncode:
            sys.exit(f"measurement with candidate {name} failed ({r.returncode})")


def _sub(src, tag, rows):
    body = "".join(r + "\n" for r in rows)
    return re.sub(rf"(// BEGIN TUNED {tag}\n)(.*?)([ \t]*// END TUNED {tag})",
                  lambda m: m.group(1) + body + m.group(3), src, flags=re.S)


def _current(src, tag, pat):
    """Entries of the table between the markers of `tag` in gla_tune.hpp."""
    m = re.search(rf"// BEGIN TUNED {tag}\n(.*?)// END TUNED {tag}", src, flags=re.S)
    out = {}
    for mm in re.finditer(pat, m.group(1) if m else ""):
        out[int(mm.group(1))] = tuple(int(v) for v in mm.groups()[1:])
    return out


def _decide(by, incumbent, margin):
    """The fastest choice when it beats the incumbent by more than `margin`
    (or the incumbent was not measured), else the incumbent."""
    best = min(by, key=by.get)
    return best if incumbent not in by or by[best] < (1 - margin) * by[incumbent] else incumbent


def table(results, margin):
    """Rewrite the tables of gla_tune.hpp from a results file.  Lengths with
    records are decided afresh (the incumbent stays unless beaten by more than
    `margin`); entries of lengths without records are kept; a part without
    records leaves its tables untouched; entries equal to the default are not
    written."""
    recs = [json.loads(line) for line in open(results) if line.strip()]
    lines = []
    src = open(os.path.join(CSRC, "gla_tune.hpp")).read()
    zpat = r"case (\d+): return \{(\d+), (\d+)\};"
    rpat = r"case (\d+): return RShape\{(\d+), (\d+), (\d+)\};"
    ppat = r"case (\d+): return (\d+);"

    def single(dt, part, tname, dflt, cost, what):
        nonlocal src
        byn = {}
        for r in recs:
            if r["part"] == part and r["dtype"] == dt:
                byn.setdefault(r["n"], {})[r["choice"]] = cost(r)
        if not byn:
            return
        vals = {n: v[0] for n, v in _current(src, tname, ppat).items()}
        for n in sorted(byn):
            d = vals.get(n, dflt)
            c = _decide(byn[n], d, margin)
            vals[n] = c
            lines.append(f"{dt} {part} {n}: {c} ({byn[n][c]:.1f} us {what}; current {d} {byn[n].get(d, 0):.1f})")
        src = _sub(src, tname, [f"            case {n}: return {v};" for n, v in sorted(vals.items()) if v != dflt])

    for dt, tag in (("c64", "C64"), ("c128", "C128")):
        # pass 3 by grid class: the large-grid table, then small-grid entries
        # where they differ from the large-grid shape
        zl, zs = {}, {}
        for r in recs:
            if r["part"] != "zpass" or r["dtype"] != dt:
                continue
            zl.setdefault(r["n"], {})[tuple(r["choice"])] = r["large_us"][2]
            zs.setdefault(r["n"], {})[tuple(r["choice"])] = r["small_us"][2]
        if zl:
            large = dict(_current(src, tag, zpat))
            small = dict(_current(src, tag + " SMALL", zpat))
            for n, by in zl.items():
                d = large.get(n, ZDEFAULT[dt])
                large[n] = _decide(by, d, margin)
                lines.append(f"{dt} zpass large {n}: {large[n]} {by[large[n]]:.1f} us (current {d} {by.get(d, 0):.1f})")
            for n, by in zs.items():
                ref = large[n]
                small[n] = _decide(by, ref, margin)
                lines.append(f"{dt} zpass small {n}: {small[n]} {by[small[n]]:.1f} us (large-grid shape {ref} "
                             f"{by.get(ref, 0):.1f})")
            src = _sub(src, tag, [f"            case {n}: return {{{s[0]}, {s[1]}}};"
                                  for n, s in sorted(large.items()) if s != ZDEFAULT[dt]])
            src = _sub(src, tag + " SMALL", [f"            case {n}: return {{{s[0]}, {s[1]}}};"
                                             for n, s in sorted(small.items())
                                             if s != large.get(n, ZDEFAULT[dt])])
        # rows: axis-1 passes (1 + 5) by ex, axis-2 passes (2 + 4) by (ey, pxy)
        rx, ry = {}, {}
        for r in recs:
            if r["part"] != "rows" or r["dtype"] != dt:
                continue
            ex, ey, pxy = r["choice"]
            t = r["rows_us"]
            rx.setdefault(r["n"], {})[ex] = min(t[0] + t[4], rx.get(r["n"], {}).get(ex, 1e30))
            ry.setdefault(r["n"], {})[(ey, pxy)] = t[1] + t[3]
        if rx:
            rows = dict(_current(src, "ROWS " + tag, rpat))
            for n in sorted(rx):
                d = rows.get(n, rdefault(dt, n))
                ex = _decide(rx[n], d[0], margin)
                ey = _decide(ry[n], (d[1], d[2]), margin)
                rows[n] = (ex, ey[0], ey[1])
                lines.append(f"{dt} rows {n}: ex {ex} ({rx[n][ex]:.1f} us), ey/pxy {ey} ({ry[n][ey]:.1f} us); "
                             f"current {d}")
            src = _sub(src, "ROWS " + tag, [f"            case {n}: return RShape{{{v[0]}, {v[1]}, {v[2]}}};"
                                            for n, v in sorted(rows.items()) if v != rdefault(dt, n)])
        # one value per length
        single(dt, "xk", "XK " + tag, 1, lambda r: r["rows_us"][0] + r["rows_us"][4], "passes 1+5")
        single(dt, "gg", "GG " + tag, 7 if dt == "c64" else 3, lambda r: r["large_us"][2], "pass 3")
        single(dt, "yk", "YK " + tag, 2, lambda r: r["rows_us"][1] + r["rows_us"][3], "passes 2+4")
        single(dt, "rl", "RELOAD " + tag, 0 if dt == "c64" else 2, lambda r: r["large_us"][2], "pass 3")
        # plan policy: pass 3 on the large grid plus the row passes
        single(dt, "plan", "PLAN " + tag, 0, lambda r: r["large_us"][2] + sum(r["rows_us"][i] for i in (0, 1, 3, 4)),
               "pass 3 + rows")
    open(os.path.join(CSRC, "gla_tune.hpp"), "w").write(src)
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["build", "measure", "table", "_one"])
    ap.add_argument("results", nargs="?")
    ap.add_argument("--out", default=os.path.join(HERE, "_tune"))
    ap.add_argument("--cand", default="")
    ap.add_argument("--only", default="", help="comma-separated candidate names to build / measure")
    ap.add_argument("--lengths", default=",".join(map(str, LENGTHS)))
    ap.add_argument("--margin", type=float, default=0.05)
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    lengths = [int(v) for v in a.lengths.split(",")]
    if a.cmd == "build":
        build(a.out, a.jobs, [s for s in a.only.split(",") if s])
    elif a.cmd == "measure":
        measure(a.out, lengths, [s for s in a.only.split(",") if s])
    elif a.cmd == "_one":
        measure_one(a.cand, lengths)
    else:
        table(a.results, a.margin)


if __name__ == "__main__":
    main()
