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


def table(results, margin):
    recs = [json.loads(line) for line in open(results) if line.strip()]
    lines = []
    src = open(os.path.join(CSRC, "gla_tune.hpp")).read()
    zpat = r"case (\d+): return \{(\d+), (\d+)\};"
    rpat = r"case (\d+): return RShape\{(\d+), (\d+), (\d+)\};"
    ppat = r"case (\d+): return (\d+);"
    for dt, tag in (("c64", "C64"), ("c128", "C128")):
        # pass 3, large class: the fastest shape where it beats the current entry
        zl, zs = {}, {}
        for r in recs:
            if r["part"] != "zpass" or r["dtype"] != dt:
                continue
            zl.setdefault(r["n"], {})[tuple(r["choice"])] = r["large_us"][2]
            zs.setdefault(r["n"], {})[tuple(r["choice"])] = r["small_us"][2]
        large, small = {}, {}
        cur = _current(src, tag, zpat)
        for n, by in zl.items():
            d = cur.get(n, ZDEFAULT[dt])
            best = min(by, key=by.get)
            large[n] = best if d not in by or by[best] < (1 - margin) * by[d] else d
            lines.append(f"{dt} zpass large {n}: {large[n]} {by[large[n]]:.1f} us (current {d} {by.get(d, 0):.1f})")
        for n, by in zs.items():
            ref = large[n]
            best = min(by, key=by.get)
            small[n] = best if ref not in by or by[best] < (1 - margin) * by[ref] else ref
            lines.append(f"{dt} zpass small {n}: {small[n]} {by[small[n]]:.1f} us (large-grid shape {ref} "
                         f"{by.get(ref, 0):.1f})")
        # a part without records in this results file keeps its table
        if zl:
            src = _sub(src, tag, [f"            case {n}: return {{{s[0]}, {s[1]}}};"
                                  for n, s in sorted(large.items()) if s != ZDEFAULT[dt]])
            src = _sub(src, tag + " SMALL", [f"            case {n}: return {{{s[0]}, {s[1]}}};"
                                             for n, s in sorted(small.items()) if s != large[n]])
        # rows: axis-1 passes (1 + 5) by ex, axis-2 passes (2 + 4) by (ey, pxy)
        rx, ry = {}, {}
        for r in recs:
            if r["part"] != "rows" or r["dtype"] != dt:
                continue
            ex, ey, pxy = r["choice"]
            t = r["rows_us"]
            rx.setdefault(r["n"], {})[ex] = min(t[0] + t[4], rx.get(r["n"], {}).get(ex, 1e30))
            ry.setdefault(r["n"], {})[(ey, pxy)] = t[1] + t[3]
        rows = []
        rcur = _current(src, "ROWS " + tag, rpat)
        for n in sorted(rx):
            d = rcur.get(n, rdefault(dt, n))
            bx = min(rx[n], key=rx[n].get)
            ex = bx if d[0] not in rx[n] or rx[n][bx] < (1 - margin) * rx[n][d[0]] else d[0]
            by_ = min(ry[n], key=ry[n].get)
            dy = (d[1], d[2])
            ey = by_ if dy not in ry[n] or ry[n][by_] < (1 - margin) * ry[n][dy] else dy
            lines.append(f"{dt} rows {n}: ex {ex} ({rx[n][ex]:.1f} us), ey/pxy {ey} ({ry[n][ey]:.1f} us); "
                         f"current {d}")
            if (ex, ey[0], ey[1]) != rdefault(dt, n):
                rows.append(f"            case {n}: return RShape{{{ex}, {ey[0]}, {ey[1]}}};")
        if rx:
            src = _sub(src, "ROWS " + tag, rows)
        # passes 1 and 5: parities together (2) or in turn (1)
        xk = {}
        for r in recs:
            if r["part"] != "xk" or r["dtype"] != dt:
                continue
            xk.setdefault(r["n"], {})[r["choice"]] = r["rows_us"][0] + r["rows_us"][4]
        xrow = []
        xcur = _current(src, "XK " + tag, ppat)
        for n in sorted(xk):
            d = xcur.get(n, (1,))[0]
            by = xk[n]
            best = min(by, key=by.get)
            c = best if d not in by or by[best] < (1 - margin) * by[d] else d
            lines.append(f"{dt} xk {n}: {c} ({by[c]:.1f} us passes 1+5; current {d} {by.get(d, 0):.1f})")
            if c != 1:
                xrow.append(f"            case {n}: return {c};")
        if xk:
            src = _sub(src, "XK " + tag, xrow)
        # pass 3: first-point G^ prefetch (7) or not (3)
        gg = {}
        for r in recs:
            if r["part"] != "gg" or r["dtype"] != dt:
                continue
            gg.setdefault(r["n"], {})[r["choice"]] = r["large_us"][2]
        grow = []
        gdef = 7 if dt == "c64" else 3
        gcur = _current(src, "GG " + tag, ppat)
        for n in sorted(gg):
            d = gcur.get(n, (gdef,))[0]
            by = gg[n]
            best = min(by, key=by.get)
            c = best if d not in by or by[best] < (1 - margin) * by[d] else d
            lines.append(f"{dt} gg {n}: {c} ({by[c]:.1f} us pass 3; current {d} {by.get(d, 0):.1f})")
            if c != gdef:
                grow.append(f"            case {n}: return {c};")
        if gg:
            src = _sub(src, "GG " + tag, grow)
        # one value per length: passes 2 + 4 for yk, pass 3 for rl
        for part, tname, dflt, cost in (("yk", "YK", 2, lambda r: r["rows_us"][1] + r["rows_us"][3]),
                                        ("rl", "RELOAD", 0 if dt == "c64" else 2, lambda r: r["large_us"][2])):
            byn = {}
            for r in recs:
                if r["part"] == part and r["dtype"] == dt:
                    byn.setdefault(r["n"], {})[r["choice"]] = cost(r)
            if not byn:
                continue
            cur = _current(src, f"{tname} {tag}", ppat)
            rows_ = []
            for n in sorted(byn):
                d = cur.get(n, (dflt,))[0]
                by = byn[n]
                best = min(by, key=by.get)
                c = best if d not in by or by[best] < (1 - margin) * by[d] else d
                lines.append(f"{dt} {part} {n}: {c} ({by[c]:.1f} us; current {d} {by.get(d, 0):.1f})")
                if c != dflt:
                    rows_.append(f"            case {n}: return {c};")
            src = _sub(src, f"{tname} {tag}", rows_)
        # plan policy: pass 3 on the large grid plus the row passes
        pol = {}
        for r in recs:
            if r["part"] != "plan" or r["dtype"] != dt:
                continue
            pol.setdefault(r["n"], {})[r["choice"]] = r["large_us"][2] + sum(r["rows_us"][i] for i in (0, 1, 3, 4))
        prow = []
        pcur = _current(src, "PLAN " + tag, ppat)
        for n in sorted(pol):
            t = pol[n]
            c = pcur.get(n, (0,))[0]
            o = 1 - c
            p = o if o in t and c in t and t[o] < (1 - margin) * t[c] else c
            lines.append(f"{dt} plan {n}: policy {p} (0: {t.get(0, 0):.1f} us, 1: {t.get(1, 0):.1f} us)")
            if p:
                prow.append(f"            case {n}: return 1;")
        if pol:
            src = _sub(src, "PLAN " + tag, prow)
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
