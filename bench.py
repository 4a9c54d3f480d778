"""Benchmark of the Green-function matvec (egoOpr!) on one MI355X.

Workload: self operator on 128x128x128 cubic cells of side 1/32 wavelength
(embedded grid 256^3 per component), one matvec of a seeded 3x128^3 complex
Gaussian field per step, ComplexF32 (the FP64 form is timed beside it).
Inputs are resident in HBM before the timed region.

Prints one JSON line.  Fields beyond the driver's contract:
  fp64            the ComplexF64 form on the same operator and input
  roofline        dominant kernel: algorithmic bytes per launch / its average
                  HIP-event duration, against the 8 TB/s HBM peak
  passes          every kernel's duration and algorithmic bytes
  cpu_baseline    the CPU oracle (a C restatement of the reference's host
                  algorithm) timed on this host on a bounded sample
  parity          relative L2 of both GPU forms against the FP64 oracle
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--counters", choices=["c64", "c128"], default=None,
                    help="short program for rocprofv3 counter passes: a 512 MiB device copy (the calibration "
                         "of FETCH_SIZE/WRITE_SIZE) and three matvecs of the given type; prints no JSON")
    return ap.parse_args()


def counters_program(a):
    """A fixed, short sequence of dispatches for --pmc runs: one copy of known
    size (512 MiB read, 512 MiB written), then three matvecs."""
    import torch
    import __graft_entry__ as ge
    gila = ge.load_package()
    from gila_amd import inputs
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    typ = np.complex64 if a.counters == "c64" else np.complex128
    N = (a.cells,) * 3
    mem = make_operator(gila, N, typ)
    x = inputs.complex_gaussian(N + (3,), seed=2).astype(typ)
    xd = torch.from_numpy(x.ravel(order="F").copy()).cuda()
    yd = torch.empty_like(xd)
    from gila_amd import _capi
    st = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        _capi.check(gila.lib().gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                             ctypes.c_void_p(yd.data_ptr()), yd.numel(), 1, ctypes.c_void_p(st)))
    torch.cuda.synchronize()
    print("counters program done", flush=True)


def make_operator(gila, N, typ):
    from fractions import Fraction
    vol = gila.GlaVol(N, (Fraction(1, 32),) * 3)
    return gila.GlaOprMem(gila.GlaKerOpt(True), vol, setTyp=typ)


def time_matvec(gila, mem, xd, yd, steps, warmup):
    import torch
    from gila_amd import _capi
    L = gila.lib()
    st = torch.cuda.current_stream()
    nx, ny = xd.numel(), yd.numel()

    def step():
        _capi.check(L.gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), nx, ctypes.c_void_p(yd.data_ptr()),
                                    ny, 1, ctypes.c_void_p(st.cuda_stream)))
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(st)
    for _ in range(steps):
        step()
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    return e0.elapsed_time(e1) / steps, wall


def profile(gila, mem, xd, yd, reps=20):
    import torch
    from gila_amd import _capi
    ms = (ctypes.c_double * 5)()
    by = (ctypes.c_int64 * 5)()
    st = torch.cuda.current_stream()
    _capi.check(gila.lib().gla_opr_profile(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                           ctypes.c_void_p(yd.data_ptr()), yd.numel(), 1,
                                           ctypes.c_void_p(st.cuda_stream), reps, ms, by))
    torch.cuda.synchronize()
    names = ["x_forward", "y_forward", "z_fused", "y_inverse", "x_inverse"]
    return [{"kernel": n, "ms": ms[i], "bytes": by[i], "gbs": by[i] / (ms[i] * 1e6)} for i, n in enumerate(names)]


def main():
    a = parse()
    if a.gpus > 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); multi-GPU runs are not part of this "
              "build", file=sys.stderr)
        sys.exit(2)
    if a.counters:
        counters_program(a)
        return
    import torch
    import __graft_entry__ as ge
    gila = ge.load_package()
    from gila_amd import inputs
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    N = (a.cells,) * 3
    x = inputs.complex_gaussian(N + (3,), seed=2)

    res = {}
    for typ in (np.complex64, np.complex128):
        log(f"setup {np.dtype(typ).name} {N}")
        t0 = time.perf_counter()
        mem = make_operator(gila, N, typ)
        setup_s = time.perf_counter() - t0
        log(f"setup done in {setup_s:.2f} s; timing {a.steps} steps")
        xd = torch.from_numpy(np.asfortranarray(x.astype(typ)).ravel(order="F").copy()).cuda()
        yd = torch.empty_like(xd)
        ms, wall = time_matvec(gila, mem, xd, yd, a.steps, a.warmup)
        log(f"timed {ms:.4f} ms/step; profiling passes")
        prof = profile(gila, mem, xd, yd)
        y = yd.cpu().numpy().reshape(N + (3,), order="F")
        res[typ] = dict(mem=mem, ms=ms, wall=wall, prof=prof, y=y, setup_s=setup_s)

    r32, r64 = res[np.complex64], res[np.complex128]
    tot_bytes = sum(p["bytes"] for p in r32["prof"])
    dom = max(r32["prof"], key=lambda p: p["ms"])
    out = {
        "metric": "G-matvecs/s (self Green matvec, 128^3 cells, ComplexF32)",
        "value": 1000.0 / r32["ms"],
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded SplitMix64 complex Gaussian input; Green data computed by the library's FP64 setup)",
        "config": {"workload": f"self Green operator, {N[0]}x{N[1]}x{N[2]} cells of side 1/32 wavelength, "
                               f"embedded {2 * N[0]}^3 per component, 3 components, one matvec per step",
                   "cells": list(N), "cell_side_wavelengths": "1/32", "embedded": [2 * n for n in N],
                   "green_setup_s": r32["setup_s"]},
        "hbm_gbs_whole_matvec": tot_bytes / (r32["ms"] * 1e6),
        "bytes_alg_per_matvec": tot_bytes,
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": None},
        "passes": r32["prof"],
        "fp64": {"value": 1000.0 / r64["ms"], "unit": "matvecs/s", "ms_per_step": r64["ms"],
                 "hbm_gbs_whole_matvec": sum(p["bytes"] for p in r64["prof"]) / (r64["ms"] * 1e6),
                 "passes": r64["prof"]},
        "fp32_over_fp64": r64["ms"] / r32["ms"],
        "wall_s_timed_region": r32["wall"],
    }
    pmc = os.path.join(ROOT, "profiles", "traffic_z_fused_128.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            tr = json.load(f)
        out["roofline"]["traffic"] = tr.get(dom["kernel"])

    if not a.no_cpu:
        log("cpu leg: FP64 oracle matvec")
        orc = ge.load_oracle()     # the checker: CPU baseline and parity only
        fur64 = np.concatenate([f.reshape(-1, order="F") for f in r64["mem"].egoFur])
        sym = (1, 1, 1)
        y_ref = orc.matvec(N, N, sym, fur64, x)
        fur32 = np.concatenate([f.reshape(-1, order="F") for f in r32["mem"].egoFur])
        x32 = x.astype(np.complex64)
        orc.matvec(N, N, sym, fur32, x32, dtype=np.complex64)   # warm-up
        ts = []
        log("cpu leg: timing the FP32 restatement")
        for _ in range(a.cpu_reps):
            t0 = time.perf_counter()
            orc.matvec(N, N, sym, fur32, x32, dtype=np.complex64)
            ts.append(time.perf_counter() - t0)
        med = float(np.median(ts))
        out["cpu_baseline"] = {"value": 1.0 / med, "unit": "matvecs/s", "cores": orc.num_threads(), "kind": "port",
                               "sample": f"{a.cpu_reps} matvecs (after 1 warm-up) of the same 128^3 ComplexF32 "
                                         f"workload, C restatement of egoBrnHst! (oracle/gla_oracle.c), median "
                                         f"{med:.3f} s"}
        out["parity"] = {"c64_rel_l2_vs_fp64_oracle": orc.rel_l2(r32["y"], y_ref),
                         "c128_rel_l2_vs_fp64_oracle": orc.rel_l2(r64["y"], y_ref)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
