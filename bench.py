"""Benchmark of the Green-function matvec (egoOpr!) on one MI355X.

Workload: self operator on 128x128x128 cubic cells of side 1/32 wavelength
(embedded grid 256^3 per component), one matvec of a seeded 3x128^3 complex
Gaussian field per step, ComplexF32 (the FP64 form is timed beside it).
Inputs are resident in HBM before the timed region.

Prints one JSON line.  Fields beyond the driver's contract:
  fp64            the ComplexF64 form on the same operator and input
  roofline        dominant kernel: algorithmic bytes per launch / its average
                  duration between HIP events at its boundaries inside the
                  timed region, against the 8 TB/s HBM peak
  passes          every kernel's duration (the same events), algorithmic
                  bytes, and its rate as a fraction of copy_gbs_same_gpu
  copy_gbs_same_gpu  a plain 1 GiB device-to-device copy in the same run
                  (bytes read + written per second)
  cpu_baseline    the CPU oracle (a C restatement of the reference's host
                  algorithm) timed on this host on a bounded sample
  parity          relative L2 of both GPU forms against the FP64 oracle
  external        the external operator with uneven volumes (256x34x34
                  source inside a 256x256x128 target), both precisions
  batched         64 right-hand sides on a 96^3 operator in one launch, and
                  a mixed batch of 16 operators replayed from a HIP graph
  serving         the served solve stream against sequential FP64 solves
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
    # about 70 ms of untimed matvecs: after a handful the chip has not reached
    # its sustained state yet (128^3 FP32: 365 us per matvec after 5, 346 us
    # after 200)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-external", action="store_true", help="skip the external-operator leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched right-hand-side leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served solve-stream leg")
    ap.add_argument("--serve-requests", type=int, default=500,
                    help="requests of the served stream at each rate (a rate may carry its own count, rate:count)")
    ap.add_argument("--serve-seq-requests", type=int, default=40,
                    help="bounded sample of the stream solved one at a time in plain FP64 (the baseline rate)")
    ap.add_argument("--serve-rates", default="4",
                    help="Poisson arrival rates of the served stream (comma-separated, each optionally rate:count), "
                         "in multiples of the measured sequential FP64 rate; the last is the headline")
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
    """Two timed regions of `steps` matvecs each, issued as gla_opr_apply
    issues them on torch's current stream, each bracketed by device
    synchronizations: (1) plain launches between two events -- ms per
    matvec, the headline; (2) the same sequence with HIP events on that
    stream at every pass boundary (gla_opr_profile_seq) -- each pass's
    average duration inside the sequence, and the evented ms per matvec (the
    markers cost a few percent, so they stay out of the headline region).
    Returns (ms, wall, passes, evented ms)."""
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
    ms_plain = e0.elapsed_time(e1) / steps
    ms = (ctypes.c_double * 5)()
    by = (ctypes.c_int64 * 5)()
    tot = ctypes.c_double()
    _capi.check(L.gla_opr_profile_seq(mem._h, ctypes.c_void_p(xd.data_ptr()), nx, ctypes.c_void_p(yd.data_ptr()),
                                      ny, 1, ctypes.c_void_p(st.cuda_stream), steps, ms, by, ctypes.byref(tot)))
    torch.cuda.synchronize()
    names = ["x_forward", "y_forward", "z_fused", "y_inverse", "x_inverse"]
    passes = [{"kernel": n, "ms": ms[i], "bytes": by[i], "gbs": by[i] / (ms[i] * 1e6)} for i, n in enumerate(names)]
    return ms_plain, wall, passes, tot.value


def event_ms(fn, reps, stream):
    """Average milliseconds of fn() over reps calls, by events on `stream`."""
    import torch
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(reps):
        fn()
    t1.record(stream)
    t1.synchronize()
    return t0.elapsed_time(t1) / reps


def batched_leg(gila, log):
    """96^3 self operator, 64 right-hand sides in one launch (32 plane waves,
    32 point dipoles), and a mixed batch of 16 operators replayed from one
    HIP graph."""
    import torch
    from fractions import Fraction
    from gila_amd import _capi, inputs
    out = {}
    N = (96, 96, 96)
    nrhs = 64
    log("batched: setup 96^3 c64")
    mem = make_operator(gila, N, np.complex64)
    fields = []
    for i in range(32):
        khat = inputs.random_unit(7000 + i)
        p = inputs.random_unit(8000 + i)
        p = p - khat * (p @ khat)
        fields.append(inputs.plane_wave(N, (Fraction(1, 32),) * 3, khat, p / np.linalg.norm(p)))
    u = inputs.uniform(9000, 3 * 32)
    for i in range(32):
        cell = [int(u[3 * i + d] * N[d]) for d in range(3)]
        fields.append(inputs.point_dipole(N, cell, inputs.random_unit(9100 + i)))
    x = np.stack([f.reshape(-1, order="F") for f in fields]).astype(np.complex64)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    st = torch.cuda.current_stream()
    lib = gila.lib()

    def launch(n=nrhs, xx=xd, yy=yd):
        _capi.check(lib.gla_opr_apply(mem._h, ctypes.c_void_p(xx.data_ptr()), xx[:n].numel(),
                                      ctypes.c_void_p(yy.data_ptr()), yy[:n].numel(), n,
                                      ctypes.c_void_p(st.cuda_stream)))
    launch()
    torch.cuda.synchronize()
    ms = event_ms(launch, 5, st)
    per_rhs = 8 * (78 * 96 ** 3 + 6 * 97 ** 3)
    # RHS 0 and 63 alone give bitwise the batch's columns
    y1 = torch.empty_like(yd[:1])
    same = True
    for r in (0, nrhs - 1):
        launch(1, xd[r:r + 1], y1)
        torch.cuda.synchronize()
        same = same and torch.equal(y1[0], yd[r])
    out["rhs64"] = {"workload": "96^3 self operator, ComplexF32, 64 right-hand sides (32 plane waves, 32 point "
                                "dipoles) in one launch",
                    "ms_per_launch": ms, "rhs_matvecs_per_s": nrhs * 1000.0 / ms,
                    "hbm_gbs_alg": nrhs * per_rhs / (ms * 1e6),
                    "ideal_ms_green_once_8tbs": (nrhs * (per_rhs - 8 * 6 * 97 ** 3) + 8 * 6 * 97 ** 3) / 8e9,
                    "single_rhs_bitwise_equal": bool(same)}
    del xd, yd, y1, mem
    torch.cuda.empty_cache()

    shapes = [(24, 24, 24), (32, 48, 64), (64, 64, 64), (80, 40, 40)]
    uu = inputs.uniform(4242, 16)
    ops = []
    for i in range(16):
        cells = shapes[int(uu[i] * len(shapes))]
        m = make_operator(gila, cells, np.complex64)
        xi = torch.from_numpy(inputs.complex_gaussian(cells + (3,), seed=500 + i).astype(np.complex64)
                              .reshape(-1, order="F").copy()).cuda()
        ops.append((cells, m, xi, torch.empty_like(xi)))
    log("batched: mixed batch of 16 operators " + ",".join("x".join(map(str, o[0])) for o in ops))

    def seq(stream):
        for cells, m, xi, yi in ops:
            _capi.check(lib.gla_opr_apply(m._h, ctypes.c_void_p(xi.data_ptr()), xi.numel(),
                                          ctypes.c_void_p(yi.data_ptr()), yi.numel(), 1,
                                          ctypes.c_void_p(stream.cuda_stream)))
    seq(st)
    torch.cuda.synchronize()
    eager = [o[3].clone() for o in ops]
    ms_eager = event_ms(lambda: seq(st), 10, st)
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(st)
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            seq(cap)
    st.wait_stream(cap)
    torch.cuda.synchronize()
    ok = True
    for _ in range(3):
        for o in ops:
            o[3].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        ok = ok and all(torch.equal(o[3], e) for o, e in zip(ops, eager))
    ms_graph = event_ms(g.replay, 10, st)
    # the library's recorded batch: one graph whose operators are dealt over
    # four parallel branches (gla_batch_create)
    n = len(ops)
    b = ctypes.c_void_p()
    _capi.check(lib.gla_batch_create((ctypes.c_void_p * n)(*[o[1]._h for o in ops]),
                                     (ctypes.c_void_p * n)(*[o[2].data_ptr() for o in ops]),
                                     (ctypes.c_void_p * n)(*[o[3].data_ptr() for o in ops]), None, n, 4,
                                     ctypes.byref(b)))
    launch_b = lambda: _capi.check(lib.gla_batch_launch(b, ctypes.c_void_p(st.cuda_stream)))
    ok_b = True
    for _ in range(3):
        for o in ops:
            o[3].zero_()
        torch.cuda.synchronize()
        launch_b()
        torch.cuda.synchronize()
        ok_b = ok_b and all(torch.equal(o[3], e) for o, e in zip(ops, eager))
    ms_batch = event_ms(launch_b, 10, st)
    _capi.check(lib.gla_batch_destroy(b))
    out["mixed16"] = {"workload": "16 self operators drawn (seeded) from 24^3, 32x48x64, 64^3, 80x40x40, "
                                  "ComplexF32, one matvec each",
                      "shapes": ["x".join(map(str, o[0])) for o in ops],
                      "ms_eager": ms_eager, "ms_hipgraph": ms_graph,
                      "matvecs_per_s_hipgraph": 16 * 1000.0 / ms_graph,
                      "graph_replays_bitwise_equal_eager": bool(ok),
                      "ms_recorded_batch_4_lanes": ms_batch,
                      "matvecs_per_s_recorded_batch": 16 * 1000.0 / ms_batch,
                      "recorded_batch_bitwise_equal_eager": bool(ok_b)}
    del ops, g
    torch.cuda.empty_cache()
    return out


def model_bytes(S, T, sym, s):
    """Algorithmic bytes of one matvec of the pruned five-pass pipeline on the
    reference's grid L = S + T (not a longer device grid a transform length
    is embedded in): s [3 (S + 2 L1 S2 S3 + 2 L1 L2 S3 + 2 L1 L2 T3 +
    2 L1 T2 T3 + T) + 6 U1 U2 U3], U = S + 1 on a symmetric axis, else L."""
    L = [a + b for a, b in zip(S, T)]
    U = [S[d] + 1 if sym[d] else L[d] for d in range(3)]
    f = (S[0] * S[1] * S[2] + 2 * L[0] * S[1] * S[2] + 2 * L[0] * L[1] * S[2] + 2 * L[0] * L[1] * T[2]
         + 2 * L[0] * T[1] * T[2] + T[0] * T[1] * T[2])
    return s * (3 * f + 6 * U[0] * U[1] * U[2])


def external_leg(gila, log, steps=20, warmup=50):
    """External operator with uneven volumes: a 256x34x34 source centred in a
    256x256x128 target at 1/32 wavelength (B = (256,145,81); built from
    geometry, axes 2 and 3 run on the next fast transform lengths 160 and 96,
    while bytes_alg_per_matvec and the GB/s count the bytes of the
    reference's 512x290x162 grid).  Input: an x-polarised point
    dipole in a seeded random source cell plus 1e-3 complex Gaussian noise.
    ComplexF32 and ComplexF64, Green data from the library's FP64 setup; the
    same two timed regions as the headline matvec."""
    import torch
    from fractions import Fraction
    from gila_amd import inputs
    S, T = (256, 34, 34), (256, 256, 128)
    h = (Fraction(1, 32),) * 3
    src, tgt = gila.GlaVol(S, h), gila.GlaVol(T, h)
    u = inputs.uniform(77, 3)
    cell = [int(u[d] * S[d]) for d in range(3)]
    x = 1e-3 * inputs.complex_gaussian(S + (3,), seed=78)
    x[tuple(cell) + (0,)] += 1.0
    out = {"workload": "external operator, source 256x34x34 centred in a 256x256x128 target, cells of side 1/32 "
                       "wavelength, one matvec per step (x-polarised point dipole plus 1e-3 noise)"}
    for typ, key in ((np.complex64, "c64"), (np.complex128, "c128")):
        log(f"external: setup {key}")
        t0 = time.perf_counter()
        mem = gila.GlaOprMem(gila.GlaKerOpt(True), tgt, src, setTyp=typ)
        setup_s = time.perf_counter() - t0
        xd = torch.from_numpy(np.asfortranarray(x.astype(typ)).ravel(order="F").copy()).cuda()
        yd = torch.empty(3 * T[0] * T[1] * T[2], dtype=xd.dtype, device=xd.device)
        ms, _, prof, _ = time_matvec(gila, mem, xd, yd, steps, warmup)
        dev_by = sum(q["bytes"] for q in prof)
        by = model_bytes(S, T, mem._dims.sym, np.dtype(typ).itemsize)
        out[key] = {"ms_per_matvec": ms, "matvecs_per_s": 1000.0 / ms, "bytes_alg_per_matvec": by,
                    "hbm_gbs_whole_matvec": by / (ms * 1e6), "branch_lengths": list(mem._dims.brn),
                    "device_lengths": list(mem._dims.dev_brn), "device_grid_bytes_per_matvec": dev_by,
                    "green_setup_s": setup_s, "passes": prof}
        log(f"external {key}: {ms:.4f} ms/matvec")
        del mem, xd, yd
        torch.cuda.empty_cache()
    out["fp32_over_fp64"] = out["c128"]["ms_per_matvec"] / out["c64"]["ms_per_matvec"]
    return out


def serve_leg(gila, a, log):
    """Served solve stream against sequential plain FP64 solves (GMRES(30) on
    (I - chi G) E = E_inc to 1e-6 by the FP64 residual): the sequential rate
    from a bounded sample of the stream solved one at a time, then the stream
    through the server with Poisson arrivals at each rate of the sweep (one
    warm server, operators resident across the sweep)."""
    S = gila.serve
    shapes = [(32, 32, 32), (64, 64, 48), (128, 64, 64), (128, 128, 128), (256, 128, 64)]
    from gila_amd import inputs
    rates = []
    for tok in a.serve_rates.split(","):
        if tok.strip():
            r, _, n = tok.partition(":")
            rates.append((float(r), int(n) if n else a.serve_requests))
    nmax = max([n for _, n in rates] + [a.serve_seq_requests])
    u = inputs.uniform(31337, nmax * (1 + len(rates)))

    def requests(n):
        return [S.random_request(20000 + i, shapes[int(u[i] * len(shapes))], tol=1e-6) for i in range(n)]
    log(f"serve: operators for {len(shapes)} grids (setup, not timed)")
    seq = S.Server(inner="fp64", max_batch=1, max_active=1, max_cycles=60)
    srv = S.Server(inner="fp32", max_batch=16, max_cycles=60)
    for sh in shapes:
        seq._grid_for(S.random_request(1, sh))
        srv._grid_for(S.random_request(1, sh))
    reqs = requests(a.serve_seq_requests)
    log(f"serve: {len(reqs)} sequential plain FP64 solves (the first requests of the stream)")
    t0 = time.perf_counter()
    seq.run(reqs, on_done=lambda r: log(f"  seq {r.rid} {'x'.join(map(str, r.cells))} steps={r.steps}"))
    t_seq = time.perf_counter() - t0
    rate_seq = len(reqs) / t_seq
    st_seq = S.latency_stats(reqs)
    sweep = []
    for k, (mult, nreq) in enumerate(rates):
        lam = mult * rate_seq
        uu = u[nmax * (1 + k):nmax * (1 + k) + nreq]
        gaps = -np.log(uu) / lam
        arrivals = np.cumsum(gaps) - gaps[0]
        reqs2 = requests(nreq)
        log(f"serve: {nreq} requests, Poisson arrivals at {lam:.2f}/s ({mult}x the sequential rate)")
        srv.run(reqs2, arrivals=list(arrivals),
                on_done=lambda r: log(f"  srv {r.rid} {'x'.join(map(str, r.cells))} steps={r.steps} "
                                      f"lat={r.latency:.2f}s"))
        st = S.latency_stats(reqs2)
        # the requests solved both ways agree (FP32 inner solves with FP64
        # refinement against plain FP64)
        agree = max(float(np.linalg.norm(r2.result - r1.result) / np.linalg.norm(r1.result))
                    for r1, r2 in zip(reqs, reqs2))
        span_arr = float(arrivals[-1] - arrivals[0]) if nreq > 1 else 0.0
        sweep.append({"rate_x_sequential": mult, "requests": nreq, "arrival_rate_per_s": lam,
                      "arrival_span_s": span_arr,
                      "realised_arrival_rate_per_s": (nreq - 1) / span_arr if span_arr > 0 else None,
                      "solves_per_s": st["solves_per_s"],
                      "p50_latency_s": st["p50_s"], "p99_latency_s": st["p99_s"],
                      "speedup_vs_sequential_fp64": st["solves_per_s"] / rate_seq,
                      "max_rel_diff_served_vs_sequential": agree,
                      "all_converged": bool(all(r.converged for r in reqs2)),
                      "max_relres": max(r.relres for r in reqs2), "steps_total": sum(r.steps for r in reqs2)})
    head = sweep[-1]
    return {"workload": f"{head['requests']} requests, domains drawn (seeded) from "
                        + ", ".join("x".join(map(str, s)) for s in shapes)
                        + " cells at 1/32 wavelength; chi 0 outside a random box of 10-40% of the domain, "
                          "U[0.5,3]+i U[0,0.1] inside; E_inc a plane wave at a random angle; Poisson arrivals "
                          "at each rate of the sweep (headline: the last)",
            "requests": head["requests"],
            "solves_per_s": head["solves_per_s"], "p50_latency_s": head["p50_latency_s"],
            "p99_latency_s": head["p99_latency_s"], "arrival_rate_per_s": head["arrival_rate_per_s"],
            "sequential_fp64_solves_per_s": rate_seq, "sequential_fp64_sample_requests": len(reqs),
            "sequential_fp64_p50_latency_s": st_seq["p50_s"],
            "sequential_fp64_all_converged": bool(all(r.converged for r in reqs)),
            "speedup_vs_sequential_fp64": head["speedup_vs_sequential_fp64"],
            "max_rel_diff_served_vs_sequential": max(r["max_rel_diff_served_vs_sequential"] for r in sweep),
            "all_converged": bool(all(r["all_converged"] for r in sweep)),
            "max_relres": max(r["max_relres"] for r in sweep),
            "sweep": sweep}


def copy_gbs(torch, nbytes=1 << 30, reps=20):
    """Streaming bandwidth of a plain device-to-device copy of nbytes on this
    GPU in the same run (bytes read + written per second), the practical
    ceiling a bandwidth-bound pass is compared with beside the 8 TB/s peak."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * nbytes / (ms * 1e6)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(a, ge, N, x, r32, r64):
    """The CPU oracle -- the C restatement of the reference's host matvec
    egoBrnHst! (oracle/gla_oracle.c), OpenMP over all the host threads it is
    given -- timed on this host: the headline 128^3 ComplexF32 workload on all
    threads (median of --cpu-reps after one warm-up), the same in ComplexF64,
    the 128^3 ComplexF32 matvec on one thread (3 runs), and the reference's
    CPU configuration, a 16^3 ComplexF64 self operator (all threads and one
    thread, median of --cpu-reps each).  Returns (cpu_baseline, parity)."""
    orc = ge.load_oracle()     # the checker: CPU baseline and parity only
    sym = (1, 1, 1)
    fur64 = np.concatenate([f.reshape(-1, order="F") for f in r64["mem"].egoFur])
    fur32 = np.concatenate([f.reshape(-1, order="F") for f in r32["mem"].egoFur])
    x32 = x.astype(np.complex64)
    nthr = orc.num_threads()

    def timed(fn, reps):
        fn()   # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))
    log("cpu leg: FP64 oracle matvec (parity reference)")
    y_ref = orc.matvec(N, N, sym, fur64, x)
    log(f"cpu leg: 128^3 ComplexF32 on {nthr} threads")
    t32 = timed(lambda: orc.matvec(N, N, sym, fur32, x32, dtype=np.complex64), a.cpu_reps)
    log(f"cpu leg: 128^3 ComplexF64 on {nthr} threads")
    t64 = timed(lambda: orc.matvec(N, N, sym, fur64, x), a.cpu_reps)
    log("cpu leg: 128^3 ComplexF32 on 1 thread")
    orc.set_threads(1)
    try:
        t32_1 = timed(lambda: orc.matvec(N, N, sym, fur32, x32, dtype=np.complex64), 3)
        n16 = (16, 16, 16)
        fur16, _ = orc.self_operator_fur(n16, model=0)
        x16 = orc.complex_gaussian(n16 + (3,), seed=1)
        t16_1 = timed(lambda: orc.matvec(n16, n16, sym, fur16, x16), a.cpu_reps)
    finally:
        orc.set_threads(nthr)
    t16 = timed(lambda: orc.matvec(n16, n16, sym, fur16, x16), a.cpu_reps)
    try:
        online = len(os.sched_getaffinity(0))
    except AttributeError:
        online = os.cpu_count()
    base = {"value": 1.0 / t32, "unit": "matvecs/s", "cores": nthr, "kind": "port",
            "sample": f"128^3 ComplexF32 self matvec, the same workload and Green data as the GPU headline: C "
                      f"restatement of egoBrnHst! (oracle/gla_oracle.c), median of {a.cpu_reps} after one warm-up, "
                      f"{t32:.3f} s per matvec",
            "cpu_model": cpu_model(), "host_cores_online": online,
            "c128_128": {"matvecs_per_s": 1.0 / t64, "s_per_matvec": t64, "threads": nthr},
            "c64_128_one_thread": {"matvecs_per_s": 1.0 / t32_1, "s_per_matvec": t32_1, "threads": 1},
            "c128_16_reference_cpu_config": {"s_per_matvec": t16, "threads": nthr, "s_per_matvec_one_thread": t16_1,
                                             "workload": "16^3 self operator of 1/32-wavelength cells, ComplexF64, "
                                                         "seeded complex Gaussian input"}}
    parity = {"c64_rel_l2_vs_fp64_oracle": orc.rel_l2(r32["y"], y_ref),
              "c128_rel_l2_vs_fp64_oracle": orc.rel_l2(r64["y"], y_ref)}
    return base, parity


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
        ms, wall, prof, ms_ev = time_matvec(gila, mem, xd, yd, a.steps, a.warmup)
        log(f"timed {ms:.4f} ms/step ({ms_ev:.4f} with pass-boundary events)")
        y = yd.cpu().numpy().reshape(N + (3,), order="F")
        res[typ] = dict(mem=mem, ms=ms, wall=wall, prof=prof, y=y, setup_s=setup_s, ms_ev=ms_ev)

    # the host-array entry (gla_opr_apply_host: device buffers allocated,
    # PCIe copies in and out), for the record; never the headline value
    xh = np.asfortranarray(x.astype(np.complex64))
    yh = np.empty_like(xh)
    m32 = res[np.complex64]["mem"]
    from gila_amd import _capi
    hts = []
    for _ in range(3):
        t0 = time.perf_counter()
        _capi.check(gila.lib().gla_opr_apply_host(m32._h, xh.ctypes.data_as(ctypes.c_void_p), xh.size,
                                                  yh.ctypes.data_as(ctypes.c_void_p), yh.size, 1))
        hts.append(time.perf_counter() - t0)
    host_ms = 1000.0 * float(np.median(hts))
    r32, r64 = res[np.complex64], res[np.complex128]
    cgbs = copy_gbs(torch)
    for r in (r32, r64):
        for p in r["prof"]:
            p["frac_of_copy"] = p["gbs"] / cgbs
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
        "copy_gbs_same_gpu": cgbs,
        "ms_per_step_with_pass_events": r32["ms_ev"],
        "fp64": {"value": 1000.0 / r64["ms"], "unit": "matvecs/s", "ms_per_step": r64["ms"],
                 "hbm_gbs_whole_matvec": sum(p["bytes"] for p in r64["prof"]) / (r64["ms"] * 1e6),
                 "passes": r64["prof"]},
        "fp32_over_fp64": r64["ms"] / r32["ms"],
        "host_buffers_ms_per_matvec": host_ms,
        "wall_s_timed_region": r32["wall"],
    }
    pmc = os.path.join(ROOT, "profiles", "traffic_z_fused_128.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            tr = json.load(f)
        out["roofline"]["traffic"] = tr.get(dom["kernel"])

    if not a.no_cpu:
        out["cpu_baseline"], out["parity"] = cpu_leg(a, ge, N, x, r32, r64)
    del res, r32, r64
    torch.cuda.empty_cache()
    if not a.no_external:
        out["external"] = external_leg(gila, log)
    if not a.no_batched:
        out["batched"] = batched_leg(gila, log)
    if not a.no_serve:
        out["serving"] = serve_leg(gila, a, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
