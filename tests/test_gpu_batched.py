"""The batched workloads at their full size: 64 right-hand sides (32 plane
waves, 32 point dipoles) on a 96^3 operator in one launch, and a mixed batch
of 16 operators of uneven volume replayed from one captured HIP graph.

Columns of the 64-RHS launch must equal single-RHS launches bit for bit (the
Green data is shared, every column's arithmetic is the same), and sampled
columns match the FP64 oracle fed the same (FP32-rounded) Green data to the
ComplexF32 bound 1e-5.  Graph replays must equal the eager sequence bit for
bit on every replay, with eager work and synchronizations between them."""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

H = (Fraction(1, 32),) * 3


def _apply(gila, mem, xd, yd, nrhs, stream):
    from gila_amd import _capi
    _capi.check(gila.lib().gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                         ctypes.c_void_p(yd.data_ptr()), yd.numel(), nrhs,
                                         ctypes.c_void_p(stream.cuda_stream)))


def _fields_96():
    from gila_amd import inputs
    N = (96, 96, 96)
    fields = []
    for i in range(32):
        khat = inputs.random_unit(7000 + i)
        p = inputs.random_unit(8000 + i)
        p = p - khat * (p @ khat)
        fields.append(inputs.plane_wave(N, H, khat, p / np.linalg.norm(p)))
    u = inputs.uniform(9000, 3 * 32)
    for i in range(32):
        cell = [int(u[3 * i + d] * N[d]) for d in range(3)]
        fields.append(inputs.point_dipole(N, cell, inputs.random_unit(9100 + i)))
    return N, fields


@pytest.mark.gpu
def test_64_rhs_launch_at_96_cubed(gila, orc, gpu):
    import torch
    print("case 96^3 x 64 right-hand sides, ComplexF32", flush=True)
    N, fields = _fields_96()
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), gila.GlaVol(N, H), setTyp=np.complex64)
    x = np.stack([f.reshape(-1, order="F") for f in fields]).astype(np.complex64)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    st = torch.cuda.current_stream()
    _apply(gila, mem, xd, yd, 64, st)
    y1 = torch.empty_like(yd[:1])
    for r in range(64):
        _apply(gila, mem, xd[r:r + 1], y1, 1, st)
        torch.cuda.synchronize()
        assert torch.equal(y1[0], yd[r]), r
    fur = np.concatenate([f.reshape(-1, order="F") for f in mem.egoFur]).astype(np.complex128)
    yh = yd.cpu().numpy()
    for r in (0, 31, 32, 63):       # two plane waves, two dipoles
        y_ref = orc.matvec(N, N, (1, 1, 1), fur, x[r].astype(np.complex128).reshape(N + (3,), order="F"))
        assert orc.rel_l2(yh[r].reshape(N + (3,), order="F"), y_ref) <= 1e-5, r


@pytest.mark.gpu
def test_mixed_batch_graph_replay(gila, gpu):
    import torch
    from gila_amd import inputs
    shapes = [(24, 24, 24), (32, 48, 64), (64, 64, 64), (80, 40, 40)]
    uu = inputs.uniform(4242, 16)
    ops = []
    for i in range(16):
        cells = shapes[int(uu[i] * len(shapes))]
        m = gila.GlaOprMem(gila.GlaKerOpt(True), gila.GlaVol(cells, H), setTyp=np.complex64)
        xi = torch.from_numpy(inputs.complex_gaussian(cells + (3,), seed=500 + i).astype(np.complex64)
                              .reshape(-1, order="F").copy()).cuda()
        ops.append((m, xi, torch.empty_like(xi)))
    print("case mixed batch of 16 operators, graph replay", flush=True)
    st = torch.cuda.current_stream()

    def seq(stream):
        for m, xi, yi in ops:
            _apply(gila, m, xi, yi, 1, stream)
    seq(st)        # eager, also grows every handle's work buffers
    torch.cuda.synchronize()
    eager = [o[2].clone() for o in ops]
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(st)
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            seq(cap)
    st.wait_stream(cap)
    torch.cuda.synchronize()
    for rep in range(3):
        for o in ops:
            o[2].zero_()
        seq(st)                    # eager work between replays
        torch.cuda.synchronize()
        for o in ops:
            o[2].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, (o, e) in enumerate(zip(ops, eager)):
            assert torch.equal(o[2], e), (rep, k)


@pytest.mark.gpu
def test_apply_under_capture_needs_reserved_buffers(gila, gpu):
    import torch
    from gila_amd import _capi
    N = (16, 16, 16)
    m = gila.GlaOprMem(gila.GlaKerOpt(True), gila.GlaVol(N, H), setTyp=np.complex64)
    n = 3 * 16 ** 3
    xd = torch.ones(4 * n, dtype=torch.complex64, device="cuda")
    yd = torch.empty_like(xd)
    st = torch.cuda.current_stream()
    _apply(gila, m, xd[:n], yd[:n], 1, st)
    torch.cuda.synchronize()
    # four right-hand sides under capture would grow the buffers: refused
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(st)
    status = []
    dummy = torch.zeros(1, device="cuda")
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            dummy.add_(1)          # keeps the captured graph non-empty
            status.append(gila.lib().gla_opr_apply(m._h, ctypes.c_void_p(xd.data_ptr()), 4 * n,
                                                   ctypes.c_void_p(yd.data_ptr()), 4 * n, 4,
                                                   ctypes.c_void_p(cap.cuda_stream)))
    torch.cuda.synchronize()
    assert status == [_capi.GLA_ERR_ARG]
    assert "reserve" in gila.lib().gla_last_error().decode()
    # after gla_opr_reserve the same capture works and replays equal eager
    _capi.check(gila.lib().gla_opr_reserve(m._h, 4))
    _apply(gila, m, xd, yd, 4, st)
    torch.cuda.synchronize()
    ref = yd.clone()
    g2 = torch.cuda.CUDAGraph()
    cap.wait_stream(st)
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g2, stream=cap):
            _apply(gila, m, xd, yd, 4, cap)
    st.wait_stream(cap)
    yd.zero_()
    torch.cuda.synchronize()
    g2.replay()
    torch.cuda.synchronize()
    assert torch.equal(yd, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 4])
def test_recorded_batch_equals_eager(gila, gpu, lanes):
    """gla_batch_create / gla_batch_launch: the mixed batch of 16 operators
    recorded once over `lanes` parallel branches; every replay (with eager
    work and synchronizations between replays) equals the eager sequence bit
    for bit; an operator listed twice shares its work buffers safely."""
    import torch
    from gila_amd import _capi, inputs
    shapes = [(24, 24, 24), (32, 48, 64), (64, 64, 64), (80, 40, 40)]
    uu = inputs.uniform(4242, 16)
    mems, xs, ys = [], [], []
    for i in range(16):
        cells = shapes[int(uu[i] * len(shapes))]
        mems.append(gila.GlaOprMem(gila.GlaKerOpt(True), gila.GlaVol(cells, H), setTyp=np.complex64))
        xs.append(torch.from_numpy(inputs.complex_gaussian(cells + (3,), seed=500 + i).astype(np.complex64)
                                   .reshape(-1, order="F").copy()).cuda())
        ys.append(torch.empty_like(xs[-1]))
    # the first operator again, on another input
    mems.append(mems[0])
    xs.append(torch.flip(xs[0], [0]).contiguous())
    ys.append(torch.empty_like(xs[0]))
    print(f"case recorded batch of {len(mems)} operators, {lanes} lanes", flush=True)
    st = torch.cuda.current_stream()
    L = gila.lib()

    def eager():
        for m, x, y in zip(mems, xs, ys):
            _apply(gila, m, x, y, 1, st)
    eager()
    torch.cuda.synchronize()
    ref = [y.clone() for y in ys]
    n = len(mems)
    ops = (ctypes.c_void_p * n)(*[m._h for m in mems])
    xp = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
    yp = (ctypes.c_void_p * n)(*[y.data_ptr() for y in ys])
    b = ctypes.c_void_p()
    _capi.check(L.gla_batch_create(ops, xp, yp, None, n, lanes, ctypes.byref(b)))
    try:
        for rep in range(3):
            for y in ys:
                y.zero_()
            eager()
            torch.cuda.synchronize()
            for y in ys:
                y.zero_()
            torch.cuda.synchronize()
            _capi.check(L.gla_batch_launch(b, ctypes.c_void_p(st.cuda_stream)))
            torch.cuda.synchronize()
            for k, (y, r) in enumerate(zip(ys, ref)):
                assert torch.equal(y, r), (rep, k)
        # the pinned work buffers cannot grow while the batch exists
        m = mems[1]
        big = torch.zeros(2 * xs[1].numel(), dtype=torch.complex64, device="cuda")
        out = torch.empty_like(big)
        st_ = L.gla_opr_apply(m._h, ctypes.c_void_p(big.data_ptr()), big.numel(), ctypes.c_void_p(out.data_ptr()),
                              out.numel(), 2, ctypes.c_void_p(st.cuda_stream))
        assert st_ == _capi.GLA_ERR_ARG and "batch" in L.gla_last_error().decode()
    finally:
        _capi.check(L.gla_batch_destroy(b))
    _apply(gila, mems[1], big, out, 2, st)
    torch.cuda.synchronize()
