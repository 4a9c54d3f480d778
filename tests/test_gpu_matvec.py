"""GPU parity of the HIP matvec against the CPU oracle, through the C ABI.

Tolerances (relative L2 against the FP64 oracle fed the same Fourier Green
data): ComplexF64 <= 1e-12, ComplexF32 <= 1e-5 (x and G^ rounded to FP32)."""
import ctypes

import numpy as np
import pytest

H = 1.0 / 32
TOL = {np.complex128: 1e-12, np.complex64: 1e-5}


def _dev_apply(gila, mem, x, nrhs=1):
    import torch
    from gila_amd import _capi
    typ = mem.setTyp
    xd = torch.from_numpy(np.asfortranarray(x.astype(typ)).ravel(order="F").copy()).cuda()
    T = tuple(mem._dims.trg_cel)
    ny = int(np.prod(T)) * 3 * nrhs
    yd = torch.empty(ny, dtype=xd.dtype, device=xd.device)
    _capi.check(gila.lib().gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                         ctypes.c_void_p(yd.data_ptr()), ny, nrhs, None))
    torch.cuda.synchronize()
    shape = T + (3,) + ((nrhs,) if nrhs > 1 else ())
    return yd.cpu().numpy().reshape(shape, order="F")


def _mem_self(gila, orc, N, typ, model=0):
    fur, crc = orc.self_operator_fur(N, model=model)
    U = orc.dims(N, N, (1, 1, 1))[1]
    vol = gila.GlaVol(N, ((1, 32),) * 3)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vol, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    return mem, fur


SELF_SHAPES = [(1, 1, 1), (2, 2, 2), (4, 4, 4), (3, 5, 6), (16, 16, 16), (8, 12, 16), (10, 20, 40), (16, 32, 16),
               (24, 48, 64), (80, 40, 40), (96, 8, 6), (24, 24, 24), (32, 48, 64), (64, 64, 64)]


@pytest.mark.gpu
@pytest.mark.parametrize("N", SELF_SHAPES)
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_self_matvec_parity(gila, orc, gpu, N, typ):
    print(f"case self N={N} {np.dtype(typ).name}", flush=True)
    mem, fur = _mem_self(gila, orc, N, typ)
    x = orc.complex_gaussian(N + (3,), seed=1)
    y_ref = orc.matvec(N, N, (1, 1, 1), fur, x)
    y = _dev_apply(gila, mem, x)
    assert orc.rel_l2(y, y_ref) <= TOL[typ]


# external pairs exercising zero-padding, spill-over and bottom-half crops on
# every axis (B = (S+T)/2 must be a compiled length)
EXT = [((2, 6, 4), (6, 2, 12)), ((6, 4, 12), (2, 12, 4)), ((16, 16, 16), (16, 16, 16)), ((1, 3, 5), (3, 1, 3)),
       ((24, 10, 30), (40, 30, 10))]


@pytest.mark.gpu
@pytest.mark.parametrize("S,T", EXT)
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_external_matvec_parity(gila, orc, gpu, S, T, typ):
    print(f"case ext S={S} T={T} {np.dtype(typ).name}", flush=True)
    delta = (0.25, -0.5, 0.75)
    crc = orc.crc_ext(S, T, (H, H, H), delta)
    sym = (0, 0, 0)
    fur = orc.fur_from_crc(S, T, sym, crc)
    B, U = orc.dims(S, T, sym)
    src = gila.GlaVol(S, ((1, 32),) * 3, (0, 0, 0))
    # a target origin that makes no axis symmetric
    trg = gila.GlaVol(T, ((1, 32),) * 3, ((1, 4), (1, 2), (3, 4)))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), trg, src, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    assert tuple(mem._dims.sym) == (0, 0, 0)
    x = orc.complex_gaussian(tuple(S) + (3,), seed=2)
    y_ref = orc.matvec(S, T, sym, fur, x)
    y = _dev_apply(gila, mem, x)
    assert orc.rel_l2(y, y_ref) <= TOL[typ]


# branch lengths with a factor beyond 5: 29 (prime, one direct radix-29
# stage), 81 = 3^4 and 145 = 5 x 29, on every axis position (pass 1/5 rows,
# pass 2/4 bundles, pass 3 columns)
ODD_SELF = [(29, 81, 4), (145, 4, 29), (4, 145, 81), (6, 29, 145)]


@pytest.mark.gpu
@pytest.mark.parametrize("N", ODD_SELF)
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_self_matvec_odd_lengths(gila, orc, gpu, N, typ):
    print(f"case self odd lengths N={N} {np.dtype(typ).name}", flush=True)
    mem, fur = _mem_self(gila, orc, N, typ)
    assert tuple(mem._dims.dev_brn) == N
    x = orc.complex_gaussian(N + (3,), seed=21)
    y_ref = orc.matvec(N, N, (1, 1, 1), fur, x)
    assert orc.rel_l2(_dev_apply(gila, mem, x), y_ref) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_external_fur_with_branch_lengths_29_and_81(gila, orc, gpu, typ):
    # Fourier data in the reference's layout for B = (4, 29, 81): accepted as
    # it is (no re-embedding) and applied to the oracle's bound
    S, T = (4, 30, 2), (4, 28, 160)
    print(f"case ext fur B=(4,29,81) {np.dtype(typ).name}", flush=True)
    crc = orc.crc_ext(S, T, (H, H, H), (0.25, 0.5, -0.75))
    sym = (0, 0, 0)
    fur = orc.fur_from_crc(S, T, sym, crc)
    B, U = orc.dims(S, T, sym)
    assert B == (4, 29, 81)
    src = gila.GlaVol(S, ((1, 32),) * 3, (0, 0, 0))
    trg = gila.GlaVol(T, ((1, 32),) * 3, ((1, 4), (1, 2), (3, 4)))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), trg, src, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    assert tuple(mem._dims.dev_brn) == B
    x = orc.complex_gaussian(S + (3,), seed=22)
    assert orc.rel_l2(_dev_apply(gila, mem, x), orc.matvec(S, T, sym, fur, x)) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_external_touching_symmetric_axes(gila, orc, gpu, typ):
    # volB <- volA of test/runtests.jl (touching along y; x and z symmetric)
    A = (16, 16, 16)
    crc = orc.crc_ext(A, A, (H, H, H), (0.0, 16 * H, 0.0))
    sym = (1, 0, 1)
    fur = orc.fur_from_crc(A, A, sym, crc)
    U = orc.dims(A, A, sym)[1]
    va = gila.GlaVol(A, ((1, 32),) * 3, (0, 0, 0))
    vb = gila.GlaVol(A, ((1, 32),) * 3, (0, (16, 32), 0))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vb, va, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    assert tuple(mem._dims.sym) == sym
    x = orc.complex_gaussian(A + (3,), seed=3)
    assert orc.rel_l2(_dev_apply(gila, mem, x), orc.matvec(A, A, sym, fur, x)) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_batched_rhs(gila, orc, gpu, typ):
    N = (12, 8, 16)
    mem, fur = _mem_self(gila, orc, N, typ)
    nrhs = 5
    x = orc.complex_gaussian(N + (3, nrhs), seed=4)
    y = _dev_apply(gila, mem, x, nrhs=nrhs)
    for r in range(nrhs):
        y_ref = orc.matvec(N, N, (1, 1, 1), fur, x[..., r])
        assert orc.rel_l2(y[..., r], y_ref) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("S,T", [((6, 4, 12), (2, 12, 4)), ((24, 10, 30), (40, 30, 10))])
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_batched_rhs_external(gila, orc, gpu, S, T, typ):
    # external pairs (source and target strides differ; pass 3 adds spill-over
    # rows or writes rows beyond B on some axis) on four right-hand sides in
    # one launch
    print(f"case ext batched S={S} T={T} {np.dtype(typ).name}", flush=True)
    crc = orc.crc_ext(S, T, (H, H, H), (0.25, -0.5, 0.75))
    sym = (0, 0, 0)
    fur = orc.fur_from_crc(S, T, sym, crc)
    B, U = orc.dims(S, T, sym)
    src = gila.GlaVol(S, ((1, 32),) * 3, (0, 0, 0))
    trg = gila.GlaVol(T, ((1, 32),) * 3, ((1, 4), (1, 2), (3, 4)))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), trg, src, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    nrhs = 4
    x = orc.complex_gaussian(tuple(S) + (3, nrhs), seed=44)
    y = _dev_apply(gila, mem, x, nrhs=nrhs)
    for r in range(nrhs):
        assert orc.rel_l2(y[..., r], orc.matvec(S, T, sym, fur, x[..., r])) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_full_size_128_cubed(gila, orc, gpu, typ):
    # the 128^3 self operator at full size against the FP64 oracle
    N = (128, 128, 128)
    print(f"case self 128^3 {np.dtype(typ).name}", flush=True)
    mem, fur = _mem_self(gila, orc, N, typ)
    x = orc.complex_gaussian(N + (3,), seed=2)
    y_ref = orc.matvec(N, N, (1, 1, 1), fur, x)
    y = _dev_apply(gila, mem, x)
    assert orc.rel_l2(y, y_ref) <= TOL[typ]


@pytest.mark.gpu
def test_operator_surface(gila, orc, gpu):
    """test/oprTest.jl generic_tests on a device operator."""
    import torch
    N = (16, 16, 16)
    typ = np.complex64
    mem, fur = _mem_self(gila, orc, N, typ)
    opr = gila.GlaOpr(mem)
    assert gila.size(opr) == (16 * 16 * 16 * 3,) * 2
    assert gila.glaSze(opr) == ((16, 16, 16, 3), (16, 16, 16, 3))
    assert gila.isselfoperator(opr) and not gila.isexternaloperator(opr)
    x = torch.ones(gila.size(opr, 2), dtype=torch.complex64, device="cuda")
    xg = x.reshape(16, 16, 16, 3)     # all-ones: layout is immaterial
    ref = gila.egoOpr_(mem, xg.clone())
    acted = opr * xg
    assert torch.allclose(acted, ref)
    assert torch.equal(x, torch.ones_like(x))          # input not mutated
    acted_flat = opr * x
    assert acted_flat.shape == x.shape
    with pytest.raises(AssertionError):
        opr * torch.ones(gila.size(opr, 2), dtype=torch.complex128, device="cuda")
    with pytest.raises(gila.DimensionMismatch):
        opr * torch.ones(gila.size(opr, 2) + 1, dtype=torch.complex64, device="cuda")
    adj = gila.adjoint(opr)
    adj2 = gila.adjoint(adj)
    assert gila.isadjoint(adj) and not gila.isadjoint(opr) and not gila.isadjoint(adj2)
    assert all(np.array_equal(a, b) for a, b in zip(adj2.mem.egoFur, opr.mem.egoFur))
    assert torch.allclose(adj2 * x, opr * x)
    assert not gila.issymmetric(opr) and not gila.isposdef(opr) and not gila.ishermitian(opr)
    assert not gila.isdiag(opr)
    assert "Self GlaOpr" in repr(opr)


@pytest.mark.gpu
def test_mul_axpby_fused(gila, orc, gpu):
    import torch
    N = (8, 8, 8)
    typ = np.complex128
    mem, fur = _mem_self(gila, orc, N, typ)
    opr = gila.GlaOpr(mem)
    x = orc.complex_gaussian(N + (3,), seed=5)
    y0 = orc.complex_gaussian(N + (3,), seed=6)
    a, b = 0.5 - 2j, 1.5 + 0.25j
    ref = a * orc.matvec(N, N, (1, 1, 1), fur, x) + b * y0
    xd = torch.from_numpy(x.ravel(order="F").copy()).cuda()
    yd = torch.from_numpy(y0.ravel(order="F").copy()).cuda()
    gila.mul_(yd, opr, xd, a, b)
    torch.cuda.synchronize()
    assert orc.rel_l2(yd.cpu().numpy().reshape(N + (3,), order="F"), ref) <= 1e-12


@pytest.mark.gpu
def test_host_arrays_and_errors(gila, orc, gpu):
    N = (6, 4, 8)
    typ = np.complex128
    mem, fur = _mem_self(gila, orc, N, typ)
    x = orc.complex_gaussian(N + (3,), seed=7)
    y = gila.egoOpr_(mem, x)          # numpy in, numpy out
    assert orc.rel_l2(y, orc.matvec(N, N, (1, 1, 1), fur, x)) <= 1e-12
    with pytest.raises(gila.GlaError) as ei:
        gila.egoOpr_(mem, np.zeros((6, 4, 7, 3), typ))
    assert ei.value.status == 1
    bad = [f.copy() for f in mem.egoFur]
    bad[3][0, 0, 0, 0] = np.nan
    with pytest.raises(gila.GlaError) as ei:
        gila.GlaOprMem(gila.GlaKerOpt(True), mem.trgVol, egoFur=bad, setTyp=typ)
    assert ei.value.status == 2


@pytest.mark.gpu
def test_profile_entry_leaves_a_valid_matvec(gila, orc, gpu):
    import torch
    from gila_amd import _capi
    N = (16, 16, 16)
    mem, fur = _mem_self(gila, orc, N, np.complex64)
    x = orc.complex_gaussian(N + (3,), seed=10)
    xd = torch.from_numpy(x.astype(np.complex64).ravel(order="F").copy()).cuda()
    yd = torch.empty_like(xd)
    ms = (ctypes.c_double * 5)()
    by = (ctypes.c_int64 * 5)()
    st = torch.cuda.current_stream().cuda_stream
    _capi.check(gila.lib().gla_opr_profile(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                           ctypes.c_void_p(yd.data_ptr()), yd.numel(), 1, ctypes.c_void_p(st), 3,
                                           ms, by))
    torch.cuda.synchronize()
    assert all(ms[i] > 0 for i in range(5)) and all(by[i] > 0 for i in range(5))
    y = yd.cpu().numpy().reshape(N + (3,), order="F")
    assert orc.rel_l2(y, orc.matvec(N, N, (1, 1, 1), fur, x)) <= 1e-5
    # the in-sequence form: the same byte figures, positive pass durations
    # whose sum fits in the whole-matvec time, and y bitwise the plain apply
    ms2 = (ctypes.c_double * 5)()
    by2 = (ctypes.c_int64 * 5)()
    tot = ctypes.c_double()
    ys = torch.empty_like(xd)
    _capi.check(gila.lib().gla_opr_profile_seq(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                               ctypes.c_void_p(ys.data_ptr()), ys.numel(), 1, ctypes.c_void_p(st), 4,
                                               ms2, by2, ctypes.byref(tot)))
    torch.cuda.synchronize()
    assert [by2[i] for i in range(5)] == [by[i] for i in range(5)]
    assert all(ms2[i] > 0 for i in range(5)) and sum(ms2[i] for i in range(5)) <= tot.value * 1.001
    assert torch.equal(ys.cpu(), yd.cpu())
    with pytest.raises(gila.GlaError):
        _capi.check(gila.lib().gla_opr_profile_seq(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                                   ctypes.c_void_p(ys.data_ptr()), ys.numel(), 1,
                                                   ctypes.c_void_p(st), 0, ms2, by2, ctypes.byref(tot)))


@pytest.mark.gpu
def test_batched_96_cubed_plane_waves_and_dipoles(gila, orc, gpu):
    # the 96^3 batched workload's right-hand sides (plane waves with random
    # directions and polarisations, point dipoles), 4 of them in one launch,
    # each column against the FP64 oracle
    from fractions import Fraction
    from gila_amd import inputs
    N = (96, 96, 96)
    print("case batched 96^3 plane waves + dipoles", flush=True)
    mem, fur = _mem_self(gila, orc, N, np.complex64)
    cols = []
    for i in range(2):
        khat = inputs.random_unit(7000 + i)
        p = inputs.random_unit(8000 + i)
        p = p - khat * (p @ khat)
        cols.append(inputs.plane_wave(N, (Fraction(1, 32),) * 3, khat, p / np.linalg.norm(p)))
        cols.append(inputs.point_dipole(N, [11 + 37 * i, 50, 95 - 40 * i], inputs.random_unit(9100 + i)))
    x = np.stack(cols, axis=4)
    y = _dev_apply(gila, mem, x, nrhs=4)
    for r in range(4):
        assert orc.rel_l2(y[..., r], orc.matvec(N, N, (1, 1, 1), fur, x[..., r])) <= 1e-5


Z_LENGTHS = [16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512]


@pytest.mark.gpu
@pytest.mark.parametrize("n", Z_LENGTHS)
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_every_tuned_pass3_length(gila, orc, gpu, n, typ):
    # pass 3 runs a length-n transform with the shape the tuning table picks
    # for (element type, n); every compiled length >= 16 against the oracle
    N = (4, 6, n)
    print(f"case pass-3 length {n} {np.dtype(typ).name}", flush=True)
    mem, fur = _mem_self(gila, orc, N, typ)
    x = orc.complex_gaussian(N + (3,), seed=n)
    assert orc.rel_l2(_dev_apply(gila, mem, x), orc.matvec(N, N, (1, 1, 1), fur, x)) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("n", Z_LENGTHS)
def test_every_tuned_pass3_length_large_grid(gila, gpu, n):
    # 128 x 64 x n cells: 256 x 128 columns, the large-grid class of pass 3
    # (the test above runs the small class).  Too large for the oracle, so the
    # two precisions, built from geometry and run on their own shapes, check
    # each other: a wrong shape in either shows as a difference far above the
    # ComplexF32 rounding.
    import torch
    from gila_amd import _capi
    N = (128, 64, n)
    print(f"case pass-3 length {n} large grid", flush=True)
    vol = gila.GlaVol(N, ((1, 32),) * 3)
    x = np.random.default_rng(n).standard_normal(N + (3, 2)).view(np.complex128)[..., 0]
    out = {}
    for typ in (np.complex128, np.complex64):
        mem = gila.GlaOprMem(gila.GlaKerOpt(True), vol, setTyp=typ)
        assert tuple(mem._dims.dev_brn) == N
        xd = torch.from_numpy(np.asfortranarray(x.astype(typ)).ravel(order="F").copy()).cuda()
        yd = torch.empty_like(xd)
        _capi.check(gila.lib().gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                             ctypes.c_void_p(yd.data_ptr()), yd.numel(), 1, None))
        torch.cuda.synchronize()
        out[typ] = yd.cpu().numpy().astype(np.complex128)
        del mem, xd, yd
    d = np.linalg.norm(out[np.complex64] - out[np.complex128]) / np.linalg.norm(out[np.complex128])
    assert d <= 2e-5, d


ROW_LENGTHS = [16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512]


@pytest.mark.gpu
@pytest.mark.parametrize("n", ROW_LENGTHS)
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_every_tuned_row_length(gila, orc, gpu, n, typ):
    # passes 1/5 run a length-n row transform with the shape and parity
    # batching the tables pick for (element type, n), passes 2/4 a length-n
    # column transform: every compiled length >= 16 on axis 1 and on axis 2
    for N in ((n, 6, 4), (4, n, 6)):
        print(f"case row length {n} N={N} {np.dtype(typ).name}", flush=True)
        mem, fur = _mem_self(gila, orc, N, typ)
        x = orc.complex_gaussian(N + (3,), seed=n + 1)
        assert orc.rel_l2(_dev_apply(gila, mem, x), orc.matvec(N, N, (1, 1, 1), fur, x)) <= TOL[typ]


SQUARE_SELF = [(32, 32, 8), (64, 64, 6), (96, 96, 4), (128, 128, 2)]
SQUARE_EXT = [((32, 40, 4), (32, 24, 4)), ((32, 24, 4), (32, 40, 4)), ((64, 70, 3), (64, 58, 5))]


@pytest.mark.gpu
@pytest.mark.parametrize("N", SQUARE_SELF)
def test_square_xy_planes_self(gila, orc, gpu, N):
    # ComplexF32 with B1 = B2 in {32, 64, 96, 128}, thin along z, two
    # right-hand sides
    print(f"case square xy self N={N}", flush=True)
    mem, fur = _mem_self(gila, orc, N, np.complex64)
    x = orc.complex_gaussian(N + (3, 2), seed=70)
    y = _dev_apply(gila, mem, x, nrhs=2)
    for r in range(2):
        assert orc.rel_l2(y[..., r], orc.matvec(N, N, (1, 1, 1), fur, x[..., r])) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("S,T", SQUARE_EXT)
def test_square_xy_external_spill_and_zero_rows(gila, orc, gpu, S, T):
    # S2 > B2 (spill-over rows added) and S2 < B2 (zero rows) with B1 = B2
    print(f"case square xy ext S={S} T={T}", flush=True)
    delta = (0.25, -0.5, 0.75)
    crc = orc.crc_ext(S, T, (H, H, H), delta)
    sym = (0, 0, 0)
    fur = orc.fur_from_crc(S, T, sym, crc)
    B, U = orc.dims(S, T, sym)
    assert B[0] == B[1]
    src = gila.GlaVol(S, ((1, 32),) * 3, (0, 0, 0))
    trg = gila.GlaVol(T, ((1, 32),) * 3, ((1, 4), (1, 2), (3, 4)))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), trg, src, egoFur=orc.fur_branches(fur.astype(np.complex64), U),
                         setTyp=np.complex64)
    x = orc.complex_gaussian(tuple(S) + (3,), seed=71)
    assert orc.rel_l2(_dev_apply(gila, mem, x), orc.matvec(S, T, sym, fur, x)) <= 1e-5
