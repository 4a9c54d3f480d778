"""Partitioned (multi-resolution) external operators on the GPU: cell scales
that differ by an integer factor per axis (GlaExtInf, src/glaMem.jl:190-216).

The setup's Fourier data (one block per (source, target) partition pair, the
reference layout (U1,U2,U3,6,Ps,Pt) per branch) must match the oracle's
restatement (genEgoExt! on the partition volumes, src/glaMemSup.jl:64-89),
whose mixed-size tensor and partition matvec are pinned on the CPU
(tests/test_oracle_partition.py).  The matvec -- pass 1 gathering a source
partition, pass 5 writing or accumulating a target partition -- must match
the oracle's partition matvec fed the same Fourier data: 1e-12 in FP64, the
ComplexF32 bound 1e-5 otherwise."""
from fractions import Fraction

import ctypes
import numpy as np
import pytest

H = Fraction(1, 32)
SETUP_TOL = 1e-10

# (target cells, target scale, source cells, source scale, source origin)
CASES = {
    "src_div_x": ((4, 6, 4), (2 * H, H, H), (12, 6, 4), (H, H, H), (14 * H, 0, 0)),
    "trg_div_yz": ((6, 8, 4), (H, H, H), (6, 4, 2), (H, 2 * H, 2 * H), (-2 * H, 0, 10 * H)),
    "both": ((4, 8, 4), (2 * H, H, H), (8, 4, 4), (H, 2 * H, H), (0, 11 * H, -H / 2)),
}


def _vols(gila, case):
    Tf, sT, Sf, sS, orgS = CASES[case]
    return gila.GlaVol(Tf, sT, (0, 0, 0)), gila.GlaVol(Sf, sS, orgS)


def _oracle_fur(orc, case):
    Tf, sT, Sf, sS, orgS = CASES[case]
    f = [float(v) for v in sT], [float(v) for v in sS], [float(v) for v in orgS]
    return orc.partition_fur(Tf, Sf, f[0], f[1], (0.0, 0.0, 0.0), f[2])


def _fur_flat(mem):
    return np.concatenate([f.reshape(-1, order="F") for f in mem.egoFur])


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_partitioned_setup_matches_oracle(gila, orc, gpu, case):
    print(f"case partition setup {case}", flush=True)
    vt, vs = _vols(gila, case)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, setTyp=np.complex128)
    ref, S, T, sdiv, tdiv = _oracle_fur(orc, case)
    d = mem._dims
    assert tuple(d.src_cel) == S and tuple(d.trg_cel) == T
    assert tuple(d.src_div) == sdiv and tuple(d.trg_div) == tdiv
    fur = _fur_flat(mem)
    assert fur.shape == ref.shape
    assert np.abs(fur - ref).max() / np.abs(ref).max() <= SETUP_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_partitioned_matvec_matches_oracle(gila, orc, gpu, case, dtype, tol):
    import torch
    print(f"case partition matvec {case} {np.dtype(dtype).name}", flush=True)
    vt, vs = _vols(gila, case)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, setTyp=dtype)
    d = mem._dims
    S, T = tuple(d.src_cel), tuple(d.trg_cel)
    sdiv, tdiv = tuple(d.src_div), tuple(d.trg_div)
    x = orc.complex_gaussian(tuple(vs.cel) + (3,), seed=123).astype(dtype)
    xd = torch.from_numpy(x.reshape(-1, order="F").copy()).cuda()
    y = gila.egoOpr_(mem, xd.reshape(tuple(reversed(vs.cel + (3,)))).permute(3, 2, 1, 0))
    yh = y.permute(3, 2, 1, 0).contiguous().cpu().numpy().reshape(-1)
    yh = yh.reshape(tuple(vt.cel) + (3,), order="F")
    fur = _fur_flat(mem).astype(np.complex128)
    ref = orc.partition_matvec(S, T, sdiv, tdiv, fur, x.astype(np.complex128))
    assert orc.rel_l2(yh, ref) <= tol


@pytest.mark.gpu
def test_partitioned_from_egoFur_and_entry_points(gila, orc, gpu, tmp_path):
    """egoFur re-entry with the oracle's data; alpha / beta accumulation, the
    allocation-free entry and a Green-data file round trip agree with the
    plain matvec."""
    import torch
    from gila_amd import _capi
    case = "both"
    print(f"case partition egoFur {case}", flush=True)
    vt, vs = _vols(gila, case)
    ref_fur, S, T, sdiv, tdiv = _oracle_fur(orc, case)
    g0 = gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, setTyp=np.complex128)
    U = tuple(g0._dims.fur_ext)
    Ps, Pt = int(np.prod(sdiv)), int(np.prod(tdiv))
    per = U[0] * U[1] * U[2] * 6 * Ps * Pt
    branches = [ref_fur[b * per:(b + 1) * per].reshape(U + (6, Ps, Pt), order="F") for b in range(8)]
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, egoFur=branches, setTyp=np.complex128)
    assert np.array_equal(_fur_flat(mem), ref_fur)
    x = orc.complex_gaussian(tuple(vs.cel) + (3,), seed=321)
    ref = orc.partition_matvec(S, T, sdiv, tdiv, ref_fur, x)
    L = gila.lib()
    xd = torch.from_numpy(x.reshape(-1, order="F").copy()).cuda()
    ny = 3 * int(np.prod(vt.cel))
    yd = torch.empty(ny, dtype=torch.complex128, device="cuda")
    st = torch.cuda.current_stream()

    def apply(m, fn="gla_opr_apply", out=yd):
        _capi.check(getattr(L, fn)(m._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(), ctypes.c_void_p(out.data_ptr()),
                                   out.numel(), 1, ctypes.c_void_p(st.cuda_stream)))
    apply(mem)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().reshape(tuple(vt.cel) + (3,), order="F")
    assert orc.rel_l2(y, ref) <= 1e-12
    # y = alpha G x + beta y0
    y0 = torch.from_numpy(orc.complex_gaussian((ny,), seed=9)).cuda()
    ya = y0.clone()
    al, be = 0.5 - 1.5j, -2.0 + 0.25j
    _capi.check(L.gla_opr_apply_axpby(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                      ctypes.c_void_p(ya.data_ptr()), ya.numel(), 1, al.real, al.imag, be.real,
                                      be.imag, ctypes.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    want = al * yd + be * y0
    assert (torch.linalg.vector_norm(ya - want) / torch.linalg.vector_norm(want)).item() <= 1e-13
    # the allocation-free entry point runs the same launches
    yn = torch.empty_like(yd)
    apply(mem, "gla_opr_apply_noalloc", yn)
    torch.cuda.synchronize()
    assert torch.equal(yn, yd)
    # Green-data file round trip
    path = tmp_path / "prt.gfur"
    mem.save(path)
    back = gila.GlaOprMem.load(path)
    assert tuple(back._dims.src_div) == sdiv and tuple(back._dims.trg_div) == tdiv
    yb = torch.empty_like(yd)
    apply(back, out=yb)
    torch.cuda.synchronize()
    assert torch.equal(yb, yd)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_partitioned_batched_rhs(gila, orc, gpu, case):
    """Three right-hand sides in one launch: every column equals its own
    single-right-hand-side launch bitwise and the oracle's partition matvec to
    1e-12 (FP64)."""
    import torch
    from gila_amd import _capi
    print(f"case partition batched rhs {case}", flush=True)
    vt, vs = _vols(gila, case)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, setTyp=np.complex128)
    d = mem._dims
    S, T = tuple(d.src_cel), tuple(d.trg_cel)
    sdiv, tdiv = tuple(d.src_div), tuple(d.trg_div)
    fur = _fur_flat(mem)
    nx, ny, nr = 3 * int(np.prod(vs.cel)), 3 * int(np.prod(vt.cel)), 3
    x = np.stack([orc.complex_gaussian(tuple(vs.cel) + (3,), seed=400 + r) for r in range(nr)], axis=4)
    xd = torch.from_numpy(x.reshape(-1, order="F").copy()).cuda()
    yd = torch.empty(ny * nr, dtype=torch.complex128, device="cuda")
    L = gila.lib()
    st = torch.cuda.current_stream()
    _capi.check(L.gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(), ctypes.c_void_p(yd.data_ptr()),
                                yd.numel(), nr, ctypes.c_void_p(st.cuda_stream)))
    for r in range(nr):
        x1 = xd[r * nx:(r + 1) * nx].clone()
        y1 = torch.empty(ny, dtype=torch.complex128, device="cuda")
        _capi.check(L.gla_opr_apply(mem._h, ctypes.c_void_p(x1.data_ptr()), nx, ctypes.c_void_p(y1.data_ptr()),
                                    ny, 1, ctypes.c_void_p(st.cuda_stream)))
        torch.cuda.synchronize()
        assert torch.equal(y1, yd[r * ny:(r + 1) * ny])
        yh = y1.cpu().numpy().reshape(tuple(vt.cel) + (3,), order="F")
        ref = orc.partition_matvec(S, T, sdiv, tdiv, fur, x[..., r])
        assert orc.rel_l2(yh, ref) <= 1e-12


@pytest.mark.gpu
def test_partitioned_volumes_in_contact_are_refused(gila, gpu):
    vt = gila.GlaVol((4, 4, 4), (2 * H, H, H), (0, 0, 0))
    vs = gila.GlaVol((8, 4, 4), (H, H, H), (8 * H, 0, 0))   # touching faces at x = 4h
    with pytest.raises(gila.GlaError) as ei:
        gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, setTyp=np.complex128)
    assert ei.value.status == 8
