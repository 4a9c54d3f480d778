"""Fourier data in the reference's layout at branch lengths without a compiled
transform (the reference plans FFTs of any length, src/glaMemSup.jl:213-228):
gla_opr_create_from_fur re-expresses it at the next fast length (expand the
stored branches, inverse-transform to the real-space circulant in FP64,
re-embed, transform forward) and the matvec must then equal the oracle's
matvec on the original data -- 1e-12 in ComplexF64 (the round trip adds FP64
rounding only), the ComplexF32 bound 1e-5 otherwise.  Cases: a self operator
with B = 7, 9, 11 (symmetric storage on every axis), an external pair with
B = 13, 7, 6 and a partitioned pair with B = 7 on its divided axis."""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

H = 1.0 / 32
TOL = {np.complex128: 1e-12, np.complex64: 1e-5}


def _apply(gila, mem, x, T):
    import torch
    from gila_amd import _capi
    typ = mem.setTyp
    xd = torch.from_numpy(np.asfortranarray(x.astype(typ)).ravel(order="F").copy()).cuda()
    ny = int(np.prod(T)) * 3
    yd = torch.empty(ny, dtype=xd.dtype, device=xd.device)
    _capi.check(gila.lib().gla_opr_apply(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(),
                                         ctypes.c_void_p(yd.data_ptr()), ny, 1, None))
    torch.cuda.synchronize()
    return yd.cpu().numpy().reshape(tuple(T) + (3,), order="F")


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_self_fur_at_uncompiled_lengths(gila, orc, gpu, typ):
    N = (7, 9, 11)
    print(f"case self fur B={N} {np.dtype(typ).name}", flush=True)
    fur, _ = orc.self_operator_fur(N)
    B, U = orc.dims(N, N, (1, 1, 1))
    vol = gila.GlaVol(N, ((1, 32),) * 3)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vol, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    assert tuple(mem._dims.brn) == B and tuple(mem._dims.dev_brn) == (8, 10, 12)
    x = orc.complex_gaussian(N + (3,), seed=31)
    ref = orc.matvec(N, N, (1, 1, 1), fur.astype(typ).astype(np.complex128), x)
    assert orc.rel_l2(_apply(gila, mem, x, N), ref) <= TOL[typ]


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_external_fur_at_uncompiled_lengths(gila, orc, gpu, typ):
    S, T = (10, 6, 4), (16, 8, 8)
    print(f"case external fur B=(13,7,6) {np.dtype(typ).name}", flush=True)
    crc = orc.crc_ext(S, T, (H, H, H), (0.5, -0.25, 0.75))
    sym = (0, 0, 0)
    fur = orc.fur_from_crc(S, T, sym, crc)
    B, U = orc.dims(S, T, sym)
    assert B == (13, 7, 6)
    src = gila.GlaVol(S, ((1, 32),) * 3, (0, 0, 0))
    trg = gila.GlaVol(T, ((1, 32),) * 3, ((1, 2), (-1, 4), (3, 4)))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), trg, src, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    assert tuple(mem._dims.dev_brn) == (16, 8, 6)
    x = orc.complex_gaussian(S + (3,), seed=32)
    ref = orc.matvec(S, T, sym, fur.astype(typ).astype(np.complex128), x)
    assert orc.rel_l2(_apply(gila, mem, x, T), ref) <= TOL[typ]


@pytest.mark.gpu
def test_partitioned_fur_at_uncompiled_length(gila, orc, gpu):
    # source divided on x: partition cells 8 + target 6 -> B = 7 on x
    h = Fraction(1, 32)
    Tf, sT, Sf, sS, orgS = (6, 4, 4), (2 * h, h, h), (16, 4, 4), (h, h, h), (20 * h, 0, 0)
    print("case partitioned fur B=(7,4,4)", flush=True)
    fur, S, T, sdiv, tdiv = orc.partition_fur(Tf, Sf, [float(v) for v in sT], [float(v) for v in sS],
                                              (0.0, 0.0, 0.0), [float(v) for v in orgS])
    B, U = orc.dims(S, T, (0, 0, 0))
    assert B == (7, 4, 4) and sdiv == (2, 1, 1)
    P = int(np.prod(sdiv)) * int(np.prod(tdiv))
    per = U[0] * U[1] * U[2] * 6 * P
    branches = [fur[b * per:(b + 1) * per].reshape(tuple(U) + (6, int(np.prod(sdiv)), int(np.prod(tdiv))), order="F")
                for b in range(8)]
    vt, vs = gila.GlaVol(Tf, sT, (0, 0, 0)), gila.GlaVol(Sf, sS, orgS)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vt, vs, egoFur=branches, setTyp=np.complex128)
    assert tuple(mem._dims.dev_brn) == (8, 4, 4)
    x = orc.complex_gaussian(Sf + (3,), seed=33)
    ref = orc.partition_matvec(S, T, sdiv, tdiv, fur, x)
    assert orc.rel_l2(_apply(gila, mem, x, Tf), ref) <= 1e-12
