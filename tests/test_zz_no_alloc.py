"""The allocation-free matvec entry (gla_opr_apply_noalloc) against
gla_opr_apply, through the C ABI: the same kernels on the same inputs, so the
outputs must agree bit for bit, on two calls in a row where the second is
smaller (the shared scratch still holds the first call's data), and both must
match the FP64 oracle to the usual tolerances."""
import ctypes

import numpy as np
import pytest

TOL = {np.complex128: 1e-12, np.complex64: 1e-5}


def _apply(gila, mem, x, nrhs, entry):
    import torch
    from gila_amd import _capi
    typ = mem.setTyp
    xd = torch.from_numpy(np.asfortranarray(x.astype(typ)).ravel(order="F").copy()).cuda()
    T = tuple(mem._dims.trg_cel)
    ny = int(np.prod(T)) * 3 * nrhs
    yd = torch.full((ny,), float("nan"), dtype=xd.dtype, device=xd.device)
    fn = getattr(gila.lib(), entry)
    _capi.check(fn(mem._h, ctypes.c_void_p(xd.data_ptr()), xd.numel(), ctypes.c_void_p(yd.data_ptr()), ny, nrhs,
                   None))
    torch.cuda.synchronize()
    return yd.cpu().numpy().reshape(T + (3, nrhs), order="F")


def _self_mem(gila, orc, N, typ):
    fur, _ = orc.self_operator_fur(N)
    U = orc.dims(N, N, (1, 1, 1))[1]
    vol = gila.GlaVol(N, ((1, 32),) * 3)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vol, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    return mem, fur, N, N, (1, 1, 1)


def _ext_mem(gila, orc, S, T, typ):
    crc = orc.crc_ext(S, T, (1 / 32,) * 3, (0.25, -0.5, 0.75))
    sym = (0, 0, 0)
    fur = orc.fur_from_crc(S, T, sym, crc)
    U = orc.dims(S, T, sym)[1]
    src = gila.GlaVol(S, ((1, 32),) * 3, (0, 0, 0))
    trg = gila.GlaVol(T, ((1, 32),) * 3, ((1, 4), (1, 2), (3, 4)))
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), trg, src, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    return mem, fur, S, T, sym


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex64, np.complex128])
def test_noalloc_matches_apply_on_two_calls(gila, orc, gpu, typ):
    # a larger self operator with two right-hand sides, then a smaller
    # external one, then a smaller self one: each call must not see the
    # previous call's leftovers in the shared scratch
    cases = [("self 32x24x16 nrhs=2", _self_mem(gila, orc, (32, 24, 16), typ), 2),
             ("ext (6,4,12)->(2,12,4) nrhs=1", _ext_mem(gila, orc, (6, 4, 12), (2, 12, 4), typ), 1),
             ("self 8x12x6 nrhs=1", _self_mem(gila, orc, (8, 12, 6), typ), 1)]
    for k, (name, (mem, fur, S, T, sym), nrhs) in enumerate(cases):
        print(f"case noalloc {name} {np.dtype(typ).name}", flush=True)
        x = orc.complex_gaussian(tuple(S) + (3, nrhs), seed=10 + k)
        y0 = _apply(gila, mem, x, nrhs, "gla_opr_apply_noalloc")
        y1 = _apply(gila, mem, x, nrhs, "gla_opr_apply")
        assert np.array_equal(y0, y1), name
        for r in range(nrhs):
            y_ref = orc.matvec(S, T, sym, fur, x[..., r])
            assert orc.rel_l2(y0[..., r], y_ref) <= TOL[typ], name


@pytest.mark.gpu
def test_noalloc_rejects_what_does_not_fit(gila, orc, gpu):
    import torch
    from gila_amd import _capi
    mem, _, S, T, _ = _self_mem(gila, orc, (32, 24, 16), np.complex64)
    L = gila.lib()
    per_rhs = L.gla_opr_workspace_bytes(mem._h, 1)
    nrhs = _capi.GLA_NOALLOC_SCRATCH_BYTES // per_rhs + 1
    print(f"case noalloc too large: nrhs={nrhs} x {per_rhs} bytes", flush=True)
    n_x = int(np.prod(S)) * 3 * nrhs
    n_y = int(np.prod(T)) * 3 * nrhs
    xd = torch.zeros(n_x, dtype=torch.complex64, device="cuda")
    yd = torch.zeros(n_y, dtype=torch.complex64, device="cuda")
    st = L.gla_opr_apply_noalloc(mem._h, ctypes.c_void_p(xd.data_ptr()), n_x, ctypes.c_void_p(yd.data_ptr()), n_y,
                                 nrhs, None)
    assert st == _capi.GLA_ERR_SIZE
    assert "scratch" in L.gla_last_error().decode()
    # nothing was launched: y is untouched
    torch.cuda.synchronize()
    assert bool((yd == 0).all())
