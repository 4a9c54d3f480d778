"""CPU checks of the C ABI: the library loads, exports every symbol that
include/gila_amd.h declares, and its host-only paths behave (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gila_amd.h")).read()
    return sorted(set(re.findall(r"\b(gla_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(gila):
    lib = gila.lib()
    names = _declared()
    assert "gla_opr_apply" in names and "gla_opr_create_from_fur" in names
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_header(gila):
    from gila_amd import _capi
    assert sorted(_capi.EXPORTED) == _declared()


def test_dims_for_self_and_external(gila):
    from gila_amd import _capi
    L = gila.lib()
    v = gila.GlaVol((16, 16, 16), ((1, 32),) * 3)._c()
    d = _capi.gla_dims()
    _capi.check(L.gla_opr_dims_for(ctypes.byref(v), None, ctypes.byref(d)))
    assert tuple(d.brn) == (16, 16, 16) and tuple(d.fur_ext) == (9, 9, 9) and tuple(d.sym) == (1, 1, 1)
    # external, as volB <- volA in test/runtests.jl:10-22 (touching in y)
    a = gila.GlaVol((16, 16, 16), ((1, 50),) * 3, (0, 0, 0))._c()
    b = gila.GlaVol((16, 16, 16), ((1, 50),) * 3, (0, (16, 50), 0))._c()
    _capi.check(L.gla_opr_dims_for(ctypes.byref(b), ctypes.byref(a), ctypes.byref(d)))
    assert tuple(d.sym) == (1, 0, 1) and tuple(d.fur_ext) == (9, 16, 9) and d.self_opr == 0


def test_dims_errors(gila):
    from gila_amd import _capi
    L = gila.lib()
    d = _capi.gla_dims()
    a = gila.GlaVol((4, 4, 4), ((1, 32),) * 3)._c()
    # an odd cell sum regenerates the odd volume (3 cells of 1/32 -> 4 of
    # 3/128, src/glaMemSup.jl:34-40), whose scale then shares no grid with
    # 1/32 (src/glaMem.jl:192-195)
    b = gila.GlaVol((3, 4, 4), ((1, 32),) * 3, (1, 0, 0))._c()
    assert L.gla_opr_dims_for(ctypes.byref(a), ctypes.byref(b), ctypes.byref(d)) == _capi.GLA_ERR_SCALE
    c = gila.GlaVol((4, 4, 4), ((1, 33),) * 3, (1, 0, 0))._c()
    assert L.gla_opr_dims_for(ctypes.byref(a), ctypes.byref(c), ctypes.byref(d)) == _capi.GLA_ERR_SCALE
    assert "common scale" in L.gla_last_error().decode()


def test_odd_sum_regenerates_volumes(gila):
    from gila_amd import _capi
    L = gila.lib()
    v = gila.GlaVol((3, 4, 5), ((1, 32),) * 3)._c()
    w = _capi.gla_vol()
    _capi.check(L.gla_vol_even_gen(ctypes.byref(v), ctypes.byref(w)))
    assert tuple(w.cel) == (4, 4, 6)
    assert [(w.scl_num[a], w.scl_den[a]) for a in range(3)] == [(3, 128), (1, 32), (5, 192)]
    # 3 cells of 1/32 become 4 of 3/128, the scale of the other volume
    t = gila.GlaVol((3, 2, 2), ((1, 32), (3, 128), (3, 128)))._c()
    s = gila.GlaVol((4, 2, 2), ((3, 128),) * 3, (1, 0, 0))._c()
    d = _capi.gla_dims()
    _capi.check(L.gla_opr_dims_for(ctypes.byref(t), ctypes.byref(s), ctypes.byref(d)))
    assert tuple(d.trg_cel) == (4, 2, 2) and tuple(d.brn) == (4, 2, 2)


def test_partition_dims(gila):
    """GlaExtInf (src/glaMem.jl:190-216): the finer volume splits into
    ratio-strided partitions of the coarser lattice."""
    from gila_amd import _capi
    L = gila.lib()
    d = _capi.gla_dims()
    t = gila.GlaVol((4, 6, 4), ((1, 16), (1, 32), (1, 32)))._c()
    s = gila.GlaVol((8, 6, 4), ((1, 32),) * 3, (1, 0, 0))._c()
    _capi.check(L.gla_opr_dims_for(ctypes.byref(t), ctypes.byref(s), ctypes.byref(d)))
    assert tuple(d.src_div) == (2, 1, 1) and tuple(d.trg_div) == (1, 1, 1)
    assert tuple(d.src_cel) == (4, 6, 4) and tuple(d.trg_cel) == (4, 6, 4)
    assert tuple(d.brn) == (4, 6, 4) and tuple(d.sym) == (0, 0, 0) and tuple(d.fur_ext) == (4, 6, 4)
    _capi.check(L.gla_opr_dims_for(ctypes.byref(s), ctypes.byref(t), ctypes.byref(d)))
    assert tuple(d.trg_div) == (2, 1, 1) and tuple(d.src_div) == (1, 1, 1)
    # source cells not a multiple of the ratio
    s8 = gila.GlaVol((8, 6, 4), ((1, 32),) * 3, (1, 0, 0))._c()
    t3 = gila.GlaVol((4, 6, 4), ((3, 32), (1, 32), (1, 32)))._c()
    assert L.gla_opr_dims_for(ctypes.byref(t3), ctypes.byref(s8), ctypes.byref(d)) == _capi.GLA_ERR_SIZE
    assert "multiple of the scale ratio" in L.gla_last_error().decode()
    # partition cells with an odd sum
    s6 = gila.GlaVol((6, 6, 4), ((1, 32),) * 3, (1, 0, 0))._c()
    assert L.gla_opr_dims_for(ctypes.byref(t), ctypes.byref(s6), ctypes.byref(d)) == _capi.GLA_ERR_SIZE
    assert "partition" in L.gla_last_error().decode()
    # a volume smaller than one coarse cell
    s1 = gila.GlaVol((1, 6, 4), ((1, 32),) * 3, (1, 0, 0))._c()
    t1 = gila.GlaVol((2, 6, 4), ((1, 16), (1, 32), (1, 32)))._c()
    assert L.gla_opr_dims_for(ctypes.byref(t1), ctypes.byref(s1), ctypes.byref(d)) == _capi.GLA_ERR_SCALE
    assert "smaller than cell" in L.gla_last_error().decode()


def test_non_complex_settyp_raises_argument_error(gila):
    # test/oprTest.jl:4-5
    import numpy as np
    with pytest.raises(gila.ArgumentError):
        gila.GlaOpr((1, 1, 1), ((1, 1),) * 3, (0, 0, 0), setTyp=np.float64)


def test_volume_regeneration_for_odd_sums(gila):
    v = gila.GlaVol((3, 4, 5), ((1, 10),) * 3)
    w = gila.glaVolEveGen(v)
    assert w.cel == (4, 4, 6)
    assert w.scl[0] * 4 == v.scl[0] * 3 and w.scl[2] * 6 == v.scl[2] * 5
