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
    b = gila.GlaVol((3, 4, 4), ((1, 32),) * 3, (1, 0, 0))._c()
    assert L.gla_opr_dims_for(ctypes.byref(a), ctypes.byref(b), ctypes.byref(d)) == _capi.GLA_ERR_SIZE
    c = gila.GlaVol((4, 4, 4), ((1, 33),) * 3, (1, 0, 0))._c()
    assert L.gla_opr_dims_for(ctypes.byref(a), ctypes.byref(c), ctypes.byref(d)) == _capi.GLA_ERR_SCALE
    assert "common scale" in L.gla_last_error().decode()


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
