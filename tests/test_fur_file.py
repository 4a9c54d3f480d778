"""Green-data files (gla_opr_save / gla_opr_load): the explicit replacement
of the reference's Julia-serialized preload cache (examples/GilaOperators.jl:
28-55, test/runtests.jl:38-49).  Malformed files are rejected before any
device work (CPU tests); saved operators reload to the same Green data and
the same matvec (GPU tests), including a re-embedded external operator."""
import struct

import numpy as np
import pytest

H = (1, 32)


def _load_status(gila, path):
    import ctypes
    h = ctypes.c_void_p()
    return gila.lib().gla_opr_load(str(path).encode(), 0, 0, ctypes.byref(h))


def test_rejects_malformed_files(gila, tmp_path):
    p = tmp_path / "x.gfur"
    p.write_bytes(b"")
    assert _load_status(gila, p) == 1                       # shorter than the header (SIZE)
    p.write_bytes(b"NOTGILA!" + bytes(504))
    assert _load_status(gila, p) == 7                       # wrong magic (ARG)
    p.write_bytes(b"GILAFUR1" + struct.pack("<II", 2, 512) + bytes(496))
    assert _load_status(gila, p) == 7                       # unknown version
    assert _load_status(gila, tmp_path / "missing.gfur") == 7
    assert "missing.gfur" in gila.lib().gla_last_error().decode()


@pytest.mark.gpu
@pytest.mark.parametrize("typ", [np.complex128, np.complex64])
def test_roundtrip_self_operator(gila, orc, gpu, typ, tmp_path):
    print(f"case fur file self {np.dtype(typ).name}", flush=True)
    N = (16, 12, 8)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), gila.GlaVol(N, (H,) * 3), setTyp=typ)
    path = tmp_path / "self.gfur"
    mem.save(path)
    back = gila.GlaOprMem.load(path)
    assert back.setTyp == typ and back.trgVol == mem.trgVol and back.cmpInf.frqPhz == mem.cmpInf.frqPhz
    assert all(np.array_equal(a, b) for a, b in zip(mem.egoFur, back.egoFur))
    x = orc.complex_gaussian(N + (3,), seed=61).astype(typ)
    assert np.array_equal(gila.egoOpr_(mem, x), gila.egoOpr_(back, x))
    if typ == np.complex128:
        conv = gila.GlaOprMem.load(path, setTyp=np.complex64)
        assert conv.setTyp == np.complex64
        assert all(np.array_equal(a.astype(np.complex64), b) for a, b in zip(mem.egoFur, conv.egoFur))
    # any flipped payload byte is caught by the checksum
    raw = bytearray(path.read_bytes())
    raw[600] ^= 0x40
    path.write_bytes(bytes(raw))
    with pytest.raises(gila.GlaError) as ei:
        gila.GlaOprMem.load(path)
    assert "checksum" in str(ei.value)
    path.write_bytes(bytes(raw[:-8]))
    with pytest.raises(gila.GlaError) as ei:
        gila.GlaOprMem.load(path)
    assert ei.value.status == 1


@pytest.mark.gpu
def test_roundtrip_reembedded_external(gila, orc, gpu, tmp_path):
    print("case fur file re-embedded external", flush=True)
    tgt = gila.GlaVol((8, 12, 10), (H,) * 3, (0, 0, 0))
    src = gila.GlaVol((4, 2, 4), (H,) * 3, (0, 0, H))
    mem = gila.GlaOprMem(gila.GlaKerOpt(False), tgt, src, setTyp=np.complex128)
    assert tuple(mem._dims.dev_brn) != tuple(mem._dims.brn)
    path = tmp_path / "ext.gfur"
    mem.save(path)
    back = gila.GlaOprMem.load(path)
    assert back.srcVol == src and back.trgVol == tgt
    x = orc.complex_gaussian((4, 2, 4, 3), seed=62)
    assert np.array_equal(gila.egoOpr_(mem, x), gila.egoOpr_(back, x))
