"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle
(oracle/_build/libgila_oracle.so, built from oracle/gla_oracle.c).

Only the tests, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
use this module, and only as the checker.  See gla_oracle.c for what each
function restates from the reference.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libgila_oracle.so")


def build():
    subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        for name, args in {
            "gor_dims": [vp, vp, vp, vp, vp],
            "gor_fur_from_crc": [vp, vp, vp, vp, vp],
            "gor_matvec_c128": [vp, vp, vp, vp, vp, vp, i64],
            "gor_matvec_c64": [vp, vp, vp, vp, vp, vp, i64],
            "gor_brute_crc": [vp, vp, vp, vp, vp],
            "gor_crc_self": [vp, vp, vp],
            "gor_toe_model": [i32, vp, vp, d, d, i32, i32, vp],
            "gor_crc_ext": [i32, vp, vp, vp, vp, d, d, i32, i32, vp],
            "gor_crc_prt": [vp, vp, vp, vp, vp, vp, d, d, vp],
            "gor_tensor_mixed": [vp, vp, vp, d, d, vp],
            "gor_tensor_ref": [vp, vp, d, d, vp],
            "gor_fft_axis_c128": [vp, vp, i64, i32, i32],
            "gor_gl_rule": [i32, vp, vp],
            "gor_gl_pm1": [i32, vp, vp],
            "gor_wek_s": [vp, d, d, i32, vp],
            "gor_wek_e": [vp, d, d, i32, vp],
            "gor_wek_v": [vp, d, d, i32, vp],
            "gor_r_srf": [i32, d, d, d, vp],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.gor_set_singular.argtypes = [vp]
        L.gor_set_singular.restype = None
        L.gor_set_threads.argtypes = [i32]
        L.gor_set_threads.restype = None
        L.gor_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _i64(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int64))


def _i32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    lib().gor_set_threads(int(n))


def num_threads():
    return lib().gor_num_threads()


def dims(S, T, sym):
    B, U = np.zeros(3, np.int64), np.zeros(3, np.int64)
    if lib().gor_dims(_p(_i64(S)), _p(_i64(T)), _p(_i32(sym)), _p(B), _p(U)):
        raise ValueError("S+T must be even and positive on every axis")
    return tuple(int(b) for b in B), tuple(int(u) for u in U)


_SNG = {}


def singular_values(scl, zeta=1.0 + 0j, order=32):
    """[wS(3), wE(9), wV(6)] divided by the cell volume, as egoFunSng! uses
    them (src/glaGen.jl:117-121); cached per (scl, zeta, order)."""
    key = (tuple(float(v) for v in scl), complex(zeta), int(order))
    if key not in _SNG:
        vol = float(np.prod(scl))
        _SNG[key] = np.concatenate([wek("S", scl, zeta, order), wek("E", scl, zeta, order),
                                    wek("V", scl, zeta, order)]) / vol
    return _SNG[key]


def _use_singular(model, scl, zeta, order):
    if model == 3:
        w = np.ascontiguousarray(singular_values(scl, zeta, order))
        lib().gor_set_singular(_p(w))


def toe_model(N, scl, model=0, zeta=1.0 + 0j, qf=4, qn=8, int_ord=32):
    """Real-space Toeplitz tensor (3,3,N1,N2,N3) of a test Green model
    (0: point collocation, 1/2: face-pair surface integrals with fixed rules,
    3: the reference's construction with the DIRECTFN singular values of
    order int_ord)."""
    _use_singular(model, scl, zeta, int_ord)
    toe = np.zeros((3, 3) + tuple(N), dtype=np.complex128, order="F")
    lib().gor_toe_model(int(model), _p(_i64(N)), _p(np.asarray(scl, np.float64)), zeta.real, zeta.imag,
                        int(qf), int(qn), _p(toe))
    return toe


def crc_self(N, toe):
    crc = np.zeros((3, 3) + tuple(2 * n for n in N), dtype=np.complex128, order="F")
    lib().gor_crc_self(_p(_i64(N)), _p(toe), _p(crc))
    return crc


def crc_ext(S, T, scl, delta, model=0, zeta=1.0 + 0j, qf=4, qn=8, int_ord=32):
    _use_singular(model, scl, zeta, int_ord)
    crc = np.zeros((3, 3) + tuple(s + t for s, t in zip(S, T)), dtype=np.complex128, order="F")
    if lib().gor_crc_ext(int(model), _p(_i64(S)), _p(_i64(T)), _p(np.asarray(scl, np.float64)),
                         _p(np.asarray(delta, np.float64)), zeta.real, zeta.imag, int(qf), int(qn), _p(crc)):
        raise ValueError("S+T must be even on every axis")
    return crc


def tensor_mixed(sep, sT, sS, zeta=1.0 + 0j):
    """3x3 tensor between a target cell of sizes sT and a separated source
    cell of sizes sS at separation sep (egoFunOut!, src/glaGen.jl:233-246)."""
    out = np.zeros((3, 3), dtype=np.complex128, order="F")
    f64 = lambda v: np.ascontiguousarray(np.asarray(v, np.float64))
    lib().gor_tensor_mixed(_p(f64(sep)), _p(f64(sT)), _p(f64(sS)), zeta.real, zeta.imag, _p(out))
    return out


def tensor_ref(sep, scl, zeta=1.0 + 0j, int_ord=32):
    """Model-3 tensor (3x3) at one separation of equal cells."""
    _use_singular(3, scl, zeta, int_ord)
    out = np.zeros((3, 3), dtype=np.complex128, order="F")
    f64 = lambda v: np.ascontiguousarray(np.asarray(v, np.float64))
    lib().gor_tensor_ref(_p(f64(sep)), _p(f64(scl)), zeta.real, zeta.imag, _p(out))
    return out


def crc_prt(S, T, sT, sS, M, delta, zeta=1.0 + 0j):
    """Circulant (3,3,S+T) of one partition pair: separations delta + n M on
    the lattice of step M, target cells sT, source cells sS."""
    crc = np.zeros((3, 3) + tuple(s + t for s, t in zip(S, T)), dtype=np.complex128, order="F")
    f64 = lambda v: np.ascontiguousarray(np.asarray(v, np.float64))
    if lib().gor_crc_prt(_p(_i64(S)), _p(_i64(T)), _p(f64(sT)), _p(f64(sS)), _p(f64(M)), _p(f64(delta)),
                         zeta.real, zeta.imag, _p(crc)):
        raise ValueError("S+T must be even on every axis")
    return crc


def partitions(div):
    """Offsets of the partitions of a divisor grid, x fastest (the
    CartesianIndices order of GlaExtInf's srcPar / trgPar,
    src/glaMem.jl:211-212)."""
    return [(i, j, k) for k in range(div[2]) for j in range(div[1]) for i in range(div[0])]


def partition_fur(Tfull, Sfull, sT, sS, orgT, orgS, zeta=1.0 + 0j):
    """Fourier Green data of a partitioned external operator in the
    reference layout (8 branches of (B1,B2,B3,6,Ps,Pt), concatenated flat):
    per (source, target) partition pair the circulant of genEgoExt! on the
    partition volumes (src/glaMemSup.jl:64-89), then the Fourier storage of
    src/glaMemSup.jl:96-156 without symmetric truncation.  Returns (fur,
    S, T, srcDiv, trgDiv) with S, T the partition cell counts."""
    M = [max(a, b) for a, b in zip(sT, sS)]
    tdiv = [int(round(m / a)) for m, a in zip(M, sT)]
    sdiv = [int(round(m / b)) for m, b in zip(M, sS)]
    T = [c // v for c, v in zip(Tfull, tdiv)]
    S = [c // v for c, v in zip(Sfull, sdiv)]
    t0 = [o - h * (c - 1) / 2 for o, h, c in zip(orgT, sT, Tfull)]
    s0 = [o - h * (c - 1) / 2 for o, h, c in zip(orgS, sS, Sfull)]
    blocks = []
    for ot in partitions(tdiv):
        for os_ in partitions(sdiv):
            delta = [a + i * h - b - j * g for a, i, h, b, j, g in zip(t0, ot, sT, s0, os_, sS)]
            crc = crc_prt(S, T, sT, sS, M, delta, zeta)
            blocks.append(fur_from_crc(S, T, (0, 0, 0), crc))
    B, U = dims(S, T, (0, 0, 0))
    per = U[0] * U[1] * U[2] * 6
    P = len(blocks)
    fur = np.zeros(8 * per * P, dtype=np.complex128)
    for p, f in enumerate(blocks):
        for b in range(8):
            fur[(b * P + p) * per:(b * P + p + 1) * per] = f[b * per:(b + 1) * per]
    return fur, tuple(S), tuple(T), tuple(sdiv), tuple(tdiv)


def partition_matvec(S, T, sdiv, tdiv, fur, x, dtype=np.complex128):
    """The partitioned matvec: source partitions gathered from the whole
    source field x (genPrtHst, src/glaActSup.jl:4-18), every (source,
    target) pair's leaves summed into its target partition (egoBrnHst!,
    src/glaAct.jl:158-164 -- the sum commutes with the inverse transforms)
    and the target partitions merged into the whole target field (mrgPrtHst,
    src/glaActSup.jl:939-953).  x is (S*sdiv..., 3) column-major."""
    x = np.asarray(x)
    Tf = tuple(t * v for t, v in zip(T, tdiv))
    y = np.zeros(Tf + (3,), dtype=dtype, order="F")
    B, U = dims(S, T, (0, 0, 0))
    per = U[0] * U[1] * U[2] * 6
    sp, tp = partitions(sdiv), partitions(tdiv)
    P = len(sp) * len(tp)
    for it, ot in enumerate(tp):
        acc = np.zeros(tuple(T) + (3,), dtype=dtype, order="F")
        for js, os_ in enumerate(sp):
            p = js + len(sp) * it
            f = np.concatenate([fur[(b * P + p) * per:(b * P + p + 1) * per] for b in range(8)])
            xs = x[os_[0]::sdiv[0], os_[1]::sdiv[1], os_[2]::sdiv[2], :]
            acc += matvec(S, T, (0, 0, 0), f, xs, dtype=dtype)
        y[ot[0]::tdiv[0], ot[1]::tdiv[1], ot[2]::tdiv[2], :] = acc
    return y


def fur_from_crc(S, T, sym, crc):
    """The 8 Fourier branches (reference layout), concatenated flat."""
    B, U = dims(S, T, sym)
    fur = np.zeros(8 * U[0] * U[1] * U[2] * 6, dtype=np.complex128)
    rc = lib().gor_fur_from_crc(_p(_i64(S)), _p(_i64(T)), _p(_i32(sym)), _p(np.asfortranarray(crc)), _p(fur))
    if rc == 2:
        raise FloatingPointError("Fourier transform of circulant contains non-numeric values.")
    if rc:
        raise ValueError("bad dimensions")
    return fur


def fur_branches(fur, U):
    per = U[0] * U[1] * U[2] * 6
    return [fur[b * per:(b + 1) * per].reshape(tuple(U) + (6,), order="F") for b in range(8)]


def matvec(S, T, sym, fur, x, dtype=np.complex128):
    """Restated egoBrnHst! matvec; x is (S...,3) or (S...,3,nrhs) column-major."""
    x = np.asfortranarray(x, dtype=dtype)
    nrhs = 1 if x.ndim == 4 else x.shape[4]
    y = np.zeros(tuple(T) + (3,) + ((nrhs,) if x.ndim == 5 else ()), dtype=dtype, order="F")
    f = np.ascontiguousarray(fur, dtype=dtype)
    fn = lib().gor_matvec_c128 if dtype == np.complex128 else lib().gor_matvec_c64
    if fn(_p(_i64(S)), _p(_i64(T)), _p(_i32(sym)), _p(f), _p(x), _p(y), nrhs):
        raise ValueError("bad dimensions")
    return y


def brute(S, T, crc, x):
    x = np.asfortranarray(x, dtype=np.complex128)
    y = np.zeros(tuple(T) + (3,), dtype=np.complex128, order="F")
    lib().gor_brute_crc(_p(_i64(S)), _p(_i64(T)), _p(np.asfortranarray(crc)), _p(x), _p(y))
    return y


def fft_axis(v, axis, sign=-1):
    """In-place batched FFT along `axis` (0..2) of a (d0,d1,d2[,ncomp]) array."""
    a = np.asfortranarray(v, dtype=np.complex128).copy(order="F")
    d = a.shape[:3]
    nc = 1 if a.ndim == 3 else a.shape[3]
    lib().gor_fft_axis_c128(_p(a), _p(_i64(d)), nc, int(axis), int(sign))
    return a


def gl_rule(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().gor_gl_rule(int(n), _p(x), _p(w))
    return x, w


def wek(kind, scl, zeta=1.0 + 0j, order=32):
    """The reference's singular panel values (wekS / wekE / wekV,
    src/glaInt.jl:75-267) for cell sizes scl (wavelengths), before the
    division by the cell volume: 3, 9 or 6 complex values."""
    n = {"S": 3, "E": 9, "V": 6}[kind]
    out = np.zeros(n, np.complex128)
    fn = {"S": lib().gor_wek_s, "E": lib().gor_wek_e, "V": lib().gor_wek_v}[kind]
    fn(_p(np.asarray(scl, np.float64)), zeta.real, zeta.imag, int(order), _p(out))
    return out


def r_srf(which, la, lb, lc=0.0):
    """Closed-form static panel integrals at zeta = 1: which 0 rSrfSlf(la, lb),
    1 rSrfEdgFlt(la, lb), 2 rSrfEdgCrn(la, lb, lc) (src/glaInt.jl:410-488)."""
    out = np.zeros(2)
    lib().gor_r_srf(int(which), float(la), float(lb), float(lc), _p(out))
    return complex(out[0], out[1])


def gl_pm1(n):
    """Gauss-Legendre nodes and weights on [-1, 1], ascending (gauQud)."""
    x, w = np.zeros(n), np.zeros(n)
    lib().gor_gl_pm1(int(n), _p(x), _p(w))
    return x, w


# ---------------------------------------------------------------------------
# portable seeded inputs: SplitMix64 -> uniform doubles -> Box-Muller
# ---------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n outputs of SplitMix64 started from `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def complex_gaussian(shape, seed):
    """Standard complex Gaussian field (re, im ~ N(0, 1/2)), column-major,
    from SplitMix64 + Box-Muller."""
    n = int(np.prod(shape))
    z = splitmix64(seed, 2 * n)
    u = ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-np.log(u1))        # |g| with E|g|^2 = 1
    g = r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)
    return np.asfortranarray(g.reshape(shape, order="F"))


def self_operator_fur(N, h=1.0 / 32, model=0, zeta=1.0 + 0j, qf=4, qn=8, int_ord=32):
    """Fourier branches of a self operator on N cells from a test Green model."""
    hs = tuple(h) if np.ndim(h) else (h, h, h)
    toe = toe_model(N, hs, model=model, zeta=zeta, qf=qf, qn=qn, int_ord=int_ord)
    crc = crc_self(N, toe)
    return fur_from_crc(N, N, (1, 1, 1), crc), crc


def rel_l2(a, b):
    a = np.asarray(a).ravel(order="F")
    b = np.asarray(b).ravel(order="F")
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
