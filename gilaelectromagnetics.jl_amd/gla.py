"""Host-side mirror of GilaElectromagnetics.jl's operator surface.

Names, argument meaning and error behaviour follow the reference
(src/glaMem.jl, src/glaMemSup.jl, src/glaAct.jl, src/glaLinAlg.jl); the
matvec itself runs in libgila_amd.so on the GPU.

Arrays follow the reference's Julia layout: a field on a volume with cells
(c1, c2, c3) has shape (c1, c2, c3, 3) in column-major (Fortran) order, which
is the same memory as a flat vector of length 3*c1*c2*c3.  A fifth trailing
axis batches right-hand sides.  numpy arrays (host) and torch tensors on a
ROCm device are both accepted; torch tensors are used where they lie, numpy
arrays are copied to the GPU and back.
"""
from fractions import Fraction
import ctypes

import numpy as np

from . import _capi as capi
from ._capi import GlaError


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


def _rat(v):
    if isinstance(v, Fraction):
        return v
    if isinstance(v, (tuple, list)) and len(v) == 2:
        return Fraction(int(v[0]), int(v[1]))
    if isinstance(v, float):
        return Fraction(v).limit_denominator(1 << 30)
    return Fraction(v)


def _rat3(v):
    return tuple(_rat(a) for a in v)


class GlaVol:
    """Cuboid volume (src/glaMem.jl:12-19, constructor :136-147).

    cel: cells per axis; scl: cell size (wavelengths, rational);
    org: centre (wavelengths, rational); grdScl: grid step (defaults to scl).
    grd holds (start, step, stop) of the cell centres per axis.
    """

    def __init__(self, cel, scl, org=(0, 0, 0), grdScl=None):
        cel = tuple(int(c) for c in cel)
        scl = _rat3(scl)
        org = _rat3(org)
        grdScl = scl if grdScl is None else _rat3(grdScl)
        if not all(s <= g for s, g in zip(scl, grdScl)):
            raise GlaError(capi.GLA_ERR_SCALE, "The cell scale must be smaller than the grid scale to avoid "
                                              "partially overlapping basis elements.")
        brd = tuple(g * (Fraction(c // 2) - Fraction(1 if c % 2 == 0 else 0, 2)) for g, c in zip(grdScl, cel))
        self.cel = cel
        self.scl = scl
        self.org = org
        self.grd = [(o - b, g, o + b) for o, b, g in zip(org, brd, grdScl)]

    def __eq__(self, other):
        return isinstance(other, GlaVol) and self.cel == other.cel and self.scl == other.scl and \
            self.org == other.org and self.grd == other.grd

    def __hash__(self):
        return hash((self.cel, self.scl, self.org))

    def __repr__(self):
        return f"GlaVol(cel={self.cel}, scl={tuple(str(s) for s in self.scl)}, org={tuple(str(o) for o in self.org)})"

    def _c(self):
        v = capi.gla_vol()
        for d in range(3):
            v.cel[d] = self.cel[d]
            v.scl_num[d], v.scl_den[d] = self.scl[d].numerator, self.scl[d].denominator
            v.org_num[d], v.org_den[d] = self.org[d].numerator, self.org[d].denominator
        return v


def glaVolEveGen(vol):
    """Regenerate a volume with an even cell count per axis, rescaling the
    cell size (src/glaMem.jl:152-176)."""
    par = [c % 2 == 1 for c in vol.cel]
    if not any(par):
        return vol
    print("Warning! A volume has been regenerated to have an even number of cells---the size of a cell has "
          "changed. The sum of the number of source and target cells must be even for the algorithm to function.")
    new = tuple(c + int(p) for c, p in zip(vol.cel, par))
    scl = tuple(Fraction(s.numerator * c, s.denominator * n) for s, c, n in zip(vol.scl, vol.cel, new))
    grd = tuple(Fraction(g[1].numerator * c, g[1].denominator * n) for g, c, n in zip(vol.grd, vol.cel, new))
    return GlaVol(new, scl, vol.org, grd)


class GlaExtInf:
    """Source/target alignment (src/glaMem.jl:39-48, constructor :190-216)."""

    def __init__(self, trgVol, srcVol):
        from math import gcd
        ok = all(((s / t).denominator == 1) or ((t / s).denominator == 1) for s, t in zip(srcVol.scl, trgVol.scl))
        if not ok:
            raise GlaError(capi.GLA_ERR_SCALE, "Volume pair must share a common scale grid.")

        def rgcd(a, b):
            return Fraction(gcd(a.numerator * b.denominator, b.numerator * a.denominator), a.denominator * b.denominator)

        def rlcm(a, b):
            return a * b / rgcd(a, b)
        self.minScl = tuple(rgcd(s, t) for s, t in zip(srcVol.scl, trgVol.scl))
        maxScl = tuple(rlcm(s, t) for s, t in zip(srcVol.scl, trgVol.scl))
        self.trgDiv = tuple(int(m // t) for m, t in zip(maxScl, trgVol.scl))
        self.srcDiv = tuple(int(m // s) for m, s in zip(maxScl, srcVol.scl))
        self.trgCel = tuple(c // v for c, v in zip(trgVol.cel, self.trgDiv))
        self.srcCel = tuple(c // v for c, v in zip(srcVol.cel, self.srcDiv))
        if 0 in self.trgCel or 0 in self.srcCel:
            raise GlaError(capi.GLA_ERR_SCALE, "Volume sizes are incompatible---volume smaller than cell.")


class GlaKerOpt:
    """Operator options (src/glaMem.jl:62-70; simplified constructor :226-234).

    GlaKerOpt(devStt) gives the reference's defaults.  devMod selects where
    the caller's arrays live (True: GPU tensors, False: host arrays); the
    matvec itself always runs on the GPU.  numTrd / numBlk are accepted and
    ignored (launch geometry is chosen per transform length).
    """

    def __init__(self, frqPhz=1.0 + 0.0j, intOrd=32, adjMod=False, devMod=False, numTrd=(), numBlk=(), device=0):
        if isinstance(frqPhz, bool):   # GlaKerOpt(devStt) form
            devMod, frqPhz = frqPhz, 1.0 + 0.0j
            numTrd, numBlk = ((128, 2, 1), (1, 128, 256)) if devMod else ((), ())
        self.frqPhz = complex(frqPhz)
        self.intOrd = int(intOrd)
        self.adjMod = bool(adjMod)
        self.devMod = bool(devMod)
        self.numTrd = tuple(numTrd)
        self.numBlk = tuple(numBlk)
        self.device = int(device)

    def _c(self, dtype):
        o = capi.gla_opts()
        o.frq_phz_re, o.frq_phz_im = self.frqPhz.real, self.frqPhz.imag
        o.int_ord = self.intOrd
        o.adj_mod = int(self.adjMod)
        o.dtype = dtype
        o.device = self.device
        return o


def _dtype_code(setTyp):
    t = np.dtype(setTyp) if not _is_torch_dtype(setTyp) else np.dtype(_torch_to_np(setTyp))
    if t == np.complex64:
        return capi.GLA_C64
    if t == np.complex128:
        return capi.GLA_C128
    raise ArgumentError("setTyp must be a subtype of Complex")


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


def _is_torch_dtype(t):
    return type(t).__module__.startswith("torch")


def _torch_to_np(t):
    import torch
    return {torch.complex64: np.complex64, torch.complex128: np.complex128}.get(t, None)


class GlaOprMem:
    """Operator storage (src/glaMem.jl:93-106; constructor src/glaMemSup.jl:16-171).

    GlaOprMem(cmpInf, trgVol, srcVol=None, egoFur=None, setTyp=complex128).
    With egoFur (the 8 Fourier branches in the reference layout, each
    (U1,U2,U3,6)) the Green data is taken as given; without it the library
    computes it from the geometry.
    """

    def __init__(self, cmpInf, trgVol, srcVol=None, egoFur=None, setTyp=np.complex128, _handle=None):
        self.cmpInf = cmpInf
        if _handle is not None:
            self._h = _handle
            self.trgVol, self.srcVol = trgVol, srcVol
            self._load_dims()
            return
        try:
            code = _dtype_code(setTyp)
        except TypeError:
            raise ArgumentError("setTyp must be a subtype of Complex")
        if srcVol is None or trgVol == srcVol:
            srcVol = trgVol
        elif any((t + s) % 2 for t, s in zip(trgVol.cel, srcVol.cel)):
            trgVol = glaVolEveGen(trgVol)
            srcVol = glaVolEveGen(srcVol)
        self.trgVol, self.srcVol = trgVol, srcVol
        self.mixInf = GlaExtInf(trgVol, srcVol)
        L = capi.lib()
        h = ctypes.c_void_p()
        opts = cmpInf._c(code)
        tv = trgVol._c()
        sv = None if srcVol is trgVol else srcVol._c()
        svp = None if sv is None else ctypes.byref(sv)
        if egoFur is None:
            capi.check(L.gla_opr_create(ctypes.byref(opts), ctypes.byref(tv), svp, ctypes.byref(h)))
        else:
            flat, fcode = _flatten_fur(egoFur)
            capi.check(L.gla_opr_create_from_fur(ctypes.byref(opts), ctypes.byref(tv), svp,
                                                 flat.ctypes.data_as(ctypes.c_void_p), fcode, flat.size,
                                                 ctypes.byref(h)))
        self._h = h
        self._load_dims()

    def _load_dims(self):
        d = capi.gla_dims()
        capi.check(capi.lib().gla_opr_get_dims(self._h, ctypes.byref(d)))
        self._dims = d
        self.dimInf = tuple(d.brn)
        self.setTyp = np.complex64 if d.dtype == capi.GLA_C64 else np.complex128
        if not hasattr(self, "mixInf"):
            self.mixInf = GlaExtInf(self.trgVol, self.srcVol)

    @property
    def egoFur(self):
        """The 8 Fourier branches, each (U1,U2,U3,6) column-major, or
        (U1,U2,U3,6,Ps,Pt) over the (source, target) partition pairs of an
        operator whose cell scales differ."""
        U = tuple(self._dims.fur_ext)
        Ps, Pt = int(np.prod(self._dims.src_div)), int(np.prod(self._dims.trg_div))
        per = U[0] * U[1] * U[2] * 6 * Ps * Pt
        buf = np.empty(8 * per, dtype=self.setTyp)
        capi.check(capi.lib().gla_opr_export_fur(self._h, buf.ctypes.data_as(ctypes.c_void_p), 8 * per))
        shp = U + (6,) + ((Ps, Pt) if Ps * Pt > 1 else ())
        return [buf[b * per:(b + 1) * per].reshape(shp, order="F") for b in range(8)]

    def save(self, path):
        """Write the Green data and the geometry it belongs to (the library's
        Green-data file; the reference's preload cache uses Julia
        serialization, examples/GilaOperators.jl:28-55)."""
        capi.check(capi.lib().gla_opr_save(self._h, str(path).encode()))

    @classmethod
    def load(cls, path, setTyp=None, device=0):
        """An operator from a Green-data file; setTyp None keeps the stored
        element type."""
        code = 0 if setTyp is None else _dtype_code(setTyp)
        L = capi.lib()
        h = ctypes.c_void_p()
        capi.check(L.gla_opr_load(str(path).encode(), code, int(device), ctypes.byref(h)))
        tv, sv = capi.gla_vol(), capi.gla_vol()
        opts = capi.gla_opts()
        capi.check(L.gla_opr_get_vols(h, ctypes.byref(tv), ctypes.byref(sv)))
        capi.check(L.gla_opr_get_opts(h, ctypes.byref(opts)))

        def vol(v):
            return GlaVol(tuple(v.cel), tuple(Fraction(v.scl_num[a], v.scl_den[a]) for a in range(3)),
                          tuple(Fraction(v.org_num[a], v.org_den[a]) for a in range(3)))
        trg, src = vol(tv), vol(sv)
        cmp = GlaKerOpt(complex(opts.frq_phz_re, opts.frq_phz_im), opts.int_ord, bool(opts.adj_mod), True,
                        device=opts.device)
        return cls(cmp, trg, src if src != trg else trg, _handle=h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                capi.lib().gla_opr_destroy(h)
            except Exception:
                pass
            self._h = None


def _flatten_fur(egoFur):
    arrs = [np.asarray(a) for a in egoFur]
    if len(arrs) != 8:
        raise GlaError(capi.GLA_ERR_SIZE, "egoFur must hold 8 branches")
    dt = arrs[0].dtype
    code = _dtype_code(dt)
    flat = np.concatenate([np.asarray(a, dtype=dt).reshape(-1, order="F") for a in arrs])
    return np.ascontiguousarray(flat), code


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _colmajor_torch(x):
    """x as a tensor whose memory is column-major for its shape."""
    nd = x.dim()
    perm = list(range(nd - 1, -1, -1))
    if x.permute(*perm).is_contiguous():
        return x
    return x.permute(*perm).contiguous().permute(*perm)


def egoOpr_(egoMem, actVec):
    """egoOpr! (src/glaAct.jl:33-49): apply G to actVec of shape
    (srcCel..., 3) (or (srcCel..., 3, nrhs)); returns a new (trgCel..., 3)
    array.  Unlike the reference, actVec is left untouched.  For cell scales
    that differ the shapes are those of the whole volumes (partition cells
    times divisors): the partitions are gathered and merged inside the
    library (genPrt / mrgPrt, src/glaActSup.jl:4-18, 939-953)."""
    d = egoMem._dims
    S = tuple(int(c) * int(v) for c, v in zip(d.src_cel, d.src_div))
    T = tuple(int(c) * int(v) for c, v in zip(d.trg_cel, d.trg_div))
    shp = tuple(actVec.shape)
    if shp[:4] != S + (3,) or len(shp) not in (4, 5):
        raise GlaError(capi.GLA_ERR_SIZE, "Size of input vector does not match size of source volume. "
                                         "Required data layout is (srcCelX, srcCelY, srcCelZ, 3).")
    nrhs = shp[4] if len(shp) == 5 else 1
    oshape = T + (3,) + ((nrhs,) if len(shp) == 5 else ())
    return _apply(egoMem, actVec, oshape, nrhs)


def _apply(mem, x, oshape, nrhs, out=None, alpha=1.0, beta=0.0, axpby=False):
    L = capi.lib()
    nx = int(np.prod(x.shape))
    ny = int(np.prod(oshape))
    if _is_torch(x):
        import torch
        if x.dtype != {capi.GLA_C64: torch.complex64, capi.GLA_C128: torch.complex128}[mem._dims.dtype]:
            raise AssertionError("Input array must have the same element type as the operator.")
        if not x.is_cuda:
            raise GlaError(capi.GLA_ERR_ARG, "torch tensors must be on the GPU")
        xc = _colmajor_torch(x)
        if out is None:
            y = torch.empty(tuple(reversed(oshape)), dtype=x.dtype, device=x.device).permute(
                *range(len(oshape) - 1, -1, -1))
        else:
            y = out
            if not y.permute(*range(y.dim() - 1, -1, -1)).is_contiguous():
                raise GlaError(capi.GLA_ERR_ARG, "output tensor must be column-major")
        st = torch.cuda.current_stream(x.device).cuda_stream
        if axpby:
            capi.check(L.gla_opr_apply_axpby(mem._h, ctypes.c_void_p(xc.data_ptr()), nx,
                                             ctypes.c_void_p(y.data_ptr()), ny, nrhs,
                                             complex(alpha).real, complex(alpha).imag,
                                             complex(beta).real, complex(beta).imag, ctypes.c_void_p(st)))
        else:
            capi.check(L.gla_opr_apply(mem._h, ctypes.c_void_p(xc.data_ptr()), nx, ctypes.c_void_p(y.data_ptr()),
                                       ny, nrhs, ctypes.c_void_p(st)))
        torch.cuda.current_stream(x.device).synchronize()   # egoOpr! blocks until done
        return y
    xa = np.asarray(x)
    if xa.dtype != mem.setTyp:
        raise AssertionError("Input array must have the same element type as the operator.")
    xa = np.asfortranarray(xa)
    y = np.empty(oshape, dtype=mem.setTyp, order="F")
    capi.check(L.gla_opr_apply_host(mem._h, xa.ctypes.data_as(ctypes.c_void_p), nx,
                                    y.ctypes.data_as(ctypes.c_void_p), ny, nrhs))
    if axpby:
        tgt = out
        res = complex(alpha) * y + (0 if beta == 0 else complex(beta) * tgt)
        tgt[...] = res
        return tgt
    return y


class GlaOpr:
    """Operator wrapper (src/glaMem.jl:116-119; constructors src/glaMemSup.jl:307-349).

    GlaOpr(mem)                                          wrap a GlaOprMem
    GlaOpr(cel, scl, org=(0,0,0), useGpu=False, setTyp=complex128)      self
    GlaOpr(celSrc, sclSrc, orgSrc, celTrg, sclTrg, orgTrg, useGpu=..., setTyp=...)  external
    """

    def __init__(self, *args, useGpu=False, setTyp=np.complex128, egoFur=None):
        if len(args) == 1 and isinstance(args[0], GlaOprMem):
            self.mem = args[0]
            return
        try:
            _dtype_code(setTyp)
        except (TypeError, ArgumentError):
            raise ArgumentError("setTyp must be a subtype of Complex")
        opt = GlaKerOpt(bool(useGpu))
        if len(args) in (2, 3):
            cel, scl = args[0], args[1]
            org = args[2] if len(args) == 3 else (0, 0, 0)
            self.mem = GlaOprMem(opt, GlaVol(cel, scl, org), setTyp=setTyp, egoFur=egoFur)
        elif len(args) == 6:
            volSrc = GlaVol(args[0], args[1], args[2])
            volTrg = GlaVol(args[3], args[4], args[5])
            self.mem = GlaOprMem(opt, volTrg, volSrc, setTyp=setTyp, egoFur=egoFur)
        else:
            raise TypeError("GlaOpr(mem) | GlaOpr(cel, scl[, org]) | GlaOpr(celSrc, sclSrc, orgSrc, celTrg, sclTrg, orgTrg)")

    # --- LinearAlgebra surface (src/glaLinAlg.jl) ---
    @property
    def shape(self):
        return size(self)

    def __mul__(self, x):
        return _opr_mul(self, x)

    __matmul__ = __mul__

    def __repr__(self):
        return show(self)


def eltype(opr):
    return opr.mem.setTyp


def size(opr, dim=None):
    s = (3 * int(np.prod(opr.mem.trgVol.cel)), 3 * int(np.prod(opr.mem.srcVol.cel)))
    return s if dim is None else s[dim - 1]


def glaSze(opr, dim=None):
    """Tensor shapes of output and input (src/glaMemSup.jl:359-371)."""
    s = (tuple(opr.mem.trgVol.cel) + (3,), tuple(opr.mem.srcVol.cel) + (3,))
    return s if dim is None else s[dim - 1]


def issymmetric(opr):
    return False


def isposdef(opr):
    return False


def ishermitian(opr):
    return False


def isdiag(opr):
    return False


def isadjoint(opr):
    return opr.mem.cmpInf.adjMod


def isselfoperator(opr):
    return opr.mem.srcVol == opr.mem.trgVol


def isexternaloperator(opr):
    return not isselfoperator(opr)


def adjoint(opr):
    """Flag-only adjoint with swapped volumes (src/glaLinAlg.jl:16-29); the
    matvec is unchanged, as in the reference."""
    h = ctypes.c_void_p()
    capi.check(capi.lib().gla_opr_adjoint(opr.mem._h, ctypes.byref(h)))
    c = opr.mem.cmpInf
    opt = GlaKerOpt(c.frqPhz, c.intOrd, not c.adjMod, c.devMod, c.numTrd, c.numBlk, c.device)
    mem = GlaOprMem(opt, opr.mem.srcVol, opr.mem.trgVol, _handle=h)
    mem.mixInf = opr.mem.mixInf
    return GlaOpr(mem)


def _opr_mul(opr, x):
    """G * x for a 4-D field or a flat vector (src/glaLinAlg.jl:33-57)."""
    if not _is_torch(x):
        xa = np.asarray(x)
        if xa.dtype != eltype(opr):
            raise AssertionError(f"Input array must have the same element type as the operator. "
                                 f"eltype(opr) = {np.dtype(eltype(opr))}")
        x = xa
    elif opr.mem.cmpInf.devMod is False and not x.is_cuda:
        pass
    src_shape = glaSze(opr, 2)
    n_in = int(np.prod(src_shape))
    if tuple(x.shape) == src_shape:
        return egoOpr_(opr.mem, x)
    total = int(np.prod(x.shape))
    if total != n_in:
        raise DimensionMismatch(f"new dimensions {src_shape} must be consistent with array size {total}")
    xr = _reshape_f(x, src_shape)
    out = egoOpr_(opr.mem, xr)
    trg_shape = glaSze(opr, 1)
    if int(np.prod(x.shape)) == int(np.prod(trg_shape)):
        return _reshape_f(out, tuple(x.shape))
    if len(x.shape) == 1:
        return _reshape_f(out, (int(np.prod(trg_shape)),))
    return out


def _reshape_f(x, shape):
    if _is_torch(x):
        xc = _colmajor_torch(x)
        return xc.permute(*range(xc.dim() - 1, -1, -1)).reshape(tuple(reversed(shape))).permute(
            *range(len(shape) - 1, -1, -1))
    return np.reshape(x, shape, order="F")


def mul_(outVec, opr, innVec, alpha=1.0, beta=0.0):
    """LinearAlgebra.mul!(y, G, x, alpha, beta) (src/glaLinAlg.jl:61-69):
    y = alpha*G*x + beta*y, in place on y."""
    S, T = glaSze(opr, 2), glaSze(opr, 1)
    if _is_torch(innVec) and innVec.is_cuda and _is_torch(outVec):
        xr = innVec if tuple(innVec.shape) == S else _reshape_f(innVec, S)
        yr = outVec if tuple(outVec.shape) == T else _reshape_f(outVec, T)
        if yr.data_ptr() != outVec.data_ptr():
            raise GlaError(capi.GLA_ERR_ARG, "output must be column-major")
        _apply(opr.mem, xr, T, 1, out=yr, alpha=alpha, beta=beta, axpby=True)
        return outVec
    res = opr * innVec
    outVec *= beta
    outVec += alpha * res
    return outVec


def show(opr):
    """Text form of an operator (src/glaLinAlg.jl:110-132)."""
    m = opr.mem
    fmt = lambda t: "×".join(str(v) for v in t)
    s = "Adjoint " if isadjoint(opr) else ""
    et = "ComplexF32" if eltype(opr) == np.complex64 else "ComplexF64"
    if isselfoperator(opr):
        s += f"Self GlaOpr for a {et} ({fmt(m.srcVol.cel)}) volume of size ({fmt(m.srcVol.scl)})λ"
    else:
        sep = ", ".join(str(t - s_) for t, s_ in zip(m.trgVol.org, m.srcVol.org))
        s += (f"External GlaOpr for {et} ({fmt(m.srcVol.cel)}) -> ({fmt(m.trgVol.cel)}) volumes of sizes "
              f"({fmt(m.srcVol.scl)})λ -> ({fmt(m.trgVol.scl)})λ with separation ({sep})λ")
    return s
