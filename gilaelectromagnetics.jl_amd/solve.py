"""Scattering solves on the GPU: the Lippmann-Schwinger operator of the
reference's examples and a restarted GMRES with FP32 Arnoldi cycles inside
FP64 iterative refinement.

The operator is A x = x - chi .* (G x) (LippmannSchwinger,
examples/GilaOperators.jl:57-96), applied by libgila_amd in one launch
sequence (the medium multiply and subtraction are fused into the last
kernel, gla_ls_apply).  The reference has no solver of its own; its
examples hand the operator to IterativeSolvers' bicgstabl / gmres
(examples/scattering_guide.jl:65-68, docs/src/usage.md:370-391).

Vectors are torch tensors on the GPU holding column-major fields; a batch of
right-hand sides is a (nrhs, 3*c1*c2*c3) tensor, which is the memory layout
(c1,c2,c3,3,nrhs) the library expects, so one launch applies A to the whole
batch.  The Krylov bookkeeping (Gram-Schmidt products, Givens rotations)
uses torch / host arithmetic; the matvec is the library's.
"""
import ctypes

import numpy as np

from . import _capi as capi
from .gla import ArgumentError, DimensionMismatch, GlaOpr, _is_torch, eltype, glaSze, isselfoperator


def _torch():
    import torch
    return torch


class LippmannSchwinger:
    """A = I - chi G for a self Green operator (examples/GilaOperators.jl:57-77).

    medium: complex array of the target cells (c1,c2,c3), column-major, one
    per right-hand side if it has a trailing batch axis."""

    def __init__(self, greens_op, medium):
        torch = _torch()
        if not isinstance(greens_op, GlaOpr):
            raise ArgumentError("greens_op must be a GlaOpr")
        if not isselfoperator(greens_op):
            raise ArgumentError("the Lippmann-Schwinger operator needs a self Green operator")
        cel = tuple(glaSze(greens_op, 1)[:3])
        if tuple(medium.shape[:3]) != cel or len(medium.shape) not in (3, 4):
            raise DimensionMismatch("Green's operator and medium must have the same size.")
        tt = {np.complex64: torch.complex64, np.complex128: torch.complex128}[np.dtype(eltype(greens_op)).type]
        if _is_torch(medium):
            if medium.dtype != tt:
                raise ArgumentError("Medium must have the same element type as the Green's operator.")
            flat = medium.permute(*range(medium.dim() - 1, -1, -1)).contiguous().reshape(-1).to("cuda")
        else:
            medium = np.asarray(medium)
            if medium.dtype != eltype(greens_op):
                raise ArgumentError("Medium must have the same element type as the Green's operator.")
            flat = torch.from_numpy(medium.reshape(-1, order="F").copy()).to("cuda")
        cells = int(np.prod(cel))
        self.greens_op = greens_op
        self.nbatch = 1 if len(medium.shape) == 3 else int(medium.shape[3])
        self.medium = flat.reshape(self.nbatch, cells)
        self.n = 3 * cells

    @property
    def dtype(self):
        return self.medium.dtype

    def apply(self, x, y=None):
        """y = x - chi .* (G x) for x of shape (nrhs, n) (or (n,)) on the GPU."""
        torch = _torch()
        xb = x.reshape(-1, self.n)
        if not xb.is_contiguous():
            xb = xb.contiguous()
        nrhs = xb.shape[0]
        if self.nbatch not in (1, nrhs):
            raise DimensionMismatch("one medium per right-hand side, or one shared medium")
        if y is None:
            y = torch.empty_like(xb)
        stride = 0 if self.nbatch == 1 else self.medium.shape[1]
        st = torch.cuda.current_stream(xb.device).cuda_stream
        capi.check(capi.lib().gla_ls_apply(self.greens_op.mem._h, ctypes.c_void_p(self.medium.data_ptr()),
                                           self.medium.numel(), stride, ctypes.c_void_p(xb.data_ptr()),
                                           xb.numel(), ctypes.c_void_p(y.data_ptr()), y.numel(), nrhs,
                                           ctypes.c_void_p(st)))
        return y.reshape(x.shape)

    def __mul__(self, x):
        return self.apply(x)

    def restrict(self, idx):
        """The operator on the right-hand sides idx (a per-RHS medium is
        subset; a shared one is kept)."""
        if self.nbatch == 1:
            return self
        sub = LippmannSchwinger.__new__(LippmannSchwinger)
        sub.greens_op = self.greens_op
        sub.medium = self.medium[idx].contiguous()
        sub.nbatch = sub.medium.shape[0]
        sub.n = self.n
        return sub


class KrylovOps:
    """Batched Gram-Schmidt products and basis-row moves on device-resident
    bases through the library's Krylov kernels (gla_krylov_dots /
    gla_krylov_axpy / gla_krylov_rows).  One instance per stream: it owns the
    partial-sum workspace."""

    def __init__(self):
        self.part = None

    @staticmethod
    def _code(t):
        torch = _torch()
        return capi.GLA_C64 if t.dtype == torch.complex64 else capi.GLA_C128

    def _part(self, k, nv, n, device):
        torch = _torch()
        need = capi.lib().gla_krylov_dots_workspace(k, nv, n)
        if self.part is None or self.part.numel() * 8 < need:
            self.part = torch.empty(need // 8 + 1, dtype=torch.float64, device=device)
        return self.part

    @staticmethod
    def _ptr(t):
        return None if t is None else ctypes.c_void_p(t.data_ptr())

    def dots(self, V, k, nv, w, h, stream, w2=None, ncol=None):
        """h[r, i] = <V[r, i], w[r]> for r < k, i < nv (h complex128 (k, nv));
        w2[r] = ||w[r]||^2 (float64 device tensor) when given; ncol (int64
        device tensor of k counts): r uses only V[r, :ncol[r]], h zero beyond."""
        n = V.shape[-1]
        part = self._part(k, nv, n, V.device)
        capi.check(capi.lib().gla_krylov_dots(self._code(V), k, nv, n, self._ptr(V), V.stride(0), V.stride(1),
                                              self._ptr(w), w.stride(0), self._ptr(part), part.numel() * 8,
                                              self._ptr(h), self._ptr(w2), self._ptr(ncol), ctypes.c_void_p(stream)))

    def axpy(self, V, k, nv, h, w, s, stream, w2=None, ncol=None):
        """w[r] += s * sum_i h[r, i] V[r, i] for r < k, i < nv (i < ncol[r]
        when given); w2[r] = ||w[r]||^2 afterwards when given."""
        n = V.shape[-1]
        part = self._part(k, 0, n, V.device) if w2 is not None else None
        capi.check(capi.lib().gla_krylov_axpy(self._code(V), k, nv, n, self._ptr(V), V.stride(0), V.stride(1),
                                              self._ptr(h), h.stride(0) if h is not None else 0, self._ptr(w),
                                              w.stride(0), float(s), self._ptr(part),
                                              0 if part is None else part.numel() * 8, self._ptr(w2),
                                              self._ptr(ncol), ctypes.c_void_p(stream)))

    def next(self, V, k, idx, w, w2, stream, stage=None):
        """V[r, idx[r]] = w[r] / max(sqrt(w2[r]), 1e-30) for r < k (w2 the
        float64 device squared norms), also into stage[r] when given.  On the
        CPU (tests of the host logic) the same in torch."""
        if not V.is_cuda:
            torch = _torch()
            inv = 1.0 / torch.clamp(torch.sqrt(w2[:k]), min=1e-30)
            ar = torch.arange(k)
            v = (w[:k].to(torch.complex128) * inv[:, None]).to(V.dtype)
            V[ar, idx[:k]] = v
            if stage is not None:
                stage[:k] = v
            return
        n = V.shape[-1]
        capi.check(capi.lib().gla_krylov_next(self._code(V), k, n, self._ptr(V), V.stride(0), V.stride(1),
                                              self._ptr(idx), self._ptr(w), w.stride(0), self._ptr(w2),
                                              self._ptr(stage), 0 if stage is None else stage.stride(0),
                                              ctypes.c_void_p(stream)))

    def rows(self, V, k, idx, b, stream, scale=None):
        """scale None: b[r] = V[r, idx[r]]; otherwise V[r, idx[r]] = b[r] / scale[r]
        (idx int64, scale float64, both device tensors of k entries)."""
        n = V.shape[-1]
        capi.check(capi.lib().gla_krylov_rows(self._code(V), k, n, self._ptr(V), V.stride(0), V.stride(1),
                                              self._ptr(idx), self._ptr(b), b.stride(0), self._ptr(scale),
                                              ctypes.c_void_p(stream)))


def _stream():
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream if torch.cuda.is_available() else 0


# DGKS criterion: a Gram-Schmidt pass that leaves less than 1/sqrt(2) of the
# vector's norm is repeated (the reorthogonalization test of IterativeSolvers'
# gmres, the solver the reference's examples use)
DGKS = 0.7071067811865476


def gs_pass(kry, V, k, nv, w, stream, ncol=None, out=None):
    """One classical Gram-Schmidt pass of w[:k] against V[:k, :nv] (against
    V[r, :ncol[r]] for each r when ncol, an int64 device tensor, is given).
    Returns (coefficients complex128 (k, nv), ||w||^2 before, ||w||^2 after)
    as device tensors -- views of `out` (float64, at least k (2 nv + 2)
    entries) when given, so one copy moves all three.  On the CPU (tests of
    the host logic) the same products run in torch."""
    torch = _torch()
    if not V.is_cuda:
        Vj = V[:k, :nv].to(torch.complex128)
        if ncol is not None:
            keep = torch.arange(nv)[None, :] < ncol[:k, None]
            Vj = Vj * keep[:, :, None]
        w64 = w[:k].to(torch.complex128)
        n0 = (w64.abs() ** 2).sum(1)
        h = torch.einsum("rin,rn->ri", Vj.conj(), w64)
        w[:k] -= torch.einsum("ri,rin->rn", h, Vj).to(w.dtype)
        n1 = (w[:k].to(torch.complex128).abs() ** 2).sum(1)
        if out is not None:
            hv = out[:2 * k * nv].view(torch.complex128).view(k, nv)
            hv.copy_(h)
            out[2 * k * nv:2 * k * nv + k].copy_(n0)
            out[2 * k * nv + k:2 * k * nv + 2 * k].copy_(n1)
            return hv, out[2 * k * nv:2 * k * nv + k], out[2 * k * nv + k:2 * k * nv + 2 * k]
        return h, n0, n1
    if out is None:
        out = torch.empty(k * (2 * nv + 2), dtype=torch.float64, device=V.device)
    h = out[:2 * k * nv].view(torch.complex128).view(k, nv)
    n0 = out[2 * k * nv:2 * k * nv + k]
    n1 = out[2 * k * nv + k:2 * k * nv + 2 * k]
    kry.dots(V, k, nv, w, h, stream, w2=n0, ncol=ncol)
    kry.axpy(V, k, nv, h, w, -1.0, stream, w2=n1, ncol=ncol)
    return h, n0, n1


def orthogonalize(kry, V, k, nv, w, stream):
    """Gram-Schmidt of w[:k] against V[:k, :nv] with a DGKS-tested second
    pass; returns (coefficients complex128 (k, nv), norms of the result as a
    host array).  Synchronizes with the device to read the norms."""
    h, n0, n1 = gs_pass(kry, V, k, nv, w, stream)
    a0, a1 = np.sqrt(n0.cpu().numpy()), np.sqrt(n1.cpu().numpy())
    if (a1 < DGKS * a0).any():
        h2, _, n2 = gs_pass(kry, V, k, nv, w, stream)
        h = h + h2
        a1 = np.sqrt(n2.cpu().numpy())
    return h, a1


def combine(kry, V, r, nv, y, out, stream):
    """out += sum_i y[i] V[r, i] (y complex128 (nv,), out (n,) of V's type)."""
    torch = _torch()
    if not V.is_cuda:
        out += (y.to(torch.complex128) @ V[r, :nv].to(torch.complex128)).to(out.dtype)
        return
    kry.axpy(V[r:r + 1], 1, nv, y.reshape(1, -1).contiguous(), out.reshape(1, -1), 1.0, stream)


class SolveInfo:
    def __init__(self, nrhs):
        self.converged = np.zeros(nrhs, bool)
        self.relres = np.full(nrhs, np.inf)
        self.matvecs32 = 0
        self.matvecs64 = 0
        self.cycles = 0
        self.history = []

    def __repr__(self):
        return (f"SolveInfo(converged={self.converged.tolist()}, relres={self.relres.tolist()}, "
                f"matvecs32={self.matvecs32}, matvecs64={self.matvecs64}, cycles={self.cycles})")


def _arnoldi_cycle(A, r0, m, inner_tol):
    """One GMRES(m) cycle on A d = r0 for a batch (nrhs, n), starting from
    d = 0.  Classical Gram-Schmidt with one reorthogonalization.  Returns d
    and the number of matvecs."""
    torch = _torch()
    nrhs, n = r0.shape
    dt = r0.dtype
    beta = torch.linalg.vector_norm(r0, dim=1)                       # (nrhs,)
    V = torch.empty((nrhs, m + 1, n), dtype=dt, device=r0.device)
    V[:, 0] = r0 / beta[:, None]
    H = np.zeros((nrhs, m + 1, m), np.complex128)
    cs = np.zeros((nrhs, m), np.complex128)
    sn = np.zeros((nrhs, m), np.complex128)
    g = np.zeros((nrhs, m + 1), np.complex128)
    g[:, 0] = beta.double().cpu().numpy()
    active = np.ones(nrhs, bool)
    steps = 0
    w = torch.empty((nrhs, n), dtype=dt, device=r0.device)
    kry = KrylovOps()
    st = _stream()
    for j in range(m):
        A.apply(V[:, j], w)
        steps += 1
        h, hnn = orthogonalize(kry, V, nrhs, j + 1, w, st)              # (nrhs, j+1)
        hcol = h.cpu().numpy()
        V[:, j + 1] = w / torch.from_numpy(np.maximum(hnn, 1e-30)).to(w.device, w.real.dtype)[:, None]
        H[:, :j + 1, j] = hcol
        H[:, j + 1, j] = hnn
        for r in range(nrhs):
            col = H[r, :, j]
            for i in range(j):
                t = cs[r, i] * col[i] + sn[r, i] * col[i + 1]
                col[i + 1] = -np.conj(sn[r, i]) * col[i] + np.conj(cs[r, i]) * col[i + 1]
                col[i] = t
            # complex Givens rotation [c s; -conj(s) c], c real, zeroing col[j+1]
            a, b = col[j], col[j + 1]
            den = np.hypot(abs(a), abs(b))
            if den == 0:
                c, s = 1.0, 0.0
            elif abs(a) == 0:
                c, s = 0.0, 1.0
            else:
                c, s = abs(a) / den, (a / abs(a)) * np.conj(b) / den
            cs[r, j], sn[r, j] = c, s
            col[j] = c * a + s * b
            col[j + 1] = 0.0
            g[r, j + 1] = -np.conj(s) * g[r, j]
            g[r, j] = c * g[r, j]
            H[r, :, j] = col
            if abs(g[r, j + 1]) <= inner_tol * g[r, 0].real:
                active[r] = False
        if not active.any():
            break
    k = j + 1
    y = np.zeros((nrhs, k), np.complex128)
    for r in range(nrhs):
        R = H[r, :k, :k]
        y[r] = _back_substitute(R, g[r, :k])
    d = torch.zeros((nrhs, n), dtype=dt, device=r0.device)
    for r in range(nrhs):
        combine(kry, V, r, k, torch.from_numpy(y[r]).to(r0.device), d[r], st)
    return d, steps


def _back_substitute(R, g):
    k = len(g)
    y = np.zeros(k, np.complex128)
    for i in range(k - 1, -1, -1):
        s = g[i] - np.dot(R[i, i + 1:k], y[i + 1:k])
        y[i] = s / R[i, i] if R[i, i] != 0 else 0.0
    return y


def gmres_ir(A32, A64, b, tol=1e-6, restart=30, inner_tol=1e-4, max_cycles=60, x0=None):
    """Solve A x = b for a batch b (nrhs, n) (complex128, GPU): FP32 GMRES(restart)
    cycles on the residual equation, FP64 residual and update after each cycle,
    until ||b - A x|| / ||b|| <= tol (measured with the FP64 operator).

    A32 / A64: LippmannSchwinger operators over ComplexF32 / ComplexF64 Green
    data of the same geometry and medium.  Returns (x, SolveInfo)."""
    torch = _torch()
    bb = b.reshape(-1, A64.n)
    nrhs = bb.shape[0]
    info = SolveInfo(nrhs)
    x = torch.zeros_like(bb) if x0 is None else x0.reshape(-1, A64.n).clone()
    bn = torch.linalg.vector_norm(bb, dim=1)
    bn = torch.where(bn > 0, bn, torch.ones_like(bn))
    if x0 is None:
        r = bb.clone()
    else:
        r = bb - A64.apply(x)
        info.matvecs64 += nrhs
    for cyc in range(max_cycles + 1):
        rn = torch.linalg.vector_norm(r, dim=1)
        rel = (rn / bn).cpu().numpy()
        info.relres = rel
        info.history.append(rel.copy())
        info.converged = rel <= tol
        if info.converged.all() or cyc == max_cycles:
            break
        act = np.nonzero(~info.converged)[0]
        ai = torch.from_numpy(act).to(bb.device)
        r32 = (r[ai] / rn[ai, None]).to(torch.complex64)
        d32, steps = _arnoldi_cycle(A32.restrict(ai), r32, restart, inner_tol)
        info.matvecs32 += steps * len(act)
        info.cycles += 1
        x[ai] += d32.to(torch.complex128) * rn[ai, None]
        r[ai] = bb[ai] - A64.restrict(ai).apply(x[ai])
        info.matvecs64 += len(act)
    return x.reshape(b.shape), info
