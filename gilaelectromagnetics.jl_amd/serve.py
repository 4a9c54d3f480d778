"""Continuous-batching server for scattering solves (I - chi G) E = E_inc.

The reference stops at the operator; its examples solve the
Lippmann-Schwinger system with IterativeSolvers (LippmannSchwinger,
examples/GilaOperators.jl:57-96; examples/scattering_guide.jl:65-68) and
cache Green data per (type, cells, scale) in preload files
(examples/GilaOperators.jl:28-55).  Here requests of different domains
arrive over time and are solved together on one GPU:

* Operators: the FP32 and FP64 Fourier Green data of every grid in use stay
  resident, keyed by (cells, scale, frqPhz), and are reused by later
  requests on the same grid.
* Slots: each grid owns a slab of request slots -- an FP32 Krylov basis
  (restart + 1 vectors), the FP64 iterate, right-hand side and residual, and
  the medium -- allocated once and reused.  Requests are admitted in arrival
  order when a slot is free or a new slab fits the memory budget; the budget
  counts every byte of a slab and of the operators (Green data and the
  library's per-RHS work buffers).
* Steps: one Arnoldi step covers all active requests of a grid in one
  batched Lippmann-Schwinger apply (gla_ls_apply, one medium per right-hand
  side).  Every grid has its own stream and at most one step or refinement
  in flight; the host waits for the grid whose work was queued first, does
  its rotations and bookkeeping and queues its next step while the other
  grids' kernels run.  A request joins at
  any round: in one batched step each request's Gram-Schmidt products run
  over its own basis columns only (its own step count), and its next basis
  vector is normalized on the device from the norm the update leaves there,
  so the only host round trip of a step is the Hessenberg column.
* Precision: GMRES(restart) cycles in FP32 on the residual equation; at the
  end of each cycle the update and the residual b - A x are formed in FP64
  (iterative refinement) and the request completes when
  ||b - A x|| / ||b|| <= tol by that FP64 residual.

A result does not depend on what else is in the batch beyond rounding: the
matvec of each right-hand side is computed by the same kernels whatever the
batch, and the Gram-Schmidt, rotations and updates of a request use only its
own slot.
"""
import contextlib
import ctypes
import time
from fractions import Fraction

import numpy as np

from . import _capi as capi
from .solve import DGKS, KrylovOps, combine, gs_pass


def _torch():
    import torch
    return torch


class SolveRequest:
    """One scattering solve: chi (c1,c2,c3) and E_inc (c1,c2,c3,3), complex,
    column-major, on a grid of `cells` cubic cells of side `scale`."""

    _next_id = 0

    def __init__(self, cells, chi, e_inc, tol=1e-6, scale=Fraction(1, 32), frq=1.0 + 0j, rid=None):
        self.cells = tuple(int(c) for c in cells)
        chi = np.asarray(chi)
        e_inc = np.asarray(e_inc)
        if chi.shape != self.cells or e_inc.shape != self.cells + (3,):
            raise ValueError("chi must be (c1,c2,c3) and e_inc (c1,c2,c3,3)")
        self.chi = chi
        self.e_inc = e_inc
        self.tol = float(tol)
        self.scale = Fraction(scale)
        self.frq = complex(frq)
        if rid is None:
            rid = SolveRequest._next_id
            SolveRequest._next_id += 1
        self.rid = rid
        self.t_arrive = None
        self.t_admit = None
        self.t_done = None
        self.result = None          # FP64 field (c1,c2,c3,3), numpy
        self.relres = None
        self.converged = False
        self.cycles = 0
        self.steps = 0

    @property
    def key(self):
        return (self.cells, self.scale, self.frq)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrive


class DeviceOps:
    """FP32 and FP64 Lippmann-Schwinger applies of one grid through libgila_amd."""

    def __init__(self, cells, scale, frq, device=0):
        from .gla import GlaKerOpt, GlaOprMem, GlaVol
        vol = GlaVol(cells, (scale,) * 3)
        self.cells = tuple(cells)
        self.ncell = int(np.prod(cells))
        self.n = 3 * self.ncell
        self.mem32 = GlaOprMem(GlaKerOpt(frq, devMod=True, device=device), vol, setTyp=np.complex64)
        self.mem64 = GlaOprMem(GlaKerOpt(frq, devMod=True, device=device), vol, setTyp=np.complex128)
        L = capi.lib()
        self._ws32 = L.gla_opr_workspace_bytes(self.mem32._h, 1)
        self._ws64 = L.gla_opr_workspace_bytes(self.mem64._h, 1)
        # device G^ bytes of both operators: 8 branches x U1 U2 U3 x 6 components
        U = [int(u) for u in self.mem64._dims.fur_ext]
        self.fixed_bytes = 8 * 6 * int(np.prod(U)) * (8 + 16)

    def ws_bytes(self, nrhs32, nrhs64):
        return self._ws32 * nrhs32 + self._ws64 * nrhs64

    def _apply(self, mem, x, chi, y, stream):
        k = x.shape[0]
        capi.check(capi.lib().gla_ls_apply(mem._h, ctypes.c_void_p(chi.data_ptr()), chi.numel(),
                                           self.ncell, ctypes.c_void_p(x.data_ptr()), x.numel(),
                                           ctypes.c_void_p(y.data_ptr()), y.numel(), k,
                                           ctypes.c_void_p(stream)))

    def apply32(self, x, chi, y, stream):
        self._apply(self.mem32, x, chi, y, stream)

    def apply64(self, x, chi, y, stream):
        self._apply(self.mem64, x, chi, y, stream)


class _Grid:
    """Slab of request slots on one grid; slots 0..k-1 are active."""

    def __init__(self, ops, cap, restart, inner, device, gs="cgs2", depth=1):
        torch = _torch()
        self.ops = ops
        self.cap = cap
        self.m = restart
        self.n = ops.n
        self.dev = device
        self.inner = inner
        self.gs = gs
        cdt = torch.complex64 if inner == "fp32" else torch.complex128
        self.cdt = cdt
        n, m = self.n, restart
        self.V = torch.zeros((cap, m + 1, n), dtype=cdt, device=device)
        self.W = torch.empty((cap, n), dtype=cdt, device=device)
        self.Xin = torch.empty((cap, n), dtype=cdt, device=device)
        self.chi_in = torch.zeros((cap, ops.ncell), dtype=cdt, device=device)
        self.chi64 = torch.zeros((cap, ops.ncell), dtype=torch.complex128, device=device)
        self.x = torch.zeros((cap, n), dtype=torch.complex128, device=device)
        self.b = torch.zeros((cap, n), dtype=torch.complex128, device=device)
        self.r = torch.zeros((cap, n), dtype=torch.complex128, device=device)
        self.t64 = torch.empty((cap, n), dtype=torch.complex128, device=device)
        self.H = np.zeros((cap, m + 1, m), np.complex128)
        self.cs = np.zeros((cap, m))
        self.sn = np.zeros((cap, m), np.complex128)
        self.g = np.zeros((cap, m + 1), np.complex128)
        self.j = np.zeros(cap, np.int64)
        self.rn = np.zeros(cap)
        self.bn = np.ones(cap)
        self.reqs = []
        self.stream = torch.cuda.Stream(device=device) if str(device).startswith("cuda") else None
        self.kry = KrylovOps()
        self.last_used = time.perf_counter()
        self.pending = None
        self.reorth = None
        self.state = None           # None, "step", "reorth" or "refine": what is in flight
        self.seq = 0
        self.event = torch.cuda.Event() if self.stream is not None else None
        # device-resident step (one or two classical passes): the next basis
        # vector is normalized on the device from the norm the update kernel
        # leaves there, and also written to Xin, the next step's matvec
        # input; each slot's products run over its own columns 0..j only
        # (stale columns past j are never read, so a restart need not clear
        # them).  Xin[s] always holds slot s's current vector.  Up to `depth`
        # steps are in flight: step j+1 is queued before the host has done
        # step j's rotations, so the GPU does not wait for the host.  A slot
        # whose cycle ends at step j is masked in the steps queued after it
        # (their work for it is discarded; they only write its columns past
        # j + 1 and its staged vector, which the restart rewrites).
        self.fused = gs != "dgks"
        self.depth = depth if self.fused else 1
        self.inflight = []          # queued steps, oldest first
        self.masked = set()         # slots whose results in queued steps are discarded
        self.ended = []             # slots whose cycle ended, waiting for the queue to drain
        self.nlaunch = 0
        if self.fused:
            pin = self.V.is_cuda
            self.jh = [torch.empty((2, cap), dtype=torch.int64, pin_memory=pin) for _ in range(2)]
            self.jd = [torch.empty((2, cap), dtype=torch.int64, device=device) for _ in range(2)]
            self.events = [torch.cuda.Event() if self.stream is not None else None for _ in range(2)]
            self.gsout = torch.empty(cap * (2 * (m + 1) + 2), dtype=torch.float64, device=device)
            self.gsout2 = torch.empty_like(self.gsout) if gs == "cgs2" else None

    @staticmethod
    def slot_bytes(n, ncell, restart, inner):
        s = 8 if inner == "fp32" else 16
        return (restart + 3) * n * s + ncell * (s + 16) + 4 * n * 16

    @property
    def k(self):
        return len(self.reqs)

    def ctx(self):
        torch = _torch()
        return torch.cuda.stream(self.stream) if self.stream is not None else contextlib.nullcontext()

    def stream_handle(self):
        torch = _torch()
        if self.stream is not None:
            return self.stream.cuda_stream
        return torch.cuda.current_stream().cuda_stream if torch.cuda.is_available() else 0

    def apply_inner(self, x, chi, y):
        (self.ops.apply32 if self.inner == "fp32" else self.ops.apply64)(x, chi, y, self.stream_handle())

    def apply64(self, x, chi, y):
        self.ops.apply64(x, chi, y, self.stream_handle())

    # --- slot management ---------------------------------------------------
    def start_cycle(self, s):
        """Arnoldi restart of slot s from its residual r (norm rn[s])."""
        if self.fused:
            self.V[s, 0].copy_((self.r[s] / self.rn[s]).to(self.cdt))
            self.Xin[s].copy_(self.V[s, 0])
        else:
            self.V[s, 1:].zero_()
            self.V[s, 0].copy_((self.r[s] / self.rn[s]).to(self.cdt))
        self.H[s] = 0
        self.g[s] = 0
        self.g[s, 0] = 1.0
        self.j[s] = 0

    def admit(self, req, now):
        torch = _torch()
        s = self.k
        self.reqs.append(req)
        req.t_admit = now
        chi = torch.from_numpy(np.ascontiguousarray(req.chi.reshape(-1, order="F")).astype(np.complex128))
        b = torch.from_numpy(np.ascontiguousarray(req.e_inc.reshape(-1, order="F")).astype(np.complex128))
        bn = float(np.linalg.norm(req.e_inc))
        with self.ctx():
            self.chi64[s].copy_(chi.to(self.dev, non_blocking=False))
            self.chi_in[s].copy_(self.chi64[s].to(self.cdt))
            self.b[s].copy_(b.to(self.dev, non_blocking=False))
            self.x[s].zero_()
            self.r[s].copy_(self.b[s])
        self.bn[s] = bn if bn > 0 else 1.0
        self.rn[s] = bn
        if bn == 0:
            return False
        with self.ctx():
            self.start_cycle(s)
        return True

    def remove(self, s):
        """Drop slot s, moving the last active slot into it."""
        last = self.k - 1
        if s != last:
            with self.ctx():
                for t in (self.V, self.chi_in, self.chi64, self.x, self.b, self.r) + ((self.Xin,) if self.fused else ()):
                    t[s].copy_(t[last])
            for a in (self.H, self.cs, self.sn, self.g, self.j, self.rn, self.bn):
                a[s] = a[last]
            self.reqs[s] = self.reqs[last]
        self.reqs.pop()

    # --- one Arnoldi step for all active slots -------------------------------
    def launch_step(self):
        """Gather each slot's current basis vector, apply A to the batch and
        run classical Gram-Schmidt over the slab's columns 0..jm (zero beyond
        each slot's own step); the coefficients and the norms before and
        after are copied to the host."""
        torch = _torch()
        k = self.k
        jm = int(self.j[:k].max())
        st = self.stream_handle()
        if self.fused:
            self._launch_step_fused(k, jm, st)
            return
        with self.ctx():
            jt = torch.from_numpy(self.j[:k].copy()).to(self.dev)
            ar = torch.arange(k, device=self.dev)
            if self.V.is_cuda:
                self.kry.rows(self.V, k, jt, self.Xin, st)
            else:
                self.Xin[:k] = self.V[ar, jt]
            self.apply_inner(self.Xin[:k], self.chi_in[:k], self.W[:k])
            h, n0, n1 = gs_pass(self.kry, self.V, k, jm + 1, self.W, st)
            if self.gs == "cgs2":
                h2, _, n1 = gs_pass(self.kry, self.V, k, jm + 1, self.W, st)
                h = h + h2
            host = [t.to("cpu", non_blocking=True) for t in (h, n0, n1)]
            self.pending = (host, jt, ar, jm, k)
            self.record()

    def can_run_ahead(self):
        """Whether another step may be queued behind the ones in flight:
        nothing ended yet, and every slot's next column stays in the slab
        (a slot reaching the restart length ends its cycle)."""
        f = len(self.inflight)
        return (self.fused and 0 < f < self.depth and not self.ended and self.k > 0
                and int((self.j[:self.k] + f).max()) <= self.m - 1)

    def _launch_step_fused(self, k, jm, st):
        """launch_step with the basis products over each slot's own columns
        and the next vector formed on the device (no gather, no host norm);
        a step queued behind others uses the step indices they will reach."""
        f = len(self.inflight)
        js = self.j[:k] + f
        jm = int(js.max())
        nv = jm + 1
        b = self.nlaunch % 2          # the buffers of the step two launches back are free
        self.nlaunch += 1
        jh = self.jh[b].numpy()
        jh[0, :k] = js
        jh[1, :k] = js + 1
        with self.ctx():
            jd = self.jd[b]
            jd.copy_(self.jh[b], non_blocking=True)
            ncol = jd[1, :k]
            self.apply_inner(self.Xin[:k], self.chi_in[:k], self.W[:k])
            h, n0, n1 = gs_pass(self.kry, self.V, k, nv, self.W, st, ncol=ncol, out=self.gsout)
            if self.gs == "cgs2":
                h2, _, n1b = gs_pass(self.kry, self.V, k, nv, self.W, st, ncol=ncol, out=self.gsout2)
                h.add_(h2)
                n1.copy_(n1b)
            self.kry.next(self.V, k, ncol, self.W, n1, st, stage=self.Xin)
            src = self.gsout[:k * (2 * nv + 2)]
            host = src.to("cpu", non_blocking=True) if src.is_cuda else src.clone()
            ev = self.events[b]
            if ev is not None:
                ev.record(self.stream)
        self.inflight.append({"host": host, "jm": jm, "k": k, "event": ev})

    def launch_reorth(self):
        """The second Gram-Schmidt pass when gs is "dgks" and a slot lost more
        than 1 - 1/sqrt(2) of its norm (cgs2 ran it in launch_step already).
        Returns whether one was launched."""
        self.reorth = None
        if self.gs != "dgks":
            return False
        (h, n0, n1), jt, ar, jm, k = self.pending
        if not (np.sqrt(n1.numpy()) < DGKS * np.sqrt(n0.numpy())).any():
            return False
        with self.ctx():
            h2, _, n2 = gs_pass(self.kry, self.V, k, jm + 1, self.W, self.stream_handle())
            self.reorth = (h2.to("cpu", non_blocking=True), n2.to("cpu", non_blocking=True))
            self.record()
        return True

    def finish_step(self):
        """Host part of the step: the next basis vector, Givens rotations and
        residual estimates.  Returns the slots whose cycle ends."""
        torch = _torch()
        if self.fused:
            item = self.inflight.pop(0)
            k, jm = item["k"], item["jm"]
            arr = item["host"].numpy()
            nv = jm + 1
            hcol = arr[:2 * k * nv].view(np.complex128).reshape(k, nv)
            hn = np.sqrt(arr[2 * k * nv + k:2 * k * nv + 2 * k])
            return self._advance(hcol, hn, k, jm, mask=self.masked)
        (h, n0, n1), jt, ar, jm, k = self.pending
        self.pending = None
        hcol = h.numpy()
        hn = np.sqrt(n1.numpy())
        if self.reorth is not None:
            h2, n2 = self.reorth
            self.reorth = None
            hcol = hcol + h2.numpy()
            hn = np.sqrt(n2.numpy())
        with self.ctx():
            scale = torch.from_numpy(np.maximum(hn, 1e-30)).to(self.dev)
            if self.V.is_cuda:
                self.kry.rows(self.V, k, jt + 1, self.W, self.stream_handle(), scale=scale)
            else:
                self.V[ar, jt + 1] = self.W[:k] / scale.to(self.W.real.dtype).unsqueeze(1)
        return self._advance(hcol, hn, k, jm)

    def _advance(self, hcol, hn, k, jm, mask=()):
        """Givens update with the step's coefficients (the library's host
        routine gla_krylov_givens, one code path whatever the batch), then
        the slots whose cycle ends.  Slots in `mask` keep their state (their
        cycle ended at an earlier step)."""
        m = self.m
        js = np.ascontiguousarray(self.j[:k], dtype=np.int64)
        hcol = np.ascontiguousarray(hcol, dtype=np.complex128)
        hn = np.ascontiguousarray(hn, dtype=np.float64)
        keep = {s: (self.cs[s].copy(), self.sn[s].copy(), self.H[s].copy(), self.g[s].copy()) for s in mask}
        ptr = lambda a: ctypes.c_void_p(a.ctypes.data)
        capi.check(capi.lib().gla_krylov_givens(k, m, ptr(js), ptr(hcol), hcol.shape[1], ptr(hn), ptr(self.cs),
                                                ptr(self.sn), ptr(self.H), ptr(self.g)))
        for s, (a, b, c, d) in keep.items():
            self.cs[s], self.sn[s], self.H[s], self.g[s] = a, b, c, d
        ar_h = np.arange(k)
        res = np.abs(self.g[ar_h, js + 1])
        est = res * self.rn[:k] / self.bn[:k]
        ended = []
        for s in range(k):
            if s in mask:
                continue
            self.j[s] = js[s] + 1
            req = self.reqs[s]
            req.steps += 1
            if js[s] + 1 == m or res[s] <= self.inner_tol or est[s] <= 0.5 * req.tol:
                ended.append(s)
        return ended

    def launch_refine(self, slots):
        """FP64 update x += V y |r| and residual r = b - A x for the slots whose
        cycle ended; residual norms go to the host."""
        torch = _torch()
        with self.ctx():
            for s in slots:
                kk = int(self.j[s])
                R = self.H[s, :kk, :kk]
                y = np.linalg.solve(np.triu(R), self.g[s, :kk]) if kk else np.zeros(0, np.complex128)
                d = self.Xin[s]
                d.zero_()
                if kk:
                    combine(self.kry, self.V, s, kk, torch.from_numpy(y).to(self.dev), d, self.stream_handle())
                self.x[s] += d.to(torch.complex128) * self.rn[s]
                self.reqs[s].cycles += 1
            idx = torch.tensor(slots, device=self.dev)
            nb = len(slots)
            xin = self.x[idx]
            self.apply64(xin, self.chi64[idx].contiguous(), self.t64[:nb])
            self.r[idx] = self.b[idx] - self.t64[:nb]
            rn = torch.linalg.vector_norm(self.r[idx], dim=1)
            self.pending_refine = (slots, rn.to("cpu", non_blocking=True))
            self.record()

    def record(self):
        """Mark the end of the work just queued on this grid's stream."""
        if self.stream is not None:
            self.event.record(self.stream)

    def oldest_seq(self):
        return self.inflight[0]["seq"] if self.inflight else self.seq

    def wait_oldest(self):
        """Block until the oldest queued step (or the refinement) is done."""
        if self.inflight:
            ev = self.inflight[0]["event"]
            if ev is not None:
                ev.synchronize()
        else:
            self.wait()

    def wait(self):
        """Block until this grid's queued work is done (other streams run on)."""
        if self.stream is not None:
            self.event.synchronize()


class Server:
    """Continuous-batching Lippmann-Schwinger solver (module docstring).

    inner: "fp32" (FP32 cycles, FP64 refinement) or "fp64" (plain FP64
    restarted GMRES).  gs: Gram-Schmidt of the Arnoldi step, "cgs1" (one
    classical pass), "cgs2" (two) or "dgks" (a second pass when the first
    removes more than 1 - 1/sqrt(2) of the norm); default cgs1 for fp32
    cycles and cgs2 for fp64.  max_batch: slots per grid.  max_active: requests in
    flight over all grids (None: no limit; 1 gives one solve after another).
    mem_budget: bytes of device memory the server may hold (default: 90% of
    what is free at start).  depth: Arnoldi steps queued per grid ahead of
    the host (2: step j+1 runs while the host rotates step j; measured on
    the bench's served stream: p50 latency 1.7-2.1 s -> 1.4-1.5 s, but ~3%
    fewer solves/s at saturation, the steps computed ahead for requests
    whose cycle had just ended being discarded)."""

    def __init__(self, restart=30, inner="fp32", inner_tol=1e-4, max_batch=16, mem_budget=None, max_cycles=200,
                 device="cuda:0", ops_factory=None, max_active=None, gs=None, depth=1):
        torch = _torch()
        self.restart = restart
        self.inner = inner
        self.inner_tol = inner_tol
        self.max_batch = max_batch
        self.max_cycles = max_cycles
        self.max_active = max_active
        # Arnoldi steps in flight per grid (1: the host finishes a step
        # before the next is queued)
        self.depth = int(depth)
        if self.depth < 1:
            raise ValueError("depth must be at least 1")
        # Gram-Schmidt of the Arnoldi step: one pass in FP32 cycles (their
        # result is corrected by the FP64 residual and update), two in plain
        # FP64 solves
        self.gs = gs or ("cgs1" if inner == "fp32" else "cgs2")
        if self.gs not in ("cgs1", "cgs2", "dgks"):
            raise ValueError("gs must be cgs1, cgs2 or dgks")
        self.device = device
        if mem_budget is None:
            free = torch.cuda.mem_get_info(torch.device(device))[0] if str(device).startswith("cuda") else 1 << 36
            mem_budget = int(0.9 * free)
        self.mem_budget = int(mem_budget)
        self.ops_factory = ops_factory or (lambda cells, scale, frq: DeviceOps(cells, scale, frq))
        self.ops = {}          # key -> ops (resident Green data)
        self.grids = {}        # key -> _Grid
        self.waiting = []
        self.done = []
        self.used = 0
        self.rounds = 0
        self.reorth_rounds = 0
        self.ahead_steps = 0        # steps queued behind another step of the same grid
        self.discarded_steps = 0    # slot-steps computed ahead for a slot whose cycle had ended
        self._seq = 0
        # host seconds per phase of the rounds, and slot-steps run
        self.times = dict.fromkeys(("launch", "wait_step", "reorth", "finish", "refine_launch", "wait_refine",
                                    "complete", "slot_rounds"), 0.0)

    # --- memory accounting --------------------------------------------------
    def _grid_bytes(self, ops, cap):
        slot = _Grid.slot_bytes(ops.n, ops.ncell, self.restart, self.inner)
        return cap * (slot + ops.ws_bytes(1 if self.inner == "fp32" else 0, 1))

    def _evict_idle(self, need):
        """Free idle slabs, then idle operators, oldest first, until `need`
        bytes fit."""
        for key, g in sorted(self.grids.items(), key=lambda kv: kv[1].last_used):
            if self.used + need <= self.mem_budget:
                return
            if g.k == 0 and g.state is None:
                self.used -= g.bytes
                del self.grids[key]
        busy = set(self.grids)
        for key in list(self.ops):
            if self.used + need <= self.mem_budget:
                return
            if key not in busy and not any(r.key == key for r in self.waiting):
                self.used -= self.ops[key].fixed_bytes
                del self.ops[key]

    def _grid_for(self, req, allow_new=True):
        """The request's grid with a free slot, creating the operators and the
        slab if the budget allows; None if the request must wait."""
        key = req.key
        g = self.grids.get(key)
        if g is not None:
            return g if (g.k < g.cap and g.state is None) else None
        if not allow_new:
            return None
        ops = self.ops.get(key)
        if ops is None:
            ops = self.ops_factory(req.cells, req.scale, req.frq)
            self.ops[key] = ops
            self.used += ops.fixed_bytes
        per = self._grid_bytes(ops, 1)
        self._evict_idle(per)
        cap = int(min(self.max_batch, max(0, (self.mem_budget - self.used) // per)))
        if cap < 1:
            return None
        g = _Grid(ops, cap, self.restart, self.inner, self.device, self.gs, self.depth)
        g.inner_tol = self.inner_tol
        g.bytes = self._grid_bytes(ops, cap)
        self.used += g.bytes
        self.grids[key] = g
        return g

    # --- request flow -------------------------------------------------------
    def submit(self, req, now=None):
        req.t_arrive = time.perf_counter() if now is None else now
        self.waiting.append(req)

    def _complete(self, g, s, now, converged):
        torch = _torch()
        req = g.reqs[s]
        with g.ctx():
            xh = g.x[s].to("cpu", copy=True)
        if g.stream is not None:
            g.stream.synchronize()
        req.result = xh.numpy().reshape(req.cells + (3,), order="F")
        req.relres = float(g.rn[s] / g.bn[s])
        req.converged = converged
        req.t_done = now
        self.done.append(req)
        g.remove(s)
        if getattr(self, "on_done", None) is not None:
            self.on_done(req)
        del torch

    def _admit(self, now):
        """Admit waiting requests in arrival order.  A request whose grid is
        full waits without holding back requests for other grids; once a
        request cannot get a new slab, later ones may only take free slots of
        existing grids (no overtaking for memory)."""
        allow_new = True
        keep = []
        for req in self.waiting:
            if self.max_active is not None and self.active >= self.max_active:
                keep.append(req)
                continue
            g = self._grid_for(req, allow_new)
            if g is None:
                if req.key not in self.grids:
                    allow_new = False
                keep.append(req)
                continue
            if not g.admit(req, now):
                req.relres = 0.0
                self._complete(g, g.k - 1, time.perf_counter(), True)
        self.waiting = keep

    @property
    def active(self):
        return sum(g.k for g in self.grids.values())

    def round(self):
        """One scheduler iteration: admit what fits, queue Arnoldi steps on
        the grids with requests (each grid on its own stream; up to the
        grid's depth of steps in flight, none ahead on a grid that requests
        are waiting for, so it drains and admits them), then wait for the
        oldest queued work of any grid and do its host part -- the other
        grids' kernels, and this grid's next step, keep the GPU busy
        meanwhile.  Returns the number of grids processed (0 when nothing is
        in flight)."""
        torch = _torch()
        tm = self.times
        t0 = time.perf_counter()
        self._admit(t0)
        wanted = {r.key for r in self.waiting}
        for key, g in self.grids.items():
            if g.state is None and g.k > 0:
                g.last_used = t0
                g.state = "step"
                self._launch(g)
            while g.state == "step" and key not in wanted and g.can_run_ahead():
                self._launch(g)
                self.ahead_steps += 1
        busy = [g for g in self.grids.values() if g.state is not None]
        if not busy:
            return 0
        g = min(busy, key=lambda x: x.oldest_seq())
        t1 = time.perf_counter()
        tm["launch"] += t1 - t0
        g.wait_oldest()
        t2 = time.perf_counter()
        tm["wait_step" if g.state != "refine" else "wait_refine"] += t2 - t1
        if g.state == "step" and not g.fused and g.launch_reorth():
            self.reorth_rounds += 1
            g.state = "reorth"
            self._seq += 1
            g.seq = self._seq
        elif g.state in ("step", "reorth"):
            ends = g.finish_step()
            g.ended += ends
            if g.inflight:
                g.masked |= set(ends)      # the queued steps' work for them is discarded
                self.discarded_steps += len(ends) * len(g.inflight)
            elif g.ended:
                g.launch_refine(sorted(g.ended))
                g.ended, g.masked = [], set()
                g.state = "refine"
                self._seq += 1
                g.seq = self._seq
            else:
                g.state = None
            tm["finish"] += time.perf_counter() - t2
        else:
            slots, rn = g.pending_refine
            rn = rn.numpy()
            finished = []
            for s, v in zip(slots, rn):
                g.rn[s] = float(v)
                req = g.reqs[s]
                if g.rn[s] / g.bn[s] <= req.tol:
                    finished.append((s, True))
                elif req.cycles >= self.max_cycles:
                    finished.append((s, False))
                else:
                    with g.ctx():
                        g.start_cycle(s)
            g.state = None
            # remove from the highest slot down so compaction keeps indices valid
            now = time.perf_counter()
            for s, conv in sorted(finished, reverse=True):
                self._complete(g, s, now, conv)
            tm["complete"] += time.perf_counter() - t2
        return 1

    def _launch(self, g):
        """Queue one Arnoldi step on grid g."""
        g.launch_step()
        self._seq += 1
        g.seq = self._seq
        if g.fused:
            g.inflight[-1]["seq"] = self._seq
        self.times["slot_rounds"] += g.k
        self.rounds += 1

    def run(self, requests, arrivals=None, timeout=None, on_done=None):
        """Serve `requests`; arrivals[i] (seconds from the start) is when
        request i is submitted (all at once if None); on_done(req) is called
        as each request completes.  Returns the completed requests in
        completion order."""
        self.on_done = on_done
        t0 = time.perf_counter()
        n0 = len(self.done)        # a server may run several streams in turn
        order = list(range(len(requests)))
        if arrivals is not None:
            order.sort(key=lambda i: arrivals[i])
        nxt = 0
        while len(self.done) - n0 < len(requests):
            now = time.perf_counter() - t0
            while nxt < len(order) and (arrivals is None or arrivals[order[nxt]] <= now):
                i = order[nxt]
                self.submit(requests[i], t0 + (0.0 if arrivals is None else arrivals[i]))
                nxt += 1
            if self.active == 0 and not self.waiting:
                if nxt < len(order):
                    time.sleep(max(0.0, arrivals[order[nxt]] - (time.perf_counter() - t0)))
                continue
            if self.round() == 0 and self.waiting and self.active == 0:
                raise RuntimeError("a request does not fit the memory budget")
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError("serving did not finish in time")
        return self.done[n0:]


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs])
    span = max(r.t_done for r in reqs) - min(r.t_arrive for r in reqs)
    return {"n": len(reqs), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "solves_per_s": len(reqs) / span if span > 0 else float("inf"), "span_s": span}


def random_request(seed, cells, scale=Fraction(1, 32), tol=1e-6):
    """A served-stream request: chi = 0 outside a random axis-aligned box
    covering 10-40% of the domain and Re chi ~ U[0.5, 3], Im chi ~ U[0, 0.1]
    per cell inside; E_inc a unit plane wave with a random direction and a
    random polarisation perpendicular to it."""
    from .inputs import plane_wave, random_unit, uniform
    u = uniform(seed, 16)
    frac = 0.1 + 0.3 * u[0]
    side = [max(1, min(c, int(round(c * frac ** (1.0 / 3.0))))) for c in cells]
    lo = [int(u[1 + d] * (c - s + 1)) for d, (c, s) in enumerate(zip(cells, side))]
    chi = np.zeros(cells, np.complex128, order="F")
    box = tuple(slice(a, a + s) for a, s in zip(lo, side))
    nbox = int(np.prod(side))
    uv = uniform(seed + 1, 2 * nbox)
    chi[box] = (0.5 + 2.5 * uv[:nbox] + 0.1j * uv[nbox:]).reshape(side, order="F")
    khat = random_unit(seed + 2)
    p = random_unit(seed + 3)
    p = p - khat * (p @ khat)
    p /= np.linalg.norm(p)
    e = plane_wave(cells, (scale,) * 3, khat, p)
    return SolveRequest(cells, chi, e, tol=tol, scale=scale)
