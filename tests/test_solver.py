"""Lippmann-Schwinger operator and the mixed-precision GMRES solver.

The operator of examples/GilaOperators.jl:57-96 (A x = x - chi .* G x) is
checked on the GPU against the oracle's matvec; the solver's host logic is
checked on the CPU with operators backed by the oracle, and on the GPU by
the residual of the FP64 oracle operator at the returned solution."""
import numpy as np
import pytest

H = 1.0 / 32


def _sphere_medium(N, chi, radius_frac=0.4):
    c = [(np.arange(n) + 0.5) / n - 0.5 for n in N]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    inside = X ** 2 + Y ** 2 + Z ** 2 <= radius_frac ** 2
    m = np.zeros(N, np.complex128, order="F")
    m[inside] = chi
    return m


class OracleLS:
    """A x = x - chi .* (G x) with the oracle's matvec; the interface the solver
    uses (apply, restrict, n, nbatch)."""

    def __init__(self, orc, N, fur, media, dtype):
        self.orc, self.N, self.fur, self.dtype = orc, N, fur, dtype
        self.media = [np.asarray(m, dtype) for m in media]
        self.nbatch = len(self.media)
        self.n = 3 * int(np.prod(N))

    def apply(self, x, y=None):
        import torch
        xb = x.reshape(-1, self.n)
        out = np.empty(tuple(xb.shape), self.dtype)
        for r in range(xb.shape[0]):
            v = xb[r].cpu().numpy().astype(self.dtype).reshape(self.N + (3,), order="F")
            gx = self.orc.matvec(self.N, self.N, (1, 1, 1), self.fur, v, dtype=self.dtype)
            chi = self.media[0 if self.nbatch == 1 else r]
            out[r] = (v - chi[..., None] * gx).reshape(-1, order="F")
        res = torch.from_numpy(out)
        if y is None:
            return res.reshape(x.shape)
        y.copy_(res.reshape(y.shape))
        return y

    def restrict(self, idx):
        if self.nbatch == 1:
            return self
        return OracleLS(self.orc, self.N, self.fur, [self.media[int(i)] for i in idx], self.dtype)


def _oracle_residual(orc, N, fur, chi, x, b):
    xa = x.reshape(N + (3,), order="F")
    ax = xa - chi[..., None] * orc.matvec(N, N, (1, 1, 1), fur, xa)
    return np.linalg.norm(ax.reshape(-1, order="F") - b) / np.linalg.norm(b)


def test_gmres_ir_host_logic(gila, orc):
    import torch
    N = (6, 6, 6)
    fur, _ = orc.self_operator_fur(N, model=0)
    # a scaled Green function so that chi G is a sizeable perturbation of I
    fur = fur * 40.0
    chi = _sphere_medium(N, 2.0 + 0.5j)
    A32 = OracleLS(orc, N, fur, [chi], np.complex64)
    A64 = OracleLS(orc, N, fur, [chi], np.complex128)
    b = torch.from_numpy(orc.complex_gaussian((2, A64.n), seed=21))
    x, info = gila.solve.gmres_ir(A32, A64, b, tol=1e-11, restart=30, inner_tol=1e-4)
    assert info.converged.all(), info
    for r in range(2):
        assert _oracle_residual(orc, N, fur, chi, x[r].numpy(), b[r].numpy()) <= 1e-11
    # FP32 cycles stall near FP32 precision; the FP64 refinement carries on
    assert info.cycles >= 2 and info.matvecs64 == sum(int((h > 1e-11).sum()) for h in info.history[:-1])


def test_gmres_ir_per_rhs_media_and_subsets(gila, orc):
    import torch
    N = (4, 6, 4)
    fur, _ = orc.self_operator_fur(N, model=0)
    fur = fur * 40.0
    media = [_sphere_medium(N, c) for c in (0.5 + 0.1j, 3.0 + 1.0j, 0j)]
    A32 = OracleLS(orc, N, fur, media, np.complex64)
    A64 = OracleLS(orc, N, fur, media, np.complex128)
    b = torch.from_numpy(orc.complex_gaussian((3, A64.n), seed=22))
    x, info = gila.solve.gmres_ir(A32, A64, b, tol=1e-10)
    assert info.converged.all()
    # chi = 0: A = I, so x = b after the first cycle
    assert np.allclose(x[2].numpy(), b[2].numpy(), rtol=0, atol=1e-10 * np.abs(b[2].numpy()).max())
    for r in range(3):
        assert _oracle_residual(orc, N, fur, media[r], x[r].numpy(), b[r].numpy()) <= 1e-10
        # alone and in the batch: same answer to the solve tolerance
        xs, _ = gila.solve.gmres_ir(A32.restrict([r]), A64.restrict([r]), b[r:r + 1], tol=1e-10)
        assert np.linalg.norm(xs[0].numpy() - x[r].numpy()) <= 1e-8 * np.linalg.norm(x[r].numpy())


def test_gmres_zero_rhs_and_max_cycles(gila, orc):
    import torch
    N = (4, 4, 4)
    fur, _ = orc.self_operator_fur(N, model=0)
    chi = _sphere_medium(N, 1.0)
    A32 = OracleLS(orc, N, fur * 40.0, [chi], np.complex64)
    A64 = OracleLS(orc, N, fur * 40.0, [chi], np.complex128)
    x, info = gila.solve.gmres_ir(A32, A64, torch.zeros((1, A64.n), dtype=torch.complex128))
    assert info.converged.all() and info.cycles == 0 and not x.abs().max()
    b = torch.from_numpy(orc.complex_gaussian((1, A64.n), seed=3))
    x, info = gila.solve.gmres_ir(A32, A64, b, tol=1e-300, restart=2, max_cycles=3)
    assert info.cycles == 3 and not info.converged.any() and len(info.history) == 4


def _ls_setup(gila, orc, N, typ, model=0):
    fur, _ = orc.self_operator_fur(N, model=model)
    U = orc.dims(N, N, (1, 1, 1))[1]
    vol = gila.GlaVol(N, ((1, 32),) * 3)
    mem = gila.GlaOprMem(gila.GlaKerOpt(True), vol, egoFur=orc.fur_branches(fur.astype(typ), U), setTyp=typ)
    return gila.GlaOpr(mem), fur


@pytest.mark.gpu
@pytest.mark.parametrize("typ,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_ls_apply_matches_oracle(gila, orc, gpu, typ, tol):
    import torch
    N = (16, 12, 8)
    print(f"case ls_apply N={N} {np.dtype(typ).name}", flush=True)
    G, fur = _ls_setup(gila, orc, N, typ)
    media = [_sphere_medium(N, 2.0 + 0.5j), _sphere_medium(N, -1.0 + 3.0j, 0.3)]
    x = orc.complex_gaussian((2,) + N + (3,), seed=31)
    xd = torch.from_numpy(np.stack([x[r].reshape(-1, order="F") for r in range(2)]).astype(typ)).cuda()
    shared = gila.solve.LippmannSchwinger(G, media[0].astype(typ))
    batched = gila.solve.LippmannSchwinger(G, np.stack(media, axis=3).astype(typ))
    ys = shared.apply(xd).cpu().numpy()
    yb = batched.apply(xd).cpu().numpy()
    for r in range(2):
        gx = orc.matvec(N, N, (1, 1, 1), fur, x[r])
        for y, chi in ((ys[r], media[0]), (yb[r], media[r])):
            ref = x[r] - chi[..., None] * gx
            assert orc.rel_l2(y.reshape(N + (3,), order="F"), ref) <= tol
    with pytest.raises(gila.DimensionMismatch):
        gila.solve.LippmannSchwinger(G, media[0][:, :, :4].astype(typ))
    with pytest.raises(gila.ArgumentError):
        gila.solve.LippmannSchwinger(G, media[0].astype(np.complex128 if typ == np.complex64 else np.complex64))
    with pytest.raises(gila.DimensionMismatch):
        batched.apply(xd[:1])


@pytest.mark.gpu
def test_gmres_ir_on_gpu(gila, orc, gpu):
    import torch
    N = (16, 16, 16)
    print(f"case gmres_ir N={N}", flush=True)
    G32, fur = _ls_setup(gila, orc, N, np.complex64, model=2)
    G64, _ = _ls_setup(gila, orc, N, np.complex128, model=2)
    media = [_sphere_medium(N, c) for c in (3.0 + 0.2j, 10.0 + 1.0j)]
    m = np.stack(media, axis=3)
    A32 = gila.solve.LippmannSchwinger(G32, m.astype(np.complex64))
    A64 = gila.solve.LippmannSchwinger(G64, m)
    b = torch.from_numpy(orc.complex_gaussian((2, A64.n), seed=41)).cuda()
    x, info = gila.solve.gmres_ir(A32, A64, b, tol=1e-10)
    assert info.converged.all(), info
    for r in range(2):
        assert _oracle_residual(orc, N, fur, media[r], x[r].cpu().numpy(), b[r].cpu().numpy()) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("dt,tol", [("complex64", 1e-6), ("complex128", 1e-13)])
def test_krylov_kernels_match_host(gila, gpu, dt, tol):
    import torch
    print(f"case krylov kernels {dt}", flush=True)
    tdt = getattr(torch, dt)
    cap, m1, n, k, nv = 4, 11, 150001, 3, 9          # nv > 8: two sweeps of the dot kernel
    g = torch.Generator().manual_seed(3)
    V = (torch.randn((cap, m1, n), generator=g, dtype=torch.float64)
         + 1j * torch.randn((cap, m1, n), generator=g, dtype=torch.float64)).to(tdt)
    w = (torch.randn((cap, n), generator=g, dtype=torch.float64)
         + 1j * torch.randn((cap, n), generator=g, dtype=torch.float64)).to(tdt)
    V64, w64 = V.to(torch.complex128), w.to(torch.complex128)
    ref = torch.einsum("rin,rn->ri", V64[:k, :nv].conj(), w64[:k])
    kry = gila.solve.KrylovOps()
    Vd, wd = V.cuda(), w.cuda()
    h = torch.empty((k, nv), dtype=torch.complex128, device="cuda")
    w2 = torch.empty(k, dtype=torch.float64, device="cuda")
    kry.dots(Vd, k, nv, wd, h, 0, w2=w2)
    torch.cuda.synchronize()
    assert (h.cpu() - ref).abs().max() <= tol * ref.abs().max()
    assert torch.allclose(w2.cpu(), (w64[:k].abs() ** 2).sum(1), rtol=1e-12, atol=0)
    # one right-hand side alone gives bitwise the same coefficients
    h1 = torch.empty((1, nv), dtype=torch.complex128, device="cuda")
    kry.dots(Vd[1:], 1, nv, wd[1:], h1, 0)
    torch.cuda.synchronize()
    assert torch.equal(h1.cpu()[0], h.cpu()[1])
    # w -= V h, with the squared norm of the result
    upd = wd.clone()
    n1 = torch.empty(k, dtype=torch.float64, device="cuda")
    kry.axpy(Vd, k, nv, h, upd, -1.0, 0, w2=n1)
    torch.cuda.synchronize()
    want = w64[:k] - torch.einsum("ri,rin->rn", ref, V64[:k, :nv])
    err = (upd[:k].cpu().to(torch.complex128) - want).abs().max() / want.abs().max()
    assert err <= (1e-5 if dt == "complex64" else 1e-12)
    assert torch.allclose(n1.cpu(), (upd[:k].cpu().to(torch.complex128).abs() ** 2).sum(1), rtol=1e-12, atol=0)
    assert torch.equal(upd[k:].cpu(), w[k:])          # rows beyond k untouched
    # row gather and scaled scatter
    idx = torch.tensor([4, 0, 10], dtype=torch.int64, device="cuda")
    b = torch.empty((k, n), dtype=tdt, device="cuda")
    kry.rows(Vd, k, idx, b, 0)
    torch.cuda.synchronize()
    assert all(torch.equal(b[r].cpu(), V[r, int(idx[r])]) for r in range(k))
    sc = torch.tensor([2.0, 0.5, 4.0], dtype=torch.float64, device="cuda")
    Vs = Vd.clone()
    kry.rows(Vs, k, idx, wd, 0, scale=sc)
    torch.cuda.synchronize()
    for r in range(k):
        got = Vs[r, int(idx[r])].cpu().to(torch.complex128)
        assert (got - w64[r] / float(sc[r])).abs().max() <= 1e-6 * w64[r].abs().max()


@pytest.mark.gpu
@pytest.mark.parametrize("dt", ["complex64", "complex128"])
def test_krylov_column_counts_and_next_vector(gila, gpu, dt):
    """Per right-hand-side column counts (a request in the middle of its own
    restart cycle inside a batch): the products over V[r, :ncol[r]] equal
    the full products with the columns beyond zeroed, bitwise, whatever the
    stale columns hold; the next-vector kernel divides by the device norm
    exactly as a host-side scale would and stages the same vector."""
    import torch
    print(f"case krylov column counts {dt}", flush=True)
    tdt = getattr(torch, dt)
    cap, m1, n, k, nv = 4, 12, 70001, 4, 11
    g = torch.Generator().manual_seed(11)
    V = (torch.randn((cap, m1, n), generator=g, dtype=torch.float64)
         + 1j * torch.randn((cap, m1, n), generator=g, dtype=torch.float64)).to(tdt)
    w = (torch.randn((cap, n), generator=g, dtype=torch.float64)
         + 1j * torch.randn((cap, n), generator=g, dtype=torch.float64)).to(tdt)
    cnt = [11, 3, 0, 9]                                   # full, partial, none, one sweep + 1
    Vz = V.clone()
    for r, c in enumerate(cnt):
        Vz[r, c:] = 0
    kry = gila.solve.KrylovOps()
    ncol = torch.tensor(cnt, dtype=torch.int64, device="cuda")
    Vd, Vzd, wd = V.cuda(), Vz.cuda(), w.cuda()
    h = torch.full((k, nv), 7.0 + 7.0j, dtype=torch.complex128, device="cuda")
    hz = torch.empty((k, nv), dtype=torch.complex128, device="cuda")
    w2, w2z = (torch.empty(k, dtype=torch.float64, device="cuda") for _ in range(2))
    kry.dots(Vd, k, nv, wd, h, 0, w2=w2, ncol=ncol)
    kry.dots(Vzd, k, nv, wd, hz, 0, w2=w2z)
    torch.cuda.synchronize()
    assert torch.equal(h.cpu(), hz.cpu())
    assert torch.equal(w2.cpu(), w2z.cpu())
    for r, c in enumerate(cnt):
        assert bool((h[r, c:] == 0).all())
    upd, updz = wd.clone(), wd.clone()
    n1, n1z = (torch.empty(k, dtype=torch.float64, device="cuda") for _ in range(2))
    hh = (h * 0.25).contiguous()
    kry.axpy(Vd, k, nv, hh, upd, -1.0, 0, w2=n1, ncol=ncol)
    kry.axpy(Vzd, k, nv, hh, updz, -1.0, 0, w2=n1z)
    torch.cuda.synchronize()
    assert torch.equal(upd.cpu(), updz.cpu())
    assert torch.equal(n1.cpu(), n1z.cpu())
    assert torch.equal(upd[2].cpu(), w[2])                # no columns: unchanged
    # next vector: V[r, idx[r]] = upd[r] / sqrt(n1[r]) and the staged copy;
    # the same as the scatter with a host-computed scale
    idx = torch.tensor([5, 0, 11, 2], dtype=torch.int64, device="cuda")
    Va, Vb = Vd.clone(), Vd.clone()
    stage = torch.zeros((cap, n), dtype=tdt, device="cuda")
    kry.next(Va, k, idx, upd, n1, 0, stage=stage)
    sc = torch.from_numpy(np.maximum(np.sqrt(n1.cpu().numpy()), 1e-30)).cuda()
    kry.rows(Vb, k, idx, upd, 0, scale=sc)
    torch.cuda.synchronize()
    assert torch.equal(Va.cpu(), Vb.cpu())
    for r in range(k):
        assert torch.equal(stage[r].cpu(), Va[r, int(idx[r])].cpu())
    assert bool((stage[k:] == 0).all())
    # a zero vector normalizes by the floor, not by zero
    z = torch.zeros((1, n), dtype=tdt, device="cuda")
    zn = torch.zeros(1, dtype=torch.float64, device="cuda")
    Vc = Vd.clone()
    kry.next(Vc, 1, idx[:1], z, zn, 0)
    torch.cuda.synchronize()
    assert bool((Vc[0, 5] == 0).all())


def _givens(gila, k, m, js, hcol, hn, cs, sn, H, g):
    import ctypes
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    gila._capi.check(gila.lib().gla_krylov_givens(k, m, p(js), p(hcol), hcol.shape[1], p(hn), p(cs), p(sn),
                                                  p(H), p(g)))


def test_krylov_givens_solves_the_hessenberg_least_squares(gila):
    """The host Givens update (gla_krylov_givens), step by step on random
    Hessenberg columns of several slots at different steps: after step j,
    |g[j+1]| equals the least-squares residual min ||beta e1 - Hbar y||
    (numpy lstsq on the unrotated Hessenberg), back substitution with the
    rotated R gives the lstsq solution, the rotated column has a zero
    subdiagonal, and a slot's numbers do not depend on the other slots."""
    rng = np.random.default_rng(5)
    m, k = 12, 3
    starts = [0, 4, 9]                                  # slots join at different steps
    cs = np.zeros((k, m))
    sn = np.zeros((k, m), np.complex128)
    H = np.zeros((k, m + 1, m), np.complex128)
    g = np.zeros((k, m + 1), np.complex128)
    beta = np.array([2.0, 0.5, 3.0])
    g[:, 0] = beta
    Hbar = np.zeros((k, m + 1, m), np.complex128)     # unrotated
    js = np.zeros(k, np.int64)
    solo = None
    for step in range(m + max(starts)):
        live = [s for s in range(k) if starts[s] <= step and js[s] < m]
        if not live:
            continue
        cols = np.zeros((k, m), np.complex128)
        hn = np.zeros(k)
        for s in live:
            j = js[s]
            cols[s, :j + 1] = rng.standard_normal(j + 1) + 1j * rng.standard_normal(j + 1)
            hn[s] = abs(rng.standard_normal()) + 0.1
            Hbar[s, :j + 1, j] = cols[s, :j + 1]
            Hbar[s, j + 1, j] = hn[s]
        # the live slots as one batch (slot order kept), then slot 1 alone
        idx = np.array(live)
        sub = [np.ascontiguousarray(a[idx]) for a in (cs, sn, H, g)]
        if 1 in live:
            one = [np.ascontiguousarray(a[1:2]) for a in (cs, sn, H, g)]
            _givens(gila, 1, m, js[1:2].copy(), np.ascontiguousarray(cols[1:2]), hn[1:2].copy(), *one)
        _givens(gila, len(live), m, js[idx].copy(), np.ascontiguousarray(cols[idx]), hn[idx].copy(), *sub)
        for a, b in zip((cs, sn, H, g), sub):
            a[idx] = b
        if 1 in live:
            pos = live.index(1)
            for a, b in zip(sub, one):
                assert np.array_equal(a[pos], b[0])
        for s in live:
            j = js[s]
            assert H[s, j + 1, j] == 0
            e1 = np.zeros(j + 2, np.complex128)
            e1[0] = beta[s]
            y, res, *_ = np.linalg.lstsq(Hbar[s, :j + 2, :j + 1], e1, rcond=None)
            r_ls = np.linalg.norm(e1 - Hbar[s, :j + 2, :j + 1] @ y)
            assert abs(abs(g[s, j + 1]) - r_ls) <= 1e-12 * beta[s]
            R = np.triu(H[s, :j + 1, :j + 1])
            yr = np.linalg.solve(R, g[s, :j + 1])
            assert np.linalg.norm(yr - y) <= 1e-9 * max(1.0, np.linalg.norm(y))
            js[s] += 1
    # zero column and zero diagonal: identity rotation, and the swap rotation
    cs1, sn1 = np.zeros((2, 2)), np.zeros((2, 2), np.complex128)
    H1, g1 = np.zeros((2, 3, 2), np.complex128), np.zeros((2, 3), np.complex128)
    g1[:, 0] = 1.0
    _givens(gila, 2, 2, np.zeros(2, np.int64), np.zeros((2, 1), np.complex128), np.array([0.0, 2.0]),
            cs1, sn1, H1, g1)
    assert cs1[0, 0] == 1 and sn1[0, 0] == 0 and g1[0, 1] == 0
    assert cs1[1, 0] == 0 and sn1[1, 0] == 1 and H1[1, 0, 0] == 2 and abs(g1[1, 1]) == 1
    with pytest.raises(Exception):
        _givens(gila, 1, 2, np.array([2], np.int64), np.zeros((1, 3), np.complex128), np.zeros(1),
                cs1, sn1, H1, g1)
