"""Continuous-batching solve server: admission, batching across grids and
request arrival times, and the alone-versus-in-the-mix identity.

CPU tests drive the server with operators backed by the oracle's matvec; the
GPU tests use the library's operators and check each result by the residual
of the FP64 oracle operator built from the same Green data."""
import numpy as np
import pytest
from fractions import Fraction


class OracleOps:
    """apply32 / apply64 of A = I - chi G through the oracle's matvec, with the
    interface the server expects of an operator pair."""

    def __init__(self, orc, cells, scale=1.0 / 32, gain=1.0):
        self.orc = orc
        self.cells = tuple(cells)
        self.ncell = int(np.prod(cells))
        self.n = 3 * self.ncell
        fur, _ = orc.self_operator_fur(self.cells, h=float(scale), model=0)
        self.fur = fur * gain
        self.fixed_bytes = 1000
        self.calls = []

    def ws_bytes(self, n32, n64):
        return 100 * (n32 + n64)

    def _apply(self, x, chi, y, dtype):
        self.calls.append((dtype, x.shape[0]))
        for r in range(x.shape[0]):
            v = x[r].numpy().astype(dtype).reshape(self.cells + (3,), order="F")
            c = chi[r].numpy().astype(dtype).reshape(self.cells, order="F")
            gx = self.orc.matvec(self.cells, self.cells, (1, 1, 1), self.fur, v, dtype=dtype)
            y[r].copy_(__import__("torch").from_numpy((v - c[..., None] * gx).reshape(-1, order="F")))

    def apply32(self, x, chi, y, stream):
        self._apply(x, chi, y, np.complex64)

    def apply64(self, x, chi, y, stream):
        self._apply(x, chi, y, np.complex128)


def _residual(orc, ops_fur, req):
    cells = req.cells
    x = req.result
    ax = x - req.chi[..., None] * orc.matvec(cells, cells, (1, 1, 1), ops_fur, x)
    return np.linalg.norm(ax - req.e_inc) / np.linalg.norm(req.e_inc)


def _cpu_server(gila, orc, registry, **kw):
    def factory(cells, scale, frq):
        ops = OracleOps(orc, cells)
        registry[cells] = ops
        return ops
    return gila.serve.Server(device="cpu", ops_factory=factory, **kw)


def test_server_solves_mixed_grids_on_cpu(gila, orc):
    reg = {}
    srv = _cpu_server(gila, orc, reg, mem_budget=1 << 30, max_batch=3)
    reqs = [gila.serve.random_request(100 + i, cells, tol=1e-8)
            for i, cells in enumerate([(4, 4, 4), (6, 4, 4), (4, 4, 4), (4, 4, 4), (6, 4, 4), (4, 4, 4)])]
    done = srv.run(reqs)
    assert len(done) == 6 and all(r.converged for r in done)
    for r in reqs:
        assert r.relres <= 1e-8
        assert _residual(orc, reg[r.cells].fur, r) <= 1e-8
    # the 4^3 grid holds 3 slots, so its 4 requests were batched up to 3 at a time
    assert max(k for d, k in reg[(4, 4, 4)].calls if d == np.complex64) == 3
    # the operators of both grids stayed resident and were reused
    assert set(srv.ops) == {((4, 4, 4), Fraction(1, 32), 1 + 0j), ((6, 4, 4), Fraction(1, 32), 1 + 0j)}


def test_wide_batch_uses_vector_rotations(gila, orc):
    # six requests on one grid: one Givens update call covers the batch;
    # results equal those solved alone
    reg = {}
    reqs = [gila.serve.random_request(500 + i, (4, 4, 4), tol=1e-9) for i in range(6)]
    _cpu_server(gila, orc, reg, mem_budget=1 << 30, max_batch=8).run(reqs)
    assert max(k for d, k in reg[(4, 4, 4)].calls if d == np.complex64) == 6
    for i, r in enumerate(reqs):
        assert r.converged and _residual(orc, reg[(4, 4, 4)].fur, r) <= 1e-9
        alone = gila.serve.random_request(500 + i, (4, 4, 4), tol=1e-9)
        _cpu_server(gila, orc, {}, mem_budget=1 << 30).run([alone])
        assert np.linalg.norm(alone.result - r.result) <= 1e-8 * np.linalg.norm(r.result)


def test_queued_steps_match_one_step_at_a_time(gila, orc):
    """Two Arnoldi steps queued per grid (the default) give the results and
    step counts of one step at a time: requests of a batch converge at
    different steps, and the step queued behind the one where a request's
    cycle ends is discarded for that request."""
    cells = [(4, 4, 4)] * 4 + [(6, 4, 4)] * 2
    make = lambda: [gila.serve.random_request(300 + i, c, tol=1e-9) for i, c in enumerate(cells)]
    one, two = make(), make()
    _cpu_server(gila, orc, {}, mem_budget=1 << 30, max_batch=4, depth=1).run(one)
    srv = _cpu_server(gila, orc, {}, mem_budget=1 << 30, max_batch=4, depth=2)
    srv.run(two)
    assert srv.ahead_steps > 0 and srv.discarded_steps > 0
    for a, b in zip(one, two):
        assert a.converged and b.converged
        assert (a.steps, a.cycles) == (b.steps, b.cycles)
        assert np.linalg.norm(a.result - b.result) <= 1e-12 * np.linalg.norm(a.result)
    with pytest.raises(ValueError):
        gila.serve.Server(device="cpu", depth=0)


def test_alone_and_in_the_mix_agree(gila, orc):
    reg = {}
    target = gila.serve.random_request(7, (4, 4, 4), tol=1e-9)
    alone = _cpu_server(gila, orc, reg, mem_budget=1 << 30).run([target])[0].result.copy()
    mix = [gila.serve.random_request(200 + i, c, tol=1e-9) for i, c in enumerate([(4, 4, 4), (6, 4, 4), (4, 4, 4)])]
    t2 = gila.serve.random_request(7, (4, 4, 4), tol=1e-9)
    # the target arrives after the others have started (joins mid-cycle)
    arrivals = [0.0, 0.0, 0.0, 0.05]
    _cpu_server(gila, orc, reg, mem_budget=1 << 30).run(mix + [t2], arrivals=arrivals)
    assert np.linalg.norm(t2.result - alone) <= 1e-8 * np.linalg.norm(alone)


def test_admission_respects_memory_budget(gila, orc):
    reg = {}
    srv = _cpu_server(gila, orc, reg, mem_budget=1 << 30)
    slot = gila.serve._Grid.slot_bytes(3 * 64, 64, 30, "fp32") + 200
    srv = _cpu_server(gila, orc, reg, mem_budget=1000 + 2 * slot + 10, max_batch=8)
    reqs = [gila.serve.random_request(300 + i, (4, 4, 4), tol=1e-6) for i in range(5)]
    done = srv.run(reqs)
    assert all(r.converged for r in done)
    assert max(k for d, k in reg[(4, 4, 4)].calls if d == np.complex64) == 2
    assert srv.used <= srv.mem_budget
    tiny = _cpu_server(gila, orc, reg, mem_budget=500)
    with pytest.raises(RuntimeError):
        tiny.run([gila.serve.random_request(1, (4, 4, 4))])


def test_zero_incident_field_completes_at_once(gila, orc):
    reg = {}
    srv = _cpu_server(gila, orc, reg, mem_budget=1 << 30)
    r = gila.serve.SolveRequest((4, 4, 4), np.ones((4, 4, 4), complex), np.zeros((4, 4, 4, 3), complex))
    done = srv.run([r])
    assert done[0].converged and not np.abs(done[0].result).max()


@pytest.mark.gpu
def test_server_on_gpu_matches_fp64_oracle_residual(gila, orc, gpu):
    print("case serve gpu mixed grids", flush=True)
    cells_list = [(16, 16, 16), (16, 16, 16), (24, 16, 8), (16, 16, 16)]
    reqs = [gila.serve.random_request(400 + i, c, tol=1e-6) for i, c in enumerate(cells_list)]
    srv = gila.serve.Server(max_batch=4, mem_budget=8 << 30)
    done = srv.run(reqs, arrivals=[0.0, 0.0, 0.0, 0.02], timeout=120)
    assert all(r.converged for r in done)
    for r in reqs:
        key = r.key
        fur = np.concatenate([f.reshape(-1, order="F") for f in srv.ops[key].mem64.egoFur])
        assert _residual(orc, fur, r) <= 2e-6
    target = gila.serve.random_request(400, (16, 16, 16), tol=1e-6)
    alone = gila.serve.Server(max_batch=4, mem_budget=8 << 30).run([target], timeout=60)[0]
    assert np.linalg.norm(alone.result - reqs[0].result) <= 1e-5 * np.linalg.norm(alone.result)


@pytest.mark.gpu
def test_server_queued_steps_on_gpu(gila, orc, gpu):
    """Two Arnoldi steps queued per grid on the GPU: the same step and cycle
    counts as one step at a time, results within rounding of each other and
    of the FP64 oracle residual."""
    print("case serve gpu depth 2", flush=True)
    cells_list = [(16, 16, 16), (16, 16, 16), (16, 16, 16), (24, 16, 8)]
    make = lambda: [gila.serve.random_request(450 + i, c, tol=1e-6) for i, c in enumerate(cells_list)]
    one, two = make(), make()
    gila.serve.Server(max_batch=4, mem_budget=8 << 30, depth=1).run(one, timeout=120)
    srv = gila.serve.Server(max_batch=4, mem_budget=8 << 30, depth=2)
    srv.run(two, timeout=120)
    assert srv.ahead_steps > 0
    for a, b in zip(one, two):
        assert a.converged and b.converged
        assert (a.steps, a.cycles) == (b.steps, b.cycles)
        assert np.linalg.norm(a.result - b.result) <= 1e-6 * np.linalg.norm(a.result)
        fur = np.concatenate([f.reshape(-1, order="F") for f in srv.ops[b.key].mem64.egoFur])
        assert _residual(orc, fur, b) <= 2e-6


@pytest.mark.gpu
def test_server_cfg5_domains_alone_and_in_the_mix(gila, orc, gpu):
    """Served-stream requests on the stream's own domain sizes (32^3,
    64x64x48, 128^3; 1/32-wavelength cells, chi in a random 10-40% box, plane
    wave): every result converged by the FP64 residual and checked by the
    FP64 oracle operator built from the same Green data (<= 1e-6 plus the
    rounding of the two FP64 residual evaluations), and the 128^3 request
    solved alone equals its result in the mix (FP32 inner solves, FP64
    refinement: agreement to 1e-5)."""
    shapes = [(32, 32, 32), (64, 64, 48), (128, 128, 128), (32, 32, 32), (64, 64, 48)]
    print("case serve cfg5 domains " + ", ".join("x".join(map(str, s)) for s in shapes), flush=True)
    reqs = [gila.serve.random_request(20000 + i, c, tol=1e-6) for i, c in enumerate(shapes)]
    srv = gila.serve.Server(max_batch=16, max_cycles=60)
    done = srv.run(reqs, arrivals=[0.0, 0.0, 0.05, 0.1, 0.1], timeout=240)
    assert all(r.converged for r in done)
    for r in reqs:
        fur = np.concatenate([f.reshape(-1, order="F") for f in srv.ops[r.key].mem64.egoFur])
        res = _residual(orc, fur, r)
        print(f"  {'x'.join(map(str, r.cells))}: steps={r.steps} relres={r.relres:.2e} oracle={res:.2e}", flush=True)
        assert res <= 1.05e-6
    target = gila.serve.random_request(20002, (128, 128, 128), tol=1e-6)
    alone = gila.serve.Server(max_batch=16, max_cycles=60).run([target], timeout=240)[0]
    assert alone.converged
    assert np.linalg.norm(alone.result - reqs[2].result) <= 1e-5 * np.linalg.norm(alone.result)
