"""Concurrent cells: batched sweeps and the serving loop (afid_amd/sweep.py,
afid_amd/serve.py over afid_steps_async / afid_steps_more / afid_batch_steps).

CPU: the request and sweep distributions, and the serving loop's admission
logic on stand-in cells: first come first served, never above the memory
budget, every request completed with all its steps, latencies measured from
arrival.
GPU: a cell stepped asynchronously equals afid_steps; cells stepped together
(batched sweep) and cells served while others run equal the same cells run
alone, bit for bit (the cells share no device data); the footprint estimate
covers what a context allocates.
"""
import time

import numpy as np
import pytest

from afid_amd.serve import Request, Server, serving_requests, smooth_sizes, summarize
from afid_amd.sweep import STEP_CTL, SWEEP_NXM, SWEEP_NYZ, SWEEP_PR, CellSpec, sweep_specs


def test_smooth_sizes():
    s = smooth_sizes(64, 384)
    assert s[0] == 64 and s[-1] == 384 and 375 in s and 77 not in s and 378 not in s
    for n in s:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        assert m == 1


def test_request_and_sweep_distributions():
    reqs = serving_requests(400, lam=4.0, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(np.diff(t).mean() - 0.25) < 0.05
    for r in reqs:
        p = r.spec.params
        assert 32 <= p.nxm <= 192 and p.nym == p.nzm and 64 <= p.nym <= 384
        assert 1e5 <= p.ray <= 1e8 and 0.7 <= p.pra <= 7.0 and 200 <= r.spec.steps <= 2000
    specs = sweep_specs(64, 500, seed=1)
    assert [s.seed for s in specs] == list(range(100, 164))
    for s in specs:
        p = s.params
        assert p.nxm in SWEEP_NXM and p.nym in SWEEP_NYZ and p.nzm == p.nym and p.pra in SWEEP_PR
        assert 1e5 <= p.ray <= 1e7 and p.ylen == p.zlen == 2.0 and s.steps == 500


class FakeCell:
    """Stand-in for engine.Cell: each step takes 0.2 ms of wall time.  A cell
    whose nxm is in HALT stops with errorcode 166 after 5 steps (its later steps
    are no-ops); one whose nxm is in FAIL raises AfidError from steps_more."""
    live = 0
    peak = 0
    HALT = ()
    FAIL = ()
    more_calls = {}

    def __init__(self, params):
        self.params = params
        self.deadline = 0.0
        self.queued = 0
        FakeCell.live += 1
        FakeCell.peak = max(FakeCell.peak, FakeCell.live)

    def tune(self):
        return 16

    def init_linear_noise(self, seed, amp):
        self.seed = seed

    def _queue(self, n):
        self.deadline = max(self.deadline, time.perf_counter()) + 2e-4 * n
        self.queued += n

    def steps_async(self, ntime0, n, *ctl):
        self._queue(n)

    def steps_more(self, n):
        FakeCell.more_calls[self.params.nxm] = FakeCell.more_calls.get(self.params.nxm, 0) + 1
        if self.params.nxm in FakeCell.FAIL:
            from afid_amd.engine import AfidError
            raise AfidError("steps_more: implicit factorisation needs a row interchange (code 1, hip 0)", 1)
        self._queue(n)

    def busy(self):
        return time.perf_counter() < self.deadline

    def steps_wait(self):
        while self.busy():
            time.sleep(1e-4)
        if self.params.nxm in FakeCell.HALT:
            return 166, 1e-3, min(self.queued, 5)
        return 0, 1e-3, self.queued

    def max_divergence(self):
        return 1e-15

    def close(self):
        FakeCell.live -= 1


def _req(rid, arrival, nxm, steps):
    from afid_amd.engine import CellParams
    return Request(rid, arrival, CellSpec(CellParams(nxm=nxm, nym=16, nzm=16), rid, steps))


def test_server_admission_by_budget():
    FakeCell.live = FakeCell.peak = 0
    est = lambda p: p.nxm * 1000               # "bytes"
    reqs = [_req(r, 0.002 * r, 40 + 10 * (r % 3), 30 + 7 * r) for r in range(8)]
    srv = Server(budget_bytes=100_000, chunk=8, make_cell=FakeCell, estimate=est)
    done = srv.run(reqs)
    assert sorted(c.rid for c in done) == list(range(8))
    by = {c.rid: c for c in done}
    for r in reqs:
        c = by[r.rid]
        assert c.steps_done == r.spec.steps and c.admitted >= r.arrival - 1e-9 and c.latency > 0
    adm = [by[r].admitted for r in range(8)]
    assert adm == sorted(adm)                 # first come, first served
    assert srv.peak_bytes <= 100_000 and FakeCell.peak == srv.peak_active >= 2
    s = summarize(done)
    assert s["requests"] == 8 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_server_ends_halted_and_failing_cells_early():
    """A cell whose step control halted gets no further chunks and gives its
    memory back at once; a library error of one cell ends that request only."""
    FakeCell.live = FakeCell.peak = 0
    FakeCell.HALT, FakeCell.FAIL, FakeCell.more_calls = (41,), (42,), {}
    try:
        reqs = [_req(0, 0.0, 41, 400), _req(1, 0.0, 42, 400), _req(2, 0.0, 40, 60), _req(3, 0.0, 43, 60)]
        srv = Server(budget_bytes=10_000, chunk=8, make_cell=FakeCell, estimate=lambda p: 1000)
        done = {c.rid: c for c in srv.run(reqs)}
        assert sorted(done) == [0, 1, 2, 3] and FakeCell.live == 0
        halted, failed = done[0], done[1]
        assert halted.errorcode == 166 and halted.steps_done == 5 and not halted.refused
        assert FakeCell.more_calls.get(41, 0) == 0          # halted after its first chunk
        assert failed.refused and failed.errorcode == 1 and failed.steps_done == 0
        assert done[2].steps_done == 60 and done[3].steps_done == 60
        assert done[2].errorcode == 0 == done[3].errorcode
        s = summarize(list(done.values()))
        assert s["refused"] == 1 and 166 in s["errorcodes"]
    finally:
        FakeCell.HALT, FakeCell.FAIL = (), ()


def test_server_rejects_request_above_budget():
    srv = Server(budget_bytes=10, make_cell=FakeCell, estimate=lambda p: 100)
    with pytest.raises(ValueError):
        srv.run([_req(0, 0.0, 32, 10)])


# ---- GPU ---------------------------------------------------------------------------
SMALL = [dict(nxm=24, nym=16, nzm=20, ray=1e5, pra=0.7), dict(nxm=17, nym=30, nzm=12, ray=1e6, pra=1.0),
         dict(nxm=31, nym=15, nzm=18, ray=2e6, pra=4.4)]


def _spec(q, steps):
    from afid_amd.engine import CellParams
    return CellSpec(CellParams(nsst=3, istr3=6, str3=6.0, alx3=1.0, ylen=2.0, zlen=2.0, **SMALL[q]), 40 + q, steps)


def _alone(spec):
    from afid_amd.engine import Cell
    c = Cell(spec.params)
    c.init_linear_noise(spec.seed, 1e-3)
    err, _, done = c.steps(0, spec.steps, **{k: STEP_CTL[k] for k in ("idtv", "dt", "dtmin", "dtmax", "limitcfl")})
    out = {nm: c.download(nm) for nm in ("vx", "vy", "vz", "temp", "pr")}
    out["time"] = c.time
    c.close()
    return out, err, done


@pytest.mark.gpu
def test_async_steps_equal_steps():
    from afid_amd.engine import Cell
    spec = _spec(0, 12)
    want, _, _ = _alone(spec)
    c = Cell(spec.params)
    c.init_linear_noise(spec.seed, 1e-3)
    ctl = STEP_CTL
    c.steps_async(0, 5, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
    c.steps_more(7)
    err, _, done = c.steps_wait()
    assert err == 0 and done == 12 and not c.busy()
    for nm in ("vx", "vy", "vz", "temp", "pr"):
        np.testing.assert_array_equal(c.download(nm), want[nm], err_msg=nm)
    assert c.time == want["time"]
    c.close()


@pytest.mark.gpu
def test_batched_cells_equal_cells_alone():
    from afid_amd.sweep import run_batched
    specs = [_spec(0, 15), _spec(1, 15), _spec(2, 22)]
    got = {}

    def keep(q, cell):
        got[q] = {nm: cell.download(nm) for nm in ("vx", "vy", "vz", "temp", "pr")}
        got[q]["time"] = cell.time

    print("batched sweep of 3 cells", flush=True)
    res = run_batched(specs, on_done=keep)
    assert res.steps_done == [15, 15, 22] and set(res.errorcodes) == {0}
    for q, s in enumerate(specs):
        want, _, _ = _alone(s)
        for nm in ("vx", "vy", "vz", "temp", "pr"):
            np.testing.assert_array_equal(got[q][nm], want[nm], err_msg=f"cell {q} {nm}")
        assert got[q]["time"] == want["time"]


@pytest.mark.gpu
def test_served_cells_equal_cells_alone():
    reqs = [Request(0, 0.0, _spec(0, 30)), Request(1, 0.0, _spec(1, 20)), Request(2, 0.01, _spec(2, 40))]
    got = {}

    def keep(req, cell):
        got[req.rid] = {nm: cell.download(nm) for nm in ("vx", "vy", "vz", "temp", "pr")}

    print("serving 3 requests", flush=True)
    srv = Server(budget_bytes=64e9, chunk=7, on_finish=keep)
    done = srv.run(reqs)
    assert srv.peak_active >= 2
    assert {c.rid: c.steps_done for c in done} == {0: 30, 1: 20, 2: 40}
    assert max(c.max_divergence for c in done) < 1e-8
    for r in reqs:
        want, _, _ = _alone(r.spec)
        for nm in ("vx", "vy", "vz", "temp", "pr"):
            np.testing.assert_array_equal(got[r.rid][nm], want[nm], err_msg=f"request {r.rid} {nm}")


@pytest.mark.gpu
def test_estimate_bytes_covers_allocation():
    from afid_amd.engine import Cell, estimate_bytes
    spec = _spec(1, 3)
    c = Cell(spec.params)
    c.init_linear_noise(1, 1e-3)
    c.steps(0, 3, 1, 1e-3)
    c.download("vx", halo=True)
    c.CalcStats()
    c.slabs([2] * 9)
    held = c.held_bytes()
    est = estimate_bytes(spec.params)
    assert held <= est <= 1.02 * held + 65536
    c.close()


class SingularCell(FakeCell):
    def steps_async(self, *a):
        from afid_amd.engine import AfidError
        raise AfidError("steps_async: singular pressure tridiagonal", 2)


def test_server_records_refused_request():
    def make(p):
        return SingularCell(p) if p.nxm == 33 else FakeCell(p)
    FakeCell.live = 0
    reqs = [_req(0, 0.0, 40, 20), _req(1, 0.0, 33, 20), _req(2, 0.001, 50, 20)]
    done = Server(budget_bytes=1e9, chunk=8, make_cell=make, estimate=lambda p: 1000).run(reqs)
    by = {c.rid: c for c in done}
    assert by[1].refused and by[1].errorcode == 2 and by[1].steps_done == 0
    assert by[0].steps_done == by[2].steps_done == 20 and FakeCell.live == 0
    s = summarize(done)
    assert s["requests"] == 3 and s["refused"] == 1


@pytest.mark.gpu
def test_tuned_cell_equals_untuned():
    """afid_tune changes only the launch shape of the stencil pass: a tuned
    cell's fields equal an untuned one's bit for bit."""
    from afid_amd.engine import Cell
    spec = _spec(0, 6)
    want, _, _ = _alone(spec)
    c = Cell(spec.params)
    zm = c.tune()
    assert zm in (8, 16, 32, 64)
    c.init_linear_noise(spec.seed, 1e-3)
    c.steps(0, spec.steps, **{k: STEP_CTL[k] for k in ("idtv", "dt", "dtmin", "dtmax", "limitcfl")})
    for nm in ("vx", "vy", "vz", "temp", "pr"):
        np.testing.assert_array_equal(c.download(nm), want[nm], err_msg=nm)
    c.close()
