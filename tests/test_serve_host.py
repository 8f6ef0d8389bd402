"""Request loop logic on the host (no GPU): admission by HBM budget, overlap of
batches on lanes, failure isolation, cancellation, oversize refusal."""
import threading
import time

import numpy
import pytest

from qmtools import Grid, Molecule
from qmtools.serve import Server, memory_estimate, workspace_estimate


class FakeExecutor:
    """Batches 'run' for a fixed time; a batch holding a request tagged 'fail' raises in finish()."""

    def __init__(self, nlanes=2, run_s=0.05):
        self.nlanes = nlanes
        self.run_s = run_s
        self.lock = threading.Lock()
        self.inflight_bytes = 0
        self.peak_bytes = 0
        self.launched = []

    def launch(self, lane, reqs):
        nbytes = sum(memory_estimate(r) for r in reqs)
        with self.lock:
            self.inflight_bytes += nbytes
            self.peak_bytes = max(self.peak_bytes, self.inflight_bytes)
            self.launched.append((lane, [r.rid for r in reqs]))
        return {"t": time.perf_counter() + self.run_s, "bytes": nbytes}

    def done(self, h):
        return time.perf_counter() >= h["t"]

    def finish(self, h, reqs):
        with self.lock:
            self.inflight_bytes -= h["bytes"]
        if any(r.tag == "fail" for r in reqs):
            raise RuntimeError("injected kernel error")
        for r in reqs:
            r.density = numpy.full(4, r.rid, dtype=float)

    def close(self):
        pass


def _mol(basis):
    return Molecule([[1, 0, 0, 0], [1, 0.74, 0, 0]], numpy.eye(10) * 0.1, basis)


def test_admission_budget_and_overlap(basis):
    mol = _mol(basis)
    grid = Grid((1, 32, 32, 32), 0.2)
    fake = type("R", (), {"npts": 32 ** 3, "mol": mol})()
    per = memory_estimate(fake) + workspace_estimate([fake])   # a request and the workspace it adds
    ex = FakeExecutor(nlanes=2, run_s=0.03)
    srv = Server(executor=ex, hbm_budget=3 * per, max_batch=2).start()
    reqs = [srv.submit(mol, grid) for _ in range(10)]
    for r in reqs:
        r.wait(10)
    srv.stop()
    assert all(r.error is None for r in reqs)
    assert ex.peak_bytes <= 3 * per                 # never above the budget
    assert srv.max_concurrent == 2                  # two batches on the device at once
    assert max(len(ids) for _, ids in ex.launched) <= 2
    # every request ran exactly once
    ran = sorted(i for _, ids in ex.launched for i in ids)
    assert ran == sorted(r.rid for r in reqs)
    st = srv.stats()
    assert st["requests"] == 10 and st["p99_s"] >= st["p50_s"] > 0


def test_failure_is_isolated_and_oversize_refused(basis):
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    ex = FakeExecutor(nlanes=2, run_s=0.01)
    srv = Server(executor=ex, hbm_budget=10 ** 9, max_batch=1).start()
    good1 = srv.submit(mol, grid)
    bad = srv.submit(mol, grid, tag="fail")
    good2 = srv.submit(mol, grid)
    huge = srv.submit(mol, Grid((1, 1024, 1024, 1024), 0.1))
    assert isinstance(huge.error, MemoryError)
    good1.wait(5)
    good2.wait(5)
    with pytest.raises(RuntimeError):
        bad.wait(5)
    srv.stop()
    assert good1.density[0] == good1.rid and good2.density[0] == good2.rid


def test_cancel_before_admission(basis):
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    ex = FakeExecutor(nlanes=1, run_s=0.05)
    srv = Server(executor=ex, hbm_budget=10 ** 9, max_batch=1).start()
    first = srv.submit(mol, grid)
    time.sleep(0.01)                 # first is running; the next waits in the queue
    second = srv.submit(mol, grid)
    assert second.cancel()
    first.wait(5)
    with pytest.raises(RuntimeError):
        second.wait(5)
    srv.stop()
    assert all(second.rid not in ids for _, ids in ex.launched)


def test_lane_workspace_is_charged(basis):
    """What a lane's context keeps resident after a batch counts against the budget: with one lane
    holding a large workspace, a second batch that would fit on its own is not admitted beside it."""
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    fake = type("R", (), {"npts": 16 ** 3, "mol": mol})()
    per = memory_estimate(fake) + workspace_estimate([fake])

    class Resident(FakeExecutor):
        def resident_bytes(self, lane):
            return 10 * per if lane == 0 else 0

    ex = Resident(nlanes=2, run_s=0.02)
    srv = Server(executor=ex, hbm_budget=11 * per, max_batch=1).start()
    first = srv.submit(mol, grid)
    first.wait(10)
    time.sleep(0.05)                      # retired: lane 0 now holds 10 per resident
    reqs = [srv.submit(mol, grid) for _ in range(6)]
    for r in reqs:
        r.wait(10)
    srv.stop()
    assert all(r.error is None for r in reqs)
    # 10 resident + one request fits in 11; a second request beside it does not
    assert srv.max_concurrent == 1


def test_cancel_after_admission_keeps_result(basis):
    """A request admitted into a batch cannot be dropped any more: cancel() says so and it completes."""
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    ex = FakeExecutor(nlanes=1, run_s=0.2)
    srv = Server(executor=ex, hbm_budget=10 ** 9, max_batch=1).start()
    r = srv.submit(mol, grid)
    t0 = time.perf_counter()
    while not r._admitted and time.perf_counter() - t0 < 5:
        time.sleep(1e-3)
    assert r.cancel() is False
    r.wait(10)
    srv.stop()
    assert r.error is None and r.density is not None


class FakeRoundExecutor:
    """Jacobi lanes in rounds: each round 'runs' for run_s; a round holding a request tagged 'fail'
    raises in finish_round()."""

    def __init__(self, nlanes=1, run_s=0.02):
        self.nlanes = nlanes
        self.run_s = run_s
        self.lock = threading.Lock()
        self.rounds = {}          # rid -> rounds it took part in
        self.sets = []            # (lane, rids) per round
        self.inflight_bytes = 0
        self.peak_bytes = 0
        self.tail = {}            # lane -> when its last queued round ends (rounds run in order)

    def launch_round(self, lane, new, running, sweeps):
        reqs = list(running) + list(new)
        t = max(time.perf_counter(), self.tail.get(lane, 0.0)) + self.run_s
        self.tail[lane] = t
        with self.lock:
            for r in new:
                r._job = "state"
                self.inflight_bytes += memory_estimate(r)
            self.peak_bytes = max(self.peak_bytes, self.inflight_bytes)
            for r in reqs:
                self.rounds[r.rid] = self.rounds.get(r.rid, 0) + 1
            self.sets.append((lane, [r.rid for r in reqs]))
        return {"t": t, "reqs": reqs}

    def done_round(self, h):
        return time.perf_counter() >= h["t"]

    def wait_round(self, h):
        while not self.done_round(h):
            time.sleep(1e-3)

    def finish_round(self, h, finished):
        if any(r.tag == "fail" for r in h["reqs"]):
            raise RuntimeError("injected kernel error")
        with self.lock:
            for r in finished:
                r.density = numpy.full(4, r.rid, dtype=float)
                r._job = None
                self.inflight_bytes -= memory_estimate(r)

    def drop_round(self, reqs):
        with self.lock:
            for r in reqs:
                if r._job is not None:
                    self.inflight_bytes -= memory_estimate(r)
                r._job = None

    def close(self):
        pass


def test_rounds_admit_between_rounds(basis):
    """A request submitted while a lane runs joins that lane at the next round boundary instead of
    waiting for the running requests to finish, and every request runs exactly sweeps / round rounds."""
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    ex = FakeRoundExecutor(nlanes=1, run_s=0.03)
    srv = Server(executor=ex, hbm_budget=10 ** 9, max_batch=4, sweeps=400, round_sweeps=100).start()
    first = srv.submit(mol, grid)
    while first.t_admit is None:
        time.sleep(1e-3)
    time.sleep(0.045)                                  # inside the second round
    late = srv.submit(mol, grid)
    for r in (first, late):
        r.wait(10)
    srv.stop()
    assert first.error is None and late.error is None
    assert late.t_admit < first.t_done                 # joined while the first was still running
    assert srv.joined_running == 1
    assert ex.rounds[first.rid] == 4 and ex.rounds[late.rid] == 4
    assert late.t_done > first.t_done
    assert any(set(ids) == {first.rid, late.rid} for _, ids in ex.sets)
    assert first._job is None and late._job is None   # device state released at retirement


def test_rounds_budget_and_max_batch(basis):
    mol = _mol(basis)
    grid = Grid((1, 32, 32, 32), 0.2)
    fake = type("R", (), {"npts": 32 ** 3, "mol": mol})()
    per = memory_estimate(fake) + workspace_estimate([fake])
    ex = FakeRoundExecutor(nlanes=2, run_s=0.01)
    srv = Server(executor=ex, hbm_budget=3 * per, max_batch=2, sweeps=200, round_sweeps=100).start()
    reqs = [srv.submit(mol, grid) for _ in range(9)]
    for r in reqs:
        r.wait(20)
    srv.stop()
    assert all(r.error is None for r in reqs)
    assert ex.peak_bytes <= 3 * per
    assert max(len(ids) for _, ids in ex.sets) <= 2
    assert all(ex.rounds[r.rid] == 2 for r in reqs)
    assert srv.stats()["requests"] == 9


def test_rounds_failure_stays_on_its_lane(basis):
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    ex = FakeRoundExecutor(nlanes=1, run_s=0.005)
    srv = Server(executor=ex, hbm_budget=10 ** 9, max_batch=1, sweeps=200, round_sweeps=100).start()
    good1 = srv.submit(mol, grid)
    bad = srv.submit(mol, grid, tag="fail")
    good2 = srv.submit(mol, grid)
    for r in (good1, bad, good2):
        r._done.wait(10)
    srv.stop()
    assert good1.error is None and good2.error is None
    assert isinstance(bad.error, RuntimeError)
    assert ex.inflight_bytes == 0


def test_round_size_must_fit_the_blocks():
    class E:
        nlanes = 1
    assert Server(executor=E(), precision="mixed", sweeps=300, round_sweeps="auto").round_sweeps == 150
    assert Server(executor=E(), precision="fp64", sweeps=2000, round_sweeps="auto").round_sweeps == 200
    assert Server(executor=E(), precision="fp64", sweeps=50, round_sweeps="auto").round_sweeps is None
    assert Server(executor=E(), precision="fp64", sweeps=2000).round_sweeps is None
    with pytest.raises(ValueError):
        Server(executor=E(), precision="fp64", sweeps=2000, round_sweeps=150)


def test_rounds_prefer_an_idle_lane(basis):
    """With a lane free, a new request starts there rather than joining a running lane."""
    mol = _mol(basis)
    grid = Grid((1, 16, 16, 16), 0.2)
    ex = FakeRoundExecutor(nlanes=2, run_s=0.01)
    srv = Server(executor=ex, hbm_budget=10 ** 9, max_batch=4, sweeps=800, round_sweeps=100).start()
    first = srv.submit(mol, grid)
    while first.t_admit is None:
        time.sleep(1e-3)
    time.sleep(0.025)
    second = srv.submit(mol, grid)
    for r in (first, second):
        r.wait(10)
    srv.stop()
    lanes = {rid: {ln for ln, ids in ex.sets if rid in ids} for rid in (first.rid, second.rid)}
    assert len(lanes[first.rid]) == 1 and len(lanes[second.rid]) == 1
    assert lanes[first.rid] != lanes[second.rid]
    assert srv.joined_running == 0 and srv.max_concurrent == 2


def test_freespace_padding_lengths():
    """The estimate pads like the native FFT: 2n when it has no prime factor above 7, else the next
    such length (176 = 16 x 11 -> 180, 304 = 16 x 19 -> 315, 368 = 16 x 23 -> 375)."""
    from qmtools.serve import freespace_len
    assert [freespace_len(n) for n in (1, 32, 48, 64, 80, 96, 128, 256)] == [2, 64, 96, 128, 160, 192, 256, 512]
    assert freespace_len(88) == 180 and freespace_len(152) == 315 and freespace_len(184) == 375
    for n in range(1, 400):
        v = freespace_len(n)
        assert v >= 2 * n
        x = v
        for r in (2, 3, 5, 7):
            while x % r == 0:
                x //= r
        assert x == 1
        for w in range(2 * n, v):
            y = w
            for r in (2, 3, 5, 7):
                while y % r == 0:
                    y //= r
            assert y != 1
