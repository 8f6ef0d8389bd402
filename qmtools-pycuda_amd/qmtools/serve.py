"""Continuous-batching request loop for density + Hartree.

Requests (molecule, grid) arrive at any time (`Server.submit`).  One loop
thread owns every GPU call:

* admission: pending requests join the next batch while the batch's device
  memory estimate plus the memory of the batches already on the GPU plus
  what every lane's native context keeps resident (its workspaces and FFT
  caches, measured after each batch, or the batch's own workspace estimate
  when that is larger) stays within `hbm_budget` (bytes), and the batch holds
  at most `max_batch` requests; a request larger than the whole budget is
  refused;
* execution: each batch runs on a free lane (its own native context and
  stream), so a batch admitted while others are running overlaps with them
  on the device;
* retirement: when a lane's completion event has fired, results are copied
  to the host (or kept on the device), latencies are stamped and the
  memory is released.

With Jacobi Poisson and `round_sweeps` set, the lanes run in rounds of that
many sweeps (whole replayed blocks): between two rounds a lane retires the requests that have
had all their sweeps and admits pending ones into its running set, so a
request waits for at most one round, not for a whole batch.  V is the only
state a block carries, so a request's result does not depend on the rounds
it shared or on when it joined; the native library reuses a lane's block
graph while its set of molecules is unchanged, so a round costs a few
graph replays.

Molecule tables and density matrices stay resident on the device between
requests (`qmtools.qmtools.device_molecule` caches them on the Molecule).
The GPU work is done by an executor object; `GpuExecutor` runs
`qmtools.batch.run`, and tests substitute a host-only executor to exercise
admission, failure isolation and cancellation without a GPU.
"""
from __future__ import annotations

import collections
import itertools
import math
import threading
import time

import numpy


class Request:
    _ids = itertools.count()

    def __init__(self, mol, grid, tag=None):
        self.rid = next(Request._ids)
        self.mol = mol
        self.grid = grid
        self.tag = tag
        self.t_submit = time.perf_counter()
        self.t_admit = None
        self.t_done = None
        self.density = None
        self.potential = None
        self.error = None
        self.cancelled = False
        self._done = threading.Event()

    @property
    def npts(self):
        return int(self.grid.npts)

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError(f"request {self.rid} not finished")
        if self.error is not None:
            raise self.error
        return self

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit

    def cancel(self):
        """Drop the request if it has not been admitted yet; returns True if it was dropped.

        Admission marks a request under the server's lock, and so does this test, so a request
        is either dropped here or runs, never both."""
        lock = getattr(self, "_lock", None)
        if lock is None:
            self.cancelled = True
            return self.t_admit is None
        with lock:
            if self._admitted:
                return False
            self.cancelled = True
            return True

    _admitted = False
    _left = 0        # sweeps still to run (round-based Jacobi lanes)
    _queued = 0      # of those, sweeps in rounds already queued on the lane
    _job = None      # the request's device state while it runs in rounds


def memory_estimate(req, precision="mixed", poisson="jacobi"):
    """Device bytes of one request in a batch: its density and potential grids and the per-item
    buffers of the Hartree step (seed, rho, held boundary and, in FP64, the sweep partner), plus P."""
    n = req.npts
    if poisson == "fft":
        per_point = 8 + 8 if precision == "fp64" else 4 + 4
    else:
        per_point = 8 * 5 if precision == "fp64" else 4 + 8 * 4 + 4 * 3
    nbf = int(req.mol.nbf)
    return n * per_point + 16 * nbf * nbf


_PHI_BUDGET = 4 << 30          # the native density keeps at most this much basis values per round
_GREEN_BUDGET = 4 << 30        # and at most this much of cached Green's-function transforms


def freespace_len(n):
    """Cyclic length of a free-space FFT axis of n points: the smallest length >= 2n with no prime
    factor above 7 (as the native library pads)."""
    v = 2 * n
    while True:
        x = v
        for r in (2, 3, 5, 7):
            while x % r == 0:
                x //= r
        if x == 1:
            return v
        v += 1


def _padded_shape(req):
    """Shape of the request's padded free-space FFT grid, or None without a grid shape."""
    shape = getattr(getattr(req, "grid", None), "shape", None)
    if shape is None or len(shape) < 3:
        return None
    return tuple(freespace_len(int(d)) for d in shape[-3:])


def _padded_points(req):
    """Points of the request's padded free-space FFT grid (an upper bound of 27 x npts without a grid
    shape: each axis pads to less than 3 n)."""
    m = _padded_shape(req)
    return 27 * req.npts if m is None else m[0] * m[1] * m[2]


def _green_points(reqs):
    """Values of the cached Green's-function transforms the batch needs: one octant,
    (m0/2 + 1)(m1/2 + 1)(m2/2 + 1), per distinct padded shape (at most 5 npts without a shape)."""
    shapes, loose = set(), 0
    for r in reqs:
        m = _padded_shape(r)
        if m is None:
            loose += 5 * r.npts
        else:
            shapes.add(m)
    return loose + sum((a // 2 + 1) * (b // 2 + 1) * (c // 2 + 1) for a, b, c in shapes)


def workspace_estimate(reqs, precision="mixed", poisson="jacobi"):
    """Device bytes a lane's native context grows to while it runs this batch: basis values of one
    round, the FP32 Jacobi correction buffers or the FFT work buffer and Green's-function cache."""
    elt = 8 if precision == "fp64" else 4
    npts = sum(r.npts for r in reqs)
    nbf_pad = max(((int(r.mol.nbf) + 63) // 64 * 64 for r in reqs), default=0)
    phi = min(_PHI_BUDGET, npts * nbf_pad * elt)
    if poisson == "fft":
        cplx = 16 if precision == "fp64" else 8
        padded = [_padded_points(r) for r in reqs]
        hartree = max(padded, default=0) * cplx + min(_GREEN_BUDGET, _green_points(reqs) * elt)
    else:
        hartree = 0 if precision == "fp64" else 12 * npts
    return phi + hartree


class GpuExecutor:
    """Runs batches with `qmtools.batch.run` on per-lane contexts and streams.

    results: "host" copies density and potential to the host, "device" keeps the device
    tensors on the request, "drop" releases them (throughput runs)."""

    def __init__(self, nlanes=2, precision="mixed", sweeps=2000, boundary="multipole", to_host=True,
                 poisson="jacobi", results=None):
        from . import batch
        self._batch = batch
        self.nlanes = nlanes
        self.lanes = [batch.Lane() for _ in range(nlanes)]
        self.precision = precision
        self.sweeps = sweeps
        self.boundary = boundary
        self.poisson = poisson
        self.results = results or ("host" if to_host else "device")

    def launch(self, lane_id, reqs):
        lane = self.lanes[lane_id]
        torch = __import__("torch")
        qs, vs = self._batch.run([r.mol for r in reqs], [r.grid for r in reqs], self.precision, self.sweeps,
                                 self.boundary, lane, poisson=self.poisson)
        ev = torch.cuda.Event()
        ev.record(lane.stream)
        return (ev, qs, vs)

    def done(self, handle):
        return handle[0].query()

    def finish(self, handle, reqs):
        ev, qs, vs = handle
        ev.synchronize()
        for r, q, v in zip(reqs, qs, vs):
            if self.results == "host":
                r.density = q.cpu().numpy()
                r.potential = v.cpu().numpy()
            elif self.results == "device":
                r.density, r.potential = q, v

    # -- rounds (Jacobi Poisson with admission between rounds) ----------------------------
    def launch_round(self, lane_id, new, running, sweeps):
        """Densities and Jacobi state of the `new` requests, then `sweeps` sweeps for every request
        of the lane (`running` + `new`); the state lives on the request until it is retired."""
        lane = self.lanes[lane_id]
        torch = __import__("torch")
        if new:
            mols = [r.mol for r in new]
            grids = [r.grid for r in new]
            qs = self._batch.density_batch(mols, grids, self.precision, lane)
            states = self._batch.jacobi_prepare(mols, grids, qs, self.precision, self.boundary, lane)
            for r, q, st in zip(new, qs, states):
                r._job = (q, st)
        reqs = list(running) + list(new)
        self._batch.jacobi_sweeps([r.grid for r in reqs], [r._job[1] for r in reqs], sweeps, self.precision, lane)
        ev = torch.cuda.Event()
        ev.record(lane.stream)
        return ev

    def done_round(self, handle):
        return handle.query()

    def wait_round(self, handle):
        handle.synchronize()

    def finish_round(self, handle, finished):
        handle.synchronize()
        for r in finished:
            q, (rho, v, tmp) = r._job
            if self.results == "host":
                r.density = q.cpu().numpy()
                r.potential = v.cpu().numpy()
            elif self.results == "device":
                r.density, r.potential = q, v
            r._job = None

    def drop_round(self, reqs):
        for r in reqs:
            r._job = None

    def resident_bytes(self, lane_id):
        """What the lane's native context keeps on the device between batches."""
        import ctypes
        from . import _native as nat
        b = ctypes.c_int64(0)
        nat.check(nat.library().qm_ctx_memory(self.lanes[lane_id].ctx, ctypes.byref(b)), self.lanes[lane_id].ctx)
        return int(b.value)

    def close(self):
        for lane in self.lanes:
            lane.close()


class Server:
    def __init__(self, executor=None, nlanes=2, precision="mixed", hbm_budget=64 << 30, max_batch=16,
                 sweeps=2000, boundary="multipole", to_host=True, idle_sleep=2e-4, poisson="jacobi",
                 results=None, round_sweeps=None):
        """round_sweeps: with Jacobi Poisson, sweeps per round between admissions, a multiple of the
        replayed block (100 in FP64, 50 in FP32) that divides `sweeps`; "auto" takes the largest such
        round up to 200 sweeps (none: whole batches); None or 0 (the default) runs every batch to
        the end before its lane admits again.  Rounds let a request start within a round of its
        arrival, but it then shares the lane's rounds with the requests already running, whose
        sweeps slow down in proportion: with 2000-sweep requests on 64^3-256^3 grids at 10 req/s
        the median latency was lower with whole batches (46 ms against 64 ms, mixed precision)."""
        self.precision = precision
        self.poisson = poisson
        self.sweeps = int(sweeps)
        self.round_sweeps = None
        if poisson == "jacobi" and round_sweeps:
            block = 100 if precision == "fp64" else 50
            if round_sweeps == "auto":
                fits = [k for k in range(block, min(200, self.sweeps) + 1, block) if self.sweeps % k == 0]
                self.round_sweeps = max(fits) if fits else None
            elif int(round_sweeps) % block or self.sweeps % int(round_sweeps):
                raise ValueError(f"round_sweeps must be a multiple of {block} that divides sweeps={sweeps}")
            else:
                self.round_sweeps = int(round_sweeps)
        self.executor = executor or GpuExecutor(nlanes, precision, sweeps, boundary, to_host, poisson, results)
        self.nlanes = nlanes if executor is None else getattr(executor, "nlanes", nlanes)
        self._resident = [0] * self.nlanes   # measured resident bytes of each lane's context
        self._reserved = [0] * self.nlanes   # what each lane is charged: max(resident, running batch's workspace)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.idle_sleep = idle_sleep
        self._pending = collections.deque()
        self._lock = threading.Lock()
        self._inflight = {}          # lane -> (handle, reqs, bytes)
        self._used = 0
        self._stop = False
        self._thread = None
        self.completed = []
        self.batches = []            # (t_launch, lane, n requests, bytes); rounds: n new requests
        self.max_concurrent = 0
        self.joined_running = 0      # requests admitted into a lane that was already running others

    # -- client side ---------------------------------------------------------
    def submit(self, mol, grid, tag=None):
        req = Request(mol, grid, tag)
        req._lock = self._lock
        if memory_estimate(req, self.precision, self.poisson) + \
                workspace_estimate([req], self.precision, self.poisson) > self.hbm_budget:
            req.error = MemoryError(f"request {req.rid} needs more than the HBM budget of {self.hbm_budget} B")
            req.t_done = time.perf_counter()
            req._done.set()
            return req
        with self._lock:
            self._pending.append(req)
        return req

    def start(self):
        self._stop = False
        self._thread = threading.Thread(target=self._loop, name="qmtools-serve", daemon=True)
        self._thread.start()
        return self

    def stop(self, drain=True):
        if drain:
            while True:
                with self._lock:
                    busy = bool(self._pending) or bool(self._inflight)
                if not busy:
                    break
                time.sleep(1e-3)
        self._stop = True
        if self._thread is not None:
            self._thread.join()
        self.executor.close()

    # -- loop ------------------------------------------------------------------
    def _retire(self):
        for lane, (handle, reqs, nbytes) in list(self._inflight.items()):
            try:
                if not self.executor.done(handle):
                    continue
                self.executor.finish(handle, reqs)
                err = None
            except Exception as e:  # a failing batch fails its own requests only
                err = e
            now = time.perf_counter()
            for r in reqs:
                r.error = err
                r.t_done = now
                r._done.set()
                self.completed.append(r)
            del self._inflight[lane]
            self._used -= nbytes
            measure = getattr(self.executor, "resident_bytes", None)
            if measure is not None:
                self._resident[lane] = max(self._resident[lane], measure(lane))
            self._reserved[lane] = self._resident[lane]

    def _admit(self, lane, running=()):
        """Requests for a batch on `lane` (joining `running`, the requests it already holds);
        returns (batch, request bytes, lane charge)."""
        batch, nbytes = [], 0
        running = list(running)
        others = sum(self._reserved) - self._reserved[lane]
        charge = max(self._resident[lane], workspace_estimate(running, self.precision, self.poisson))
        with self._lock:
            while self._pending and len(running) + len(batch) < self.max_batch:
                r = self._pending[0]
                if r.cancelled:
                    self._pending.popleft()
                    r.error = RuntimeError("cancelled")
                    r.t_done = time.perf_counter()
                    r._done.set()
                    continue
                need = memory_estimate(r, self.precision, self.poisson)
                lane_charge = max(self._resident[lane],
                                  workspace_estimate(running + batch + [r], self.precision, self.poisson))
                if self._used + nbytes + need + others + lane_charge > self.hbm_budget:
                    break
                self._pending.popleft()
                r._admitted = True   # under the lock: cancel() can no longer drop it
                batch.append(r)
                nbytes += need
                charge = lane_charge
        return batch, nbytes, charge

    def _loop(self):
        if self.round_sweeps and hasattr(self.executor, "launch_round"):
            return self._loop_rounds()
        while not self._stop:
            self._retire()
            launched = False
            for lane in range(self.nlanes):
                if lane in self._inflight:
                    continue
                batch, nbytes, charge = self._admit(lane)
                if not batch:
                    break
                self._reserved[lane] = charge
                now = time.perf_counter()
                for r in batch:
                    r.t_admit = now
                try:
                    handle = self.executor.launch(lane, batch)
                except Exception as e:
                    for r in batch:
                        r.error = e
                        r.t_done = time.perf_counter()
                        r._done.set()
                        self.completed.append(r)
                    continue
                self._inflight[lane] = (handle, batch, nbytes)
                self._used += nbytes
                self.batches.append((now, lane, len(batch), nbytes))
                self.max_concurrent = max(self.max_concurrent, len(self._inflight))
                launched = True
            if not launched:
                time.sleep(self.idle_sleep)
        self._retire()

    def _fail(self, reqs, err):
        now = time.perf_counter()
        for r in reqs:
            r.error = err
            r.t_done = now
            r._done.set()
            self.completed.append(r)

    def _loop_rounds(self):
        """Jacobi lanes in rounds.  Each lane keeps up to two rounds queued on its stream, so the
        device never waits for the loop; a round sweeps the lane's requests that still need sweeps
        beyond those already queued, and pending requests join at the next round launched (idle
        lanes first, then the least loaded lane).  A request retires when its last round is done."""
        from collections import deque
        queued = {lane: deque() for lane in range(self.nlanes)}   # (handle, requests, sweeps) in stream order
        members = {lane: [] for lane in range(self.nlanes)}       # requests whose state lives on the lane
        depth = 2
        while not self._stop or self._inflight:
            ready = []
            for lane in range(self.nlanes):
                dq = queued[lane]
                while dq and self._round_done(lane, dq, members):
                    pass
                if len(dq) < depth:
                    ready.append(lane)
            launched = False
            for lane in sorted(ready, key=lambda ln: len(members[ln])):
                dq = queued[lane]
                need = [r for r in members[lane] if r._left - r._queued > 0]
                if self._stop:
                    new, nbytes = [], 0
                    charge = max(self._resident[lane], workspace_estimate(members[lane], self.precision,
                                                                          self.poisson))
                else:
                    new, nbytes, charge = self._admit(lane, members[lane])
                if not need and not new:
                    if not dq:
                        self._inflight.pop(lane, None)
                        self._reserved[lane] = self._resident[lane]
                    continue
                self._reserved[lane] = charge
                now = time.perf_counter()
                for r in new:
                    r.t_admit = now
                    r._left = self.sweeps
                    r._queued = 0
                if new and members[lane]:
                    self.joined_running += len(new)
                length = self.round_sweeps
                try:
                    handle = self.executor.launch_round(lane, new, need, length)
                except Exception as e:
                    self._drain_failed(lane, dq, members, e, extra=new)
                    continue
                self._used += nbytes
                members[lane] = members[lane] + new
                for r in need + new:
                    r._queued += length
                dq.append((handle, need + new, length))
                self._inflight[lane] = (handle, members[lane], length)
                if new:
                    self.batches.append((now, lane, len(new), nbytes))
                self.max_concurrent = max(self.max_concurrent, len(self._inflight))
                launched = True
            if not launched:
                time.sleep(self.idle_sleep)

    def _round_done(self, lane, dq, members):
        """Completes the lane's oldest round if it has finished; returns whether it did."""
        handle, swept, length = dq[0]
        try:
            if not self.executor.done_round(handle):
                return False
            dq.popleft()
            for r in swept:
                r._left -= length
                r._queued -= length
            finished = [r for r in swept if r._left <= 0]
            self.executor.finish_round(handle, finished)
        except Exception as e:   # a failing round fails the lane's requests only
            self._drain_failed(lane, dq, members, e)
            return False
        now = time.perf_counter()
        for r in finished:
            r.t_done = now
            r._done.set()
            self.completed.append(r)
        members[lane] = [r for r in members[lane] if r._left > 0]
        self._used -= sum(memory_estimate(r, self.precision, self.poisson) for r in finished)
        measure = getattr(self.executor, "resident_bytes", None)
        if measure is not None:
            self._resident[lane] = max(self._resident[lane], measure(lane))
        return True

    def _drain_failed(self, lane, dq, members, err, extra=()):
        """Fails every request of the lane (and `extra`) after the rounds still queued have run."""
        for handle, _, _ in dq:
            try:
                self.executor.wait_round(handle)
            except Exception:
                pass
        dq.clear()
        reqs = list(members[lane]) + list(extra)
        self.executor.drop_round(reqs)
        self._fail(reqs, err)
        self._used -= sum(memory_estimate(r, self.precision, self.poisson) for r in members[lane])
        members[lane] = []
        self._inflight.pop(lane, None)
        self._reserved[lane] = self._resident[lane]

    # -- statistics ------------------------------------------------------------
    def stats(self):
        ok = [r for r in self.completed if r.error is None]
        lat = sorted(r.latency for r in ok)
        if not ok:
            return {"requests": 0}
        t0 = min(r.t_submit for r in ok)
        t1 = max(r.t_done for r in ok)

        def q(p):
            return lat[min(len(lat) - 1, max(0, int(math.ceil(p / 100.0 * len(lat))) - 1))]
        return {"requests": len(ok), "failed": len(self.completed) - len(ok),
                "grid_pts_per_s": float(numpy.sum([r.npts for r in ok]) / max(t1 - t0, 1e-9)),
                "p50_s": q(50), "p99_s": q(99), "mean_s": float(numpy.mean(lat)),
                "batches": len(self.batches), "max_concurrent_batches": self.max_concurrent}


def poisson_arrivals(server, items, rate, seed=0):
    """Submit (mol, grid) items with exponential inter-arrival times at `rate` per second."""
    rng = numpy.random.default_rng(seed)
    reqs = []
    t_next = time.perf_counter()
    for mol, grid in items:
        t_next += rng.exponential(1.0 / rate)
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        reqs.append(server.submit(mol, grid))
    return reqs
