"""Continuous-batching request loop for density + Hartree.

Requests (molecule, grid) arrive at any time (`Server.submit`).  One loop
thread owns every GPU call:

* admission: pending requests join the next batch while the batch's device
  memory estimate plus the memory of the batches already on the GPU stays
  within `hbm_budget` (bytes) and the batch holds at most `max_batch`
  requests; a request larger than the whole budget is refused;
* execution: each batch runs on a free lane (its own native context and
  stream), so a batch admitted while others are running overlaps with them
  on the device;
* retirement: when a lane's completion event has fired, results are copied
  to the host (or kept on the device), latencies are stamped and the
  memory is released.

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
        """Drop the request if it has not been admitted yet; returns True if it was dropped."""
        self.cancelled = True
        return self.t_admit is None


def memory_estimate(req, precision="mixed"):
    """Device bytes of one request in a batch (grids of the density and Hartree steps, P)."""
    n = req.npts
    per_point = 8 * 5 if precision == "fp64" else 4 + 8 * 4 + 4 * 3
    nbf = int(req.mol.nbf)
    return n * per_point + 16 * nbf * nbf


class GpuExecutor:
    """Runs batches with `qmtools.batch.run` on per-lane contexts and streams."""

    def __init__(self, nlanes=2, precision="mixed", sweeps=2000, boundary="multipole", to_host=True):
        from . import batch
        self._batch = batch
        self.lanes = [batch.Lane() for _ in range(nlanes)]
        self.precision = precision
        self.sweeps = sweeps
        self.boundary = boundary
        self.to_host = to_host

    def launch(self, lane_id, reqs):
        lane = self.lanes[lane_id]
        torch = __import__("torch")
        qs, vs = self._batch.run([r.mol for r in reqs], [r.grid for r in reqs], self.precision, self.sweeps,
                                 self.boundary, lane)
        ev = torch.cuda.Event()
        ev.record(lane.stream)
        return (ev, qs, vs)

    def done(self, handle):
        return handle[0].query()

    def finish(self, handle, reqs):
        ev, qs, vs = handle
        ev.synchronize()
        for r, q, v in zip(reqs, qs, vs):
            if self.to_host:
                r.density = q.cpu().numpy()
                r.potential = v.cpu().numpy()
            else:
                r.density, r.potential = q, v

    def close(self):
        for lane in self.lanes:
            lane.close()


class Server:
    def __init__(self, executor=None, nlanes=2, precision="mixed", hbm_budget=64 << 30, max_batch=16,
                 sweeps=2000, boundary="multipole", to_host=True, idle_sleep=2e-4):
        self.precision = precision
        self.executor = executor or GpuExecutor(nlanes, precision, sweeps, boundary, to_host)
        self.nlanes = nlanes if executor is None else getattr(executor, "nlanes", nlanes)
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
        self.batches = []            # (t_launch, lane, n requests, bytes)
        self.max_concurrent = 0

    # -- client side ---------------------------------------------------------
    def submit(self, mol, grid, tag=None):
        req = Request(mol, grid, tag)
        if memory_estimate(req, self.precision) > self.hbm_budget:
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

    def _admit(self, lane):
        batch, nbytes = [], 0
        with self._lock:
            while self._pending and len(batch) < self.max_batch:
                r = self._pending[0]
                if r.cancelled:
                    self._pending.popleft()
                    r.error = RuntimeError("cancelled")
                    r.t_done = time.perf_counter()
                    r._done.set()
                    continue
                need = memory_estimate(r, self.precision)
                if self._used + nbytes + need > self.hbm_budget:
                    break
                self._pending.popleft()
                batch.append(r)
                nbytes += need
        return batch, nbytes

    def _loop(self):
        while not self._stop:
            self._retire()
            launched = False
            for lane in range(self.nlanes):
                if lane in self._inflight:
                    continue
                batch, nbytes = self._admit(lane)
                if not batch:
                    break
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
