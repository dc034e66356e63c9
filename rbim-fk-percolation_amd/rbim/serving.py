"""Continuous-batching scan server on one GPU.

Requests are temperature scans (the arguments of :func:`run_temperature_scan`).
A scheduler thread admits queued requests while the device memory budget and
the slot count allow -- admission is by the batch's HBM footprint, as reported
by ``rbim_batch_footprint`` -- and gives every admitted scan its own
:class:`rbim.engine.Batch` (state resident in HBM, its own HIP stream).  The
scheduler feeds each batch blocks of sweeps; batches on different streams run
concurrently on the device, so a request that arrives while others are running
starts as soon as memory and a slot free up, and its kernels overlap theirs.

Every lattice of a request is keyed by the request's own seed and its task
index, never by slot, stream or arrival time, so a request's result is the same
whether it runs alone or in any mix (``tests/test_gpu_serving.py``).

    server = ScanServer()
    ticket = server.submit(ScanRequest(L=32, p=0.05, temperatures=np.linspace(1, 3, 8),
                                       n_realizations=8, n_sweeps=2000, seed=1))
    results = ticket.result()          # the run_temperature_scan results dict
    server.stats()                     # p50 / p99 latency, site-updates/s
    server.close()
"""

from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from rbim import _hip
from rbim.disorder import generate_disorder
from rbim.engine import Batch, engine_seed, footprint
from rbim.simulation import (
    METADATA_CONVENTIONS,
    _arrays_from_batch,
    _assemble_arrays,
    _check_bins,
    scan_lattices,
)

__all__ = ["ScanRequest", "Ticket", "ScanServer", "poisson_requests", "replay",
           "mean_request_site_updates"]


@dataclass
class ScanRequest:
    """One temperature scan; fields mirror run_temperature_scan's arguments."""

    L: int
    p: float
    temperatures: Sequence[float]
    n_realizations: int = 1
    n_sweeps: int = 10_000
    n_therm: Optional[int] = None
    n_bins: int = 10
    seed: int = 0
    measure_percolation: bool = True
    start: str = "hot"
    disorder: Optional[Sequence[tuple]] = None

    def prepare(self):
        """(lattice specs, n_therm, sweeps_per_bin, metadata) exactly as
        run_temperature_scan builds them."""
        temps = [float(t) for t in self.temperatures]
        n_therm = self.n_therm if self.n_therm is not None else max(1, self.n_sweeps // 10)
        per_bin = _check_bins(self.n_sweeps, self.n_bins)
        disorder_seeds, run_seeds = np.random.SeedSequence(self.seed).spawn(2)
        if self.disorder is None:
            rng = np.random.default_rng(disorder_seeds)
            realizations = [generate_disorder(self.L, self.p, rng) for _ in range(self.n_realizations)]
        else:
            realizations = list(self.disorder)[: self.n_realizations]
            if len(realizations) < self.n_realizations:
                raise ValueError("not enough disorder realizations supplied")
        specs = scan_lattices(self.L, realizations, temps, engine_seed(run_seeds))
        metadata = {
            "L": self.L, "p": self.p, "n_realizations": len(realizations),
            "n_sweeps": self.n_sweeps, "n_therm": n_therm, "n_bins": self.n_bins,
            "seed": self.seed, "measure_percolation": self.measure_percolation,
            "start": self.start, **METADATA_CONVENTIONS,
            "engine": "rbim-hip gfx950", "rng": "Philox4x32-10 keyed by (seed, lattice, sweep, site pair)",
        }
        return specs, n_therm, per_bin, metadata, temps, len(realizations)


@dataclass
class Ticket:
    """Handle of a submitted request."""

    request: ScanRequest
    submitted: float = field(default_factory=time.perf_counter)
    admitted: Optional[float] = None
    finished: Optional[float] = None
    site_updates: int = 0
    _event: threading.Event = field(default_factory=threading.Event)
    _result: Optional[Dict[str, object]] = None
    _error: Optional[BaseException] = None

    def done(self) -> bool:
        return self._event.is_set()

    def result(self, timeout: Optional[float] = None) -> Dict[str, object]:
        if not self._event.wait(timeout):
            raise TimeoutError("scan not finished")
        if self._error is not None:
            raise self._error
        return self._result

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.submitted


class _Active:
    def __init__(self, ticket: Ticket, batch: Batch, total: int, R: int, temps, metadata, nbytes):
        self.ticket = ticket
        self.batch = batch
        self.remaining = total
        self.R = R
        self.temps = temps
        self.metadata = metadata
        self.nbytes = nbytes


class ScanServer:
    """Admits scan requests by HBM footprint and runs them concurrently."""

    def __init__(self, memory_budget: Optional[int] = None, max_active: int = 8,
                 chunk_sweeps: int = 1000, poll_s: float = 0.0005):
        free, _ = _hip.device_memory()
        self.memory_budget = int(memory_budget) if memory_budget else int(0.8 * free)
        self.max_active = int(max_active)
        self.chunk = int(chunk_sweeps)
        self.poll_s = poll_s
        self._queue: deque = deque()
        self._active: List[_Active] = []
        self._used = 0
        self._lock = threading.Lock()
        self._wake = threading.Event()
        self._stop = False
        self._finished: List[Ticket] = []
        self._started = time.perf_counter()
        self.peak_active = 0
        self.peak_used = 0
        self._thread = threading.Thread(target=self._loop, name="rbim-scan-server", daemon=True)
        self._thread.start()

    # ------------------------------------------------------------ client side
    def submit(self, request: ScanRequest) -> Ticket:
        ticket = Ticket(request)
        with self._lock:
            if self._stop:
                raise RuntimeError("server is closed")
            self._queue.append(ticket)
        self._wake.set()
        return ticket

    def close(self, wait: bool = True) -> None:
        if wait:
            while True:
                with self._lock:
                    if not self._queue and not self._active:
                        break
                time.sleep(0.001)
        with self._lock:
            self._stop = True
        self._wake.set()
        self._thread.join()

    def stats(self) -> Dict[str, float]:
        with self._lock:
            done = list(self._finished)
        lat = np.array([t.latency for t in done if t.latency is not None])
        updates = sum(t.site_updates for t in done)
        span = (max((t.finished for t in done), default=self._started) -
                min((t.submitted for t in done), default=self._started))
        return {
            "completed": len(done),
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else float("nan"),
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else float("nan"),
            "site_updates": float(updates),
            "site_updates_per_s": float(updates / span) if span > 0 else float("nan"),
            "peak_active": float(self.peak_active),
            "peak_device_bytes": float(self.peak_used),
            "memory_budget": float(self.memory_budget),
        }

    # ------------------------------------------------------------ scheduler
    def _admit(self) -> None:
        while True:
            with self._lock:
                if not self._queue or len(self._active) >= self.max_active:
                    return
                ticket = self._queue[0]
            try:
                specs, n_therm, per_bin, metadata, temps, R = ticket.request.prepare()
                nbytes = footprint(specs, n_bins=ticket.request.n_bins)
            except BaseException as exc:  # bad request: fail it, keep serving
                with self._lock:
                    self._queue.popleft()
                self._fail(ticket, exc)
                continue
            if nbytes > self.memory_budget:
                with self._lock:
                    self._queue.popleft()
                self._fail(ticket, MemoryError(
                    f"scan needs {nbytes} bytes of device memory, budget is {self.memory_budget}"))
                continue
            if self._used + nbytes > self.memory_budget:
                return  # wait for a running scan to release memory
            try:
                batch = Batch(specs, n_therm=n_therm, n_bins=ticket.request.n_bins,
                              sweeps_per_bin=per_bin, start=ticket.request.start)
            except BaseException as exc:
                with self._lock:
                    self._queue.popleft()
                self._fail(ticket, exc)
                continue
            with self._lock:
                self._queue.popleft()
                ticket.admitted = time.perf_counter()
                ticket.site_updates = batch.site_updates(batch.total_sweeps)
                self._active.append(_Active(ticket, batch, batch.total_sweeps, R, temps, metadata, nbytes))
                self._used += nbytes
                self.peak_active = max(self.peak_active, len(self._active))
                self.peak_used = max(self.peak_used, self._used)

    def _fail(self, ticket: Ticket, exc: BaseException) -> None:
        ticket._error = exc
        ticket.finished = time.perf_counter()
        ticket._event.set()

    def _step(self, act: _Active) -> bool:
        """Advance one admitted scan; True when it has finished."""
        if not act.batch.done():
            return False
        if act.remaining > 0:
            n = min(self.chunk, act.remaining)
            act.batch.run(n)
            act.remaining -= n
            return False
        req = act.ticket.request
        prims, bins = _arrays_from_batch(act.batch, req.measure_percolation)
        act.batch.close()
        act.ticket._result = _assemble_arrays(req.L, act.temps, prims, bins, act.R, act.metadata)
        act.ticket.finished = time.perf_counter()
        act.ticket._event.set()
        return True

    def _loop(self) -> None:
        while True:
            with self._lock:
                stop = self._stop
            if stop:
                for act in list(self._active):
                    act.batch.close()
                return
            self._admit()
            progressed = False
            for act in list(self._active):
                try:
                    finished = self._step(act)
                except BaseException as exc:
                    act.batch.close()
                    self._fail(act.ticket, exc)
                    finished = True
                if finished:
                    progressed = True
                    with self._lock:
                        self._active.remove(act)
                        self._used -= act.nbytes
                        self._finished.append(act.ticket)
            if not progressed:
                self._wake.wait(self.poll_s)
                self._wake.clear()


def replay(server: "ScanServer", arrivals, sweep_scale: float = 1.0):
    """Submit (arrival time, ScanRequest) pairs at their times (each request's
    sweeps multiplied by ``sweep_scale``), wait for all of them, and return
    (server stats, offered site-updates, wall seconds)."""
    offered = 0
    for _, req in arrivals:
        req.n_sweeps = max(req.n_bins, int(req.n_sweeps * sweep_scale))
        therm = req.n_therm if req.n_therm is not None else max(1, req.n_sweeps // 10)
        offered += req.L * req.L * req.n_realizations * len(req.temperatures) * (req.n_sweeps + therm)
    t0 = time.perf_counter()
    tickets = []
    for at, req in arrivals:
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets.append(server.submit(req))
    for t in tickets:
        t.result()
    return server.stats(), offered, time.perf_counter() - t0


def mean_request_site_updates(Ls=(32, 64, 128, 256), n_temps=(8, 40), realizations=(4, 32),
                              sweeps=(1000, 20000)) -> float:
    """Expected site-updates of one request of :func:`poisson_requests`'
    stream (independent uniform draws; n_therm = n_sweeps // 10)."""
    mean_sites = float(np.mean(np.asarray(Ls, dtype=float) ** 2))
    mean_t = (n_temps[0] + n_temps[1]) / 2.0
    mean_r = (realizations[0] + realizations[1]) / 2.0
    mean_sw = float(np.mean([k + k // 10 for k in range(sweeps[0], sweeps[1] + 1)]))
    return mean_sites * mean_t * mean_r * mean_sw


def poisson_requests(rate_hz: float, duration_s: float, seed: int = 0,
                     Ls=(32, 64, 128, 256), p_max: float = 0.12, n_temps=(8, 40),
                     realizations=(4, 32), sweeps=(1000, 20000)):
    """(arrival time, ScanRequest) pairs of a Poisson stream (the serving workload:
    L from ``Ls``, p ~ U[0, p_max], n_T, R and sweeps uniform in their ranges)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    k = 0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            return out
        n_t = int(rng.integers(n_temps[0], n_temps[1] + 1))
        out.append((t, ScanRequest(
            L=int(rng.choice(Ls)), p=float(rng.uniform(0.0, p_max)),
            temperatures=np.linspace(1.0, 3.0, n_t),
            n_realizations=int(rng.integers(realizations[0], realizations[1] + 1)),
            n_sweeps=int(rng.integers(sweeps[0], sweeps[1] + 1)), seed=1000 + k)))
        k += 1
