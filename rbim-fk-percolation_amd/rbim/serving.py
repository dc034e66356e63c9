"""Continuous-batching scan server on one GPU.

Requests are temperature scans (the arguments of :func:`run_temperature_scan`).
Each is prepared once, at submission: its disorder, lattices and result
layout, exactly as run_temperature_scan builds them.  The lattices of all
requests of one L share a :class:`rbim.pool.SlotPool` -- one set of resident
slots on the device, advanced by shared launches on one HIP stream per L.  A
scheduler thread admits a queued request as soon as its pool has free slots
(pools are created on first use, within the device memory budget, by their
HBM footprint), so requests that arrive while others run join the running
launches at the next run boundary; a run ends when the first running request
of the pool completes (or after ``chunk_sweeps`` sweeps), its results are read
and its slots freed for the next requests.  Pools of different L run
concurrently on their own streams.

Every lattice of a request is keyed by the request's own seed and its task
index, never by slot, pool, run boundary or arrival time, so a request's
result is the same whether it runs alone or in any mix
(``tests/test_gpu_serving.py``).

    server = ScanServer()
    ticket = server.submit(ScanRequest(L=32, p=0.05, temperatures=np.linspace(1, 3, 8),
                                       n_realizations=8, n_sweeps=2000, seed=1))
    results = ticket.result()          # the run_temperature_scan results dict
    ticket.cancel()                    # (withdraws a queued or running request)
    server.stats()                     # p50 / p99 latency, site-updates/s
    server.close()
"""

from __future__ import annotations

import threading
import time
import weakref
from collections import deque
from concurrent.futures import CancelledError
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from rbim import _hip
from rbim.disorder import generate_disorder
from rbim.engine import engine_seed, moments_to_primitive_arrays
from rbim.pool import SlotPool, pool_footprint
from rbim.simulation import METADATA_CONVENTIONS, _assemble_arrays, _check_bins, scan_lattices

__all__ = ["ScanRequest", "Ticket", "ScanServer", "poisson_requests", "replay",
           "mean_request_site_updates", "CancelledError"]


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
        """(lattice specs, n_therm, sweeps_per_bin, metadata, temperatures, R)
        exactly as run_temperature_scan builds them."""
        if self.start not in ("hot", "cold"):
            raise ValueError(f"start must be 'hot' or 'cold', got {self.start!r}")
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


class _Prepared:
    """A request's lattices and result layout, built once."""

    def __init__(self, request: ScanRequest):
        (self.specs, self.n_therm, self.per_bin, self.metadata, self.temps,
         self.R) = request.prepare()
        self.n_bins = int(request.n_bins)
        self.total = self.n_therm + self.n_bins * self.per_bin
        self.need = len(self.specs)
        self.packed = SlotPool.pack(self.specs)


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
    _cancelled: bool = False
    _server: Optional[object] = None
    _prep: Optional[_Prepared] = None

    def done(self) -> bool:
        return self._event.is_set()

    def cancel(self) -> bool:
        """Withdraw the request: a queued one never starts, a running one
        leaves its slots at its pool's next run boundary; its result() then
        raises CancelledError.  False when it has already finished."""
        if self._event.is_set():
            return False
        self._cancelled = True
        server = self._server() if self._server is not None else None
        if server is not None:
            server._wake.set()
        return True

    def result(self, timeout: Optional[float] = None) -> Dict[str, object]:
        if not self._event.wait(timeout):
            raise TimeoutError("scan not finished")
        if self._error is not None:
            raise self._error
        return self._result

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.submitted


class _Running:
    def __init__(self, ticket: Ticket, slots: np.ndarray):
        self.ticket = ticket
        self.prep = ticket._prep
        self.slots = slots
        self.queued = 0  # sweeps of its slots queued so far


class _PoolState:
    def __init__(self, pool: SlotPool):
        self.pool = pool
        self.running: List[_Running] = []


def default_slots(L: int) -> int:
    """Slots of a pool created for side L: room for several of the serving
    stream's largest requests (40 temperatures x 32 realizations)."""
    return 4096 if L <= 64 else 2048 if L <= 128 else 1536


class ScanServer:
    """Admits scan requests into per-L slot pools and runs them concurrently."""

    def __init__(self, memory_budget: Optional[int] = None, max_active: int = 64,
                 chunk_sweeps: int = 1000, poll_s: float = 0.0005,
                 slots_per_pool: Optional[int] = None, max_bins: int = 32):
        free, _ = _hip.device_memory()
        self.memory_budget = int(memory_budget) if memory_budget else int(0.8 * free)
        self.max_active = int(max_active)
        self.chunk = int(chunk_sweeps)
        self.poll_s = poll_s
        self.slots_per_pool = slots_per_pool
        self.max_bins = int(max_bins)
        self._queue: deque = deque()
        self._pools: Dict[int, _PoolState] = {}
        self._used = 0
        self._lock = threading.Lock()
        self._wake = threading.Event()
        self._stop = False
        self._finished: List[Ticket] = []
        self._started = time.perf_counter()
        self.peak_active = 0
        self.peak_used = 0
        self.runs = 0
        self._thread = threading.Thread(target=self._loop, name="rbim-scan-server", daemon=True)
        self._thread.start()

    # ------------------------------------------------------------ client side
    def submit(self, request: ScanRequest) -> Ticket:
        """Queue a scan; it is prepared here (a bad request fails its ticket)."""
        ticket = Ticket(request, _server=weakref.ref(self))
        try:
            ticket._prep = _Prepared(request)
            ticket.site_updates = request.L * request.L * ticket._prep.need * ticket._prep.total
        except BaseException as exc:
            self._fail(ticket, exc)
            return ticket
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
                    if not self._queue and not self._n_running():
                        break
                time.sleep(0.001)
        with self._lock:
            self._stop = True
        self._wake.set()
        self._thread.join()

    def stats(self) -> Dict[str, float]:
        with self._lock:
            done = [t for t in self._finished if t._error is None]
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
            "pools": float(len(self._pools)),
            "runs": float(self.runs),
        }

    # ------------------------------------------------------------ scheduler
    def _n_running(self) -> int:
        return sum(len(ps.running) for ps in self._pools.values())

    def _fail(self, ticket: Ticket, exc: BaseException) -> None:
        ticket._error = exc
        ticket.finished = time.perf_counter()
        ticket._event.set()

    def _drop_pool(self, L: int) -> None:
        ps = self._pools.pop(L)
        self._used -= ps.pool.nbytes
        ps.pool.close()

    def _pool_for(self, prep: _Prepared, L: int) -> Optional[_PoolState]:
        """The pool that can take this request now, creating (or re-creating a
        drained, too small) one within the memory budget; None: wait."""
        ps = self._pools.get(L)
        if ps is not None:
            if ps.pool.n_slots >= prep.need and ps.pool.max_bins >= prep.n_bins:
                return ps
            if ps.running:
                return None  # too small: wait for it to drain, then grow it
            self._drop_pool(L)
        n_slots = max(self.slots_per_pool or default_slots(L), prep.need)
        max_bins = max(self.max_bins, prep.n_bins)
        nbytes = pool_footprint(L, n_slots, max_bins)
        if nbytes > self.memory_budget:
            # the smallest pool that holds this request alone
            n_slots, nbytes = prep.need, pool_footprint(L, prep.need, max_bins)
            if nbytes > self.memory_budget:
                raise MemoryError(f"a pool for this scan needs {nbytes} bytes of device memory, "
                                  f"budget is {self.memory_budget}")
        # make room by closing drained pools of other sizes
        for other in [k for k, v in self._pools.items() if not v.running]:
            if self._used + nbytes <= self.memory_budget:
                break
            self._drop_pool(other)
        if self._used + nbytes > self.memory_budget:
            return None  # wait for running scans to release memory
        pool = SlotPool(L, n_slots, max_bins)
        ps = self._pools[L] = _PoolState(pool)
        self._used += pool.nbytes
        self.peak_used = max(self.peak_used, self._used)
        return ps

    def _admit(self) -> bool:
        """Load queued requests into free slots (FIFO within each L)."""
        progressed = False
        blocked = set()
        with self._lock:
            queue = list(self._queue)
        for ticket in queue:
            L = ticket.request.L
            if ticket._cancelled:
                self._dequeue(ticket)
                self._fail(ticket, CancelledError("scan cancelled before it started"))
                progressed = True
                continue
            if L in blocked or self._n_running() >= self.max_active:
                continue
            prep = ticket._prep
            try:
                ps = self._pool_for(prep, L)
                if ps is None or ps.pool.free_slots < prep.need:
                    blocked.add(L)
                    continue
                slots = ps.pool.load(prep.packed, n_therm=prep.n_therm, n_bins=prep.n_bins,
                                     sweeps_per_bin=prep.per_bin, start=ticket.request.start)
            except BaseException as exc:  # e.g. MemoryError: fail this one, keep serving
                self._dequeue(ticket)
                self._fail(ticket, exc)
                progressed = True
                continue
            self._dequeue(ticket)
            ticket.admitted = time.perf_counter()
            ps.running.append(_Running(ticket, slots))
            self.peak_active = max(self.peak_active, self._n_running())
            progressed = True
        return progressed

    def _dequeue(self, ticket: Ticket) -> None:
        with self._lock:
            self._queue.remove(ticket)

    def _harvest(self, ps: _PoolState) -> bool:
        """With the pool idle: finish completed requests, drop cancelled ones."""
        progressed = False
        for run in list(ps.running):
            t = run.ticket
            if t._cancelled:
                ps.pool.release(run.slots)
                ps.running.remove(run)
                self._fail(t, CancelledError("scan cancelled"))
                progressed = True
                continue
            if run.queued < run.prep.total:
                continue
            req, prep = t.request, run.prep
            acc = ps.pool.moments(run.slots, prep.n_bins)
            ps.pool.release(run.slots)
            ps.running.remove(run)
            prims, bins = moments_to_primitive_arrays(acc, req.measure_percolation)
            t._result = _assemble_arrays(req.L, prep.temps, prims, bins, prep.R, prep.metadata)
            t.finished = time.perf_counter()
            with self._lock:
                self._finished.append(t)
            t._event.set()
            progressed = True
        return progressed

    def _launch(self, ps: _PoolState) -> bool:
        """Queue the pool's next run: until its first running request completes,
        at most chunk_sweeps sweeps."""
        if not ps.running:
            return False
        n = min(self.chunk, min(r.prep.total - r.queued for r in ps.running))
        ps.pool.run(n)
        for r in ps.running:
            r.queued = min(r.prep.total, r.queued + n)
        self.runs += 1
        return True

    def _loop(self) -> None:
        while True:
            with self._lock:
                stop = self._stop
            if stop:
                for L in list(self._pools):
                    self._drop_pool(L)
                return
            progressed = False
            for ps in list(self._pools.values()):
                if not ps.running or not ps.pool.idle():
                    continue
                try:
                    progressed |= self._harvest(ps)
                except BaseException as exc:  # a failed read: fail the pool's requests
                    for run in ps.running:
                        self._fail(run.ticket, exc)
                    ps.running.clear()
            progressed |= self._admit()
            for ps in list(self._pools.values()):
                if ps.running and ps.pool.idle():
                    try:
                        progressed |= self._launch(ps)
                    except BaseException as exc:
                        for run in ps.running:
                            self._fail(run.ticket, exc)
                        ps.running.clear()
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
