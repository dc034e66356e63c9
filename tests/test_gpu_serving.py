"""Continuous batching: requests join free slots of per-L pools while others
run, pools are admitted by HBM footprint, requests can be cancelled, a failed
device allocation fails only its request, and each result is identical to the
same scan run alone."""

import time

import numpy as np
import pytest

from rbim.serving import CancelledError, ScanRequest, ScanServer, poisson_requests
from rbim.simulation import OBSERVABLES, run_temperature_scan

pytestmark = pytest.mark.gpu


def _requests():
    return [
        ScanRequest(L=32, p=0.05, temperatures=np.linspace(1.5, 3.0, 6), n_realizations=6,
                    n_sweeps=1200, seed=11),
        ScanRequest(L=16, p=0.10, temperatures=np.linspace(1.0, 3.0, 4), n_realizations=3,
                    n_sweeps=3000, n_therm=300, n_bins=6, seed=12),
        ScanRequest(L=64, p=0.0, temperatures=[2.0, 2.3], n_realizations=4, n_sweeps=800,
                    seed=13, start="cold"),
        ScanRequest(L=256, p=0.08, temperatures=[1.8, 2.4], n_realizations=2, n_sweeps=60,
                    n_bins=3, seed=14),
    ]


def _alone(req):
    return run_temperature_scan(req.L, req.p, req.temperatures, n_realizations=req.n_realizations,
                                n_sweeps=req.n_sweeps, n_therm=req.n_therm, n_bins=req.n_bins,
                                seed=req.seed, start=req.start,
                                measure_percolation=req.measure_percolation)


def _same(a, b):
    assert np.array_equal(a["temperatures"], b["temperatures"])
    for name in OBSERVABLES:
        for key in ("value", "error", "thermal_error"):
            assert np.array_equal(a["observables"][name][key], b["observables"][name][key],
                                  equal_nan=True), (name, key)
    for key in ("L", "p", "n_realizations", "n_sweeps", "n_therm", "n_bins", "seed"):
        assert a["metadata"][key] == b["metadata"][key]


def test_results_in_the_mix_equal_results_alone():
    reqs = _requests()
    server = ScanServer(max_active=4, chunk_sweeps=200)
    try:
        tickets = []
        for i, req in enumerate(reqs):
            tickets.append(server.submit(req))
            time.sleep(0.01 * i)  # staggered arrivals while others are running
        mixed = [t.result(timeout=120) for t in tickets]
        stats = server.stats()
    finally:
        server.close()
    for req, res in zip(reqs, mixed):
        _same(res, _alone(req))
    assert stats["completed"] == len(reqs)
    assert stats["peak_active"] >= 2  # requests overlapped on the device
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0


def test_admission_by_memory_budget():
    from rbim.pool import pool_footprint

    reqs = [ScanRequest(L=64, p=0.05, temperatures=[1.5, 2.5], n_realizations=3, n_sweeps=600,
                        seed=20 + i) for i in range(3)]
    # room for one pool that holds one of them at a time: the others wait for its slots
    size = pool_footprint(64, 6, 32)
    server = ScanServer(memory_budget=size * 3 // 2, max_active=4, chunk_sweeps=300)
    try:
        tickets = [server.submit(r) for r in reqs]
        for t in tickets:
            t.result(timeout=120)
        stats = server.stats()
        assert stats["peak_active"] == 1 and stats["peak_device_bytes"] <= size * 3 // 2
        assert tickets[1].admitted >= tickets[0].finished
        too_big = server.submit(ScanRequest(L=256, p=0.1, temperatures=[2.0] * 8, n_realizations=8,
                                            n_sweeps=100))
        with pytest.raises(MemoryError):
            too_big.result(timeout=60)
        bad = server.submit(ScanRequest(L=2, p=0.1, temperatures=[2.0], n_sweeps=10))
        with pytest.raises(ValueError):
            bad.result(timeout=60)
    finally:
        server.close()
    for req, t in zip(reqs, tickets):
        _same(t.result(), _alone(req))


def test_new_requests_join_a_running_pool():
    """A short request submitted while a long one of the same L runs is loaded
    into free slots of the same pool, shares its launches and finishes first."""
    long_req = ScanRequest(L=32, p=0.05, temperatures=np.linspace(1.5, 3.0, 8), n_realizations=8,
                           n_sweeps=60000, seed=31)
    short_req = ScanRequest(L=32, p=0.10, temperatures=[2.0, 2.5], n_realizations=2, n_sweeps=400,
                            seed=32)
    server = ScanServer(chunk_sweeps=20000)
    try:
        t_long = server.submit(long_req)
        while t_long.admitted is None:
            time.sleep(0.001)
        t_short = server.submit(short_req)
        res_short = t_short.result(timeout=120)
        assert not t_long.done()
        res_long = t_long.result(timeout=120)
        stats = server.stats()
    finally:
        server.close()
    assert t_short.admitted < t_long.finished and t_short.finished < t_long.finished
    assert stats["pools"] == 1 and stats["peak_active"] == 2
    _same(res_short, _alone(short_req))
    _same(res_long, _alone(long_req))


def test_cancel_queued_and_running_requests():
    running = ScanRequest(L=16, p=0.05, temperatures=[2.0] * 4, n_realizations=4,
                          n_sweeps=400000, seed=41)
    queued = ScanRequest(L=16, p=0.05, temperatures=[2.0] * 4, n_realizations=4, n_sweeps=100,
                         seed=42)
    other = ScanRequest(L=16, p=0.05, temperatures=[2.0, 2.2], n_realizations=2, n_sweeps=300,
                        seed=43)
    # a pool of 16 slots: the second request waits behind the first
    server = ScanServer(slots_per_pool=16, chunk_sweeps=2000)
    try:
        t_run = server.submit(running)
        t_queued = server.submit(queued)
        while t_run.admitted is None:
            time.sleep(0.001)
        assert t_queued.cancel()
        time.sleep(0.05)
        assert t_run.cancel()
        with pytest.raises(CancelledError):
            t_queued.result(timeout=60)
        with pytest.raises(CancelledError):
            t_run.result(timeout=60)
        assert t_queued.admitted is None
        t_other = server.submit(other)  # the freed slots serve the next request
        res = t_other.result(timeout=60)
        assert not t_other.cancel()
    finally:
        server.close()
    _same(res, _alone(other))


def test_injected_allocation_failure_fails_only_its_request():
    from rbim import _hip

    first = ScanRequest(L=48, p=0.05, temperatures=[2.0, 2.4], n_realizations=2, n_sweeps=200,
                        seed=51)
    second = ScanRequest(L=48, p=0.05, temperatures=[2.1], n_realizations=3, n_sweeps=200, seed=52)
    server = ScanServer()
    try:
        _hip.lib().rbim_inject_alloc_failure(0)  # the next device allocation fails
        t1 = server.submit(first)
        with pytest.raises(MemoryError):
            t1.result(timeout=60)
        t2 = server.submit(second)  # the pool is created again, and serves
        res = t2.result(timeout=60)
    finally:
        _hip.lib().rbim_inject_alloc_failure(-1)
        server.close()
    _same(res, _alone(second))


def test_poisson_stream_reports_latency():
    arrivals = poisson_requests(rate_hz=40.0, duration_s=0.5, seed=3, Ls=(16, 32), n_temps=(2, 4),
                                realizations=(2, 4), sweeps=(100, 400))
    assert len(arrivals) >= 5
    server = ScanServer(max_active=8, chunk_sweeps=200)
    t0 = time.perf_counter()
    try:
        tickets = []
        for at, req in arrivals:
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(server.submit(req))
        for t in tickets:
            t.result(timeout=120)
        stats = server.stats()
    finally:
        server.close()
    assert stats["completed"] == len(arrivals)
    assert np.isfinite(stats["p50_latency_s"]) and np.isfinite(stats["site_updates_per_s"])
