"""Continuous batching: requests admitted by HBM footprint run concurrently,
and each result is identical to the same scan run alone."""

import time

import numpy as np
import pytest

from rbim.serving import ScanRequest, ScanServer, poisson_requests
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
    from rbim.engine import footprint

    reqs = [ScanRequest(L=64, p=0.05, temperatures=[1.5, 2.5], n_realizations=3, n_sweeps=600,
                        seed=20 + i) for i in range(3)]
    size = footprint(reqs[0].prepare()[0], n_bins=reqs[0].n_bins)
    # room for one of them at a time: the others wait for memory to be released
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
