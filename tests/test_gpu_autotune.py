"""The autotuner picks a launch shape, caches it, and Batch uses it; any
shape gives the same bits."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tune_cache_and_invariance(tmp_path, monkeypatch):
    monkeypatch.setenv("RBIM_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    from rbim import autotune
    from rbim.engine import Batch

    entry = autotune.tune(16, n_lattices=48, sweeps=300)
    assert entry["threads"] in autotune.LDS_THREADS
    assert entry["sweeps_per_launch"] in autotune.LDS_SWEEPS
    # shapes whose sweeps per launch exceed the run (1000, 2000 > 300) are not tried
    assert len(entry["trials"]) == len(autotune.LDS_THREADS)
    assert autotune.lookup(16, 48)["threads"] == entry["threads"]
    specs = autotune._synthetic(16, 48)
    with Batch(specs, n_therm=20, n_bins=4, sweeps_per_bin=20) as tuned:
        assert (tuned.threads, tuned.sweeps_per_launch) == (entry["threads"], entry["sweeps_per_launch"])
        tuned.run()
        a = tuned.moments()
    with Batch(specs, n_therm=20, n_bins=4, sweeps_per_bin=20, threads=64, sweeps_per_launch=7) as other:
        other.run()
        assert np.array_equal(other.moments(), a)
    big = autotune.tune(192, n_lattices=4, sweeps=40)
    assert big["threads"] == 0 and big["sweeps_per_launch"] in autotune.GRAPH_SWEEPS
    # both tile sides divide 192 and both were timed
    assert {t["tile"] for t in big["trials"]} == {32, 64} and big["tile"] in (32, 64)
    specs = autotune._synthetic(192, 4)
    out = []
    for tile in (32, 64):
        with Batch(specs, n_therm=3, n_bins=2, sweeps_per_bin=3, tile=tile) as b:
            b.run()
            out.append((b.moments(), b.spins()))
    assert np.array_equal(out[0][0], out[1][0])
    assert all(np.array_equal(x, y) for x, y in zip(out[0][1], out[1][1]))
