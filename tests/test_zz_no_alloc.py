"""The allocation-free sweep (rbim_sweep_noalloc) against rbim_sweep, through
the C ABI: identical spins and sweep stats, bit for bit, for both one-lattice
engines, on consecutive calls whose second input is smaller than the first
(the shared scratch keeps the first call's contents), with and without the
measurement; lattices beyond the scratch's limit are refused."""

import ctypes

import numpy as np
import pytest

from conftest import random_disorder

pytestmark = pytest.mark.gpu

PLAIN, BATCHED = 1, 2


@pytest.fixture(scope="module")
def hip():
    from rbim import _hip

    _hip.context()
    return _hip


def sweep(hip, fn, L, spins, J_h, J_v, thr, seed, lid, g, measure, engine):
    out = spins.copy()
    st = hip.SweepStats()
    ctx = hip.context()
    hip.check(getattr(hip.lib(), fn)(ctx, L, hip.ptr(out), hip.ptr(J_h), hip.ptr(J_v), thr, seed,
                                     lid, g, measure, engine, ctypes.byref(st)), ctx)
    return out, {name: getattr(st, name) for name, _ in hip.SweepStats._fields_}


# (engine, sizes in call order): each second call is smaller than the first
CASES = [(PLAIN, (256, 100)), (PLAIN, (200, 37)), (BATCHED, (128, 16)), (BATCHED, (64, 33))]


@pytest.mark.parametrize("engine,sizes", CASES)
def test_noalloc_sweep_equals_sweep(hip, engine, sizes):
    from rbim.clusters import activation_threshold

    rng = np.random.default_rng(sum(sizes) + engine)
    for L in sizes:
        print(f"noalloc: engine={engine} L={L}", flush=True)
        J_h, J_v = random_disorder(rng, L, 0.1)
        spins = np.where(rng.random(L * L) < 0.5, -1, 1).astype(np.int8)
        thr = activation_threshold(1.0 / 2.2)
        seed, lid = int(rng.integers(1, 2**63)), int(rng.integers(0, 2**31))
        s_ref, s_new = spins, spins
        for g in range(3):  # consecutive sweeps, the first unmeasured
            measure = int(g > 0)
            s_ref, st_ref = sweep(hip, "rbim_sweep", L, s_ref, J_h, J_v, thr, seed, lid, g,
                                  measure, engine)
            s_new, st_new = sweep(hip, "rbim_sweep_noalloc", L, s_new, J_h, J_v, thr, seed, lid, g,
                                  measure, engine)
            assert np.array_equal(s_new, s_ref), (engine, L, g)
            if measure:
                assert st_new == st_ref, (engine, L, g)
                assert st_new["largest_size"] > 0
        assert not np.array_equal(s_new, spins)


def test_noalloc_refuses_lattices_beyond_its_scratch(hip):
    L = 257
    z = np.ones(L * L, np.int8)
    with pytest.raises(ValueError, match="L <= 256"):
        sweep(hip, "rbim_sweep_noalloc", L, z, z, z, 1 << 30, 1, 0, 0, 1, PLAIN)
    with pytest.raises(ValueError, match="LDS engine"):
        sweep(hip, "rbim_sweep_noalloc", 200, z[:40000], z[:40000], z[:40000], 1 << 30, 1, 0, 0,
              1, BATCHED)
