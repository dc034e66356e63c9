"""Slot pools through the C ABI: lattices of several loads with different
schedules share launches, slots are reused after release, and every lattice's
moment sums and spins equal those of the same lattice in a Batch, bit for bit
(the LDS kernel, the edge-slot sweep with the slot merge in LDS and in HBM,
and the multi-kernel tiled sweep)."""

import numpy as np
import pytest

from conftest import random_disorder

pytestmark = pytest.mark.gpu


def specs_for(L, n, seed, p=0.1, lid0=0, spins=False):
    from rbim.engine import LatticeSpec

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        J_h, J_v = random_disorder(rng, L, p)
        s0 = np.where(rng.random(L * L) < 0.5, -1, 1).astype(np.int8) if spins else None
        out.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / (1.4 + 0.3 * i), seed=seed * 7 + 1,
                               lattice_id=lid0 + i, spins0=s0))
    return out


def batch_result(specs, n_therm, n_bins, spb, start, engine):
    from rbim.engine import Batch

    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, start=start,
               engine=engine) as b:
        b.run()
        return b.moments(), np.stack(b.spins())


# (L, engine): LDS kernel (compile-time and run-time L), edge-slot sweep with the
# slot merge in LDS (L=64, 256) and in HBM (L=288), multi-kernel tiled sweep
CASES = [(32, 0), (24, 0), (64, 1), (256, 0), (288, 0), (40, 3)]


@pytest.mark.parametrize("L,engine", CASES)
def test_pool_equals_batch(L, engine):
    from rbim.pool import SlotPool

    # three loads with different schedules and starts; the second arrives while
    # the first is running, the third reuses the first one's released slots
    loads = [
        (specs_for(L, 3, 1), 5, 4, 3, "hot"),
        (specs_for(L, 2, 2, lid0=10, spins=True), 0, 2, 7, "hot"),
        (specs_for(L, 3, 3, lid0=20), 2, 3, 2, "cold"),
    ]
    print(f"pool: L={L} engine={engine}", flush=True)
    expected = [batch_result(sp, nt, nb, spb, st, engine) for sp, nt, nb, spb, st in loads]
    with SlotPool(L, n_slots=7, max_bins=4, engine=engine) as pool:
        sp, nt, nb, spb, st = loads[0]
        s0 = pool.load(sp, n_therm=nt, n_bins=nb, sweeps_per_bin=spb, start=st)
        pool.run(6)
        sp, nt, nb, spb, st = loads[1]
        s1 = pool.load(sp, n_therm=nt, n_bins=nb, sweeps_per_bin=spb, start=st)
        pool.run(5)
        pool.run(100)  # each slot stops at its own total
        done, total = pool.progress(np.concatenate([s0, s1]))
        assert np.array_equal(done, total) and list(total) == [17] * 3 + [14] * 2
        got0 = pool.moments(s0, 4), pool.spins(s0)
        got1 = pool.moments(s1, 2), pool.spins(s1)
        pool.release(s0)
        sp, nt, nb, spb, st = loads[2]
        s2 = pool.load(sp, n_therm=nt, n_bins=nb, sweeps_per_bin=spb, start=st)
        assert set(s2) == set(s0)  # the released slots, cleared by the load
        pool.run(3)
        pool.run(5)
        got2 = pool.moments(s2, 3), pool.spins(s2)
    for (acc, spins), (acc_ref, spins_ref) in zip((got0, got1, got2), expected):
        assert np.array_equal(acc, acc_ref)
        assert np.array_equal(spins, spins_ref)


def test_pool_checks_its_arguments():
    from rbim.pool import SlotPool

    with SlotPool(16, n_slots=4, max_bins=2) as pool:
        with pytest.raises(ValueError, match="n_bins"):
            pool.load(specs_for(16, 1, 5), n_therm=1, n_bins=3, sweeps_per_bin=1)
        with pytest.raises(RuntimeError, match="free slots"):
            pool.load(specs_for(16, 5, 5), n_therm=1, n_bins=1, sweeps_per_bin=1)
        with pytest.raises(ValueError, match="L=16"):
            pool.load(specs_for(32, 1, 5), n_therm=1, n_bins=1, sweeps_per_bin=1)
        slots = pool.load(specs_for(16, 2, 5), n_therm=1, n_bins=1, sweeps_per_bin=1)
        assert pool.free_slots == 2
        pool.run(0)
        pool.sync()
        done, total = pool.progress(slots)
        assert list(done) == [0, 0] and list(total) == [2, 2]
