"""The reference's own tests (tests/test_clusters.py, tests/test_wrapping.py,
tests/test_physics.py), restated against the drop-in ``rbim`` package, whose
every sweep, label and measurement runs on the GPU."""

import numpy as np
import pytest

import rbim
from rbim import _hip
from rbim.clusters import activate_bonds, cluster_statistics, label_and_wrap, label_clusters
from rbim.clusters import spans_along, wraps_along
from rbim.lattice import build_lattice, couplings_from_disorder
from rbim.simulation import derive_observables, run_single_realization, run_temperature_scan
from rbim.swendsen_wang import random_spins, sw_sweep

pytestmark = pytest.mark.gpu

TC_EXACT = 2.0 / np.log(1.0 + np.sqrt(2.0))


def ferro(L):
    ones = np.ones((L, L), np.int8)
    return couplings_from_disorder(ones, ones)


# ------------------------------------------------ test_clusters.py:30-111
def test_infinite_temperature_activates_nothing():
    L = 8
    lattice = build_lattice(L)
    spins = np.ones(lattice.n_sites, np.int64)
    active = activate_bonds(spins, lattice, ferro(L), 0.0, np.random.default_rng(0))
    assert not active.any()
    n, _ = label_clusters(active, lattice)
    assert n == lattice.n_sites
    assert not wraps_along(active, lattice, "x") and not spans_along(active, lattice, "y")


def test_zero_temperature_ferromagnet_is_one_wrapping_cluster():
    L = 8
    lattice = build_lattice(L)
    spins = np.ones(lattice.n_sites, np.int64)
    active = activate_bonds(spins, lattice, ferro(L), 50.0, np.random.default_rng(0))
    assert active.all()
    assert label_clusters(active, lattice)[0] == 1
    assert wraps_along(active, lattice, "x") and wraps_along(active, lattice, "y")


def test_only_satisfied_bonds_are_activated():
    L = 12
    lattice = build_lattice(L)
    rng = np.random.default_rng(7)
    J = couplings_from_disorder(*rbim.generate_disorder(L, 0.3, rng))
    spins = rng.integers(0, 2, size=lattice.n_sites) * 2 - 1
    active = activate_bonds(spins, lattice, J, 0.6, rng)
    ends = lattice.bonds[active]
    assert active.any()
    assert np.all(J[active] * spins[ends[:, 0]] * spins[ends[:, 1]] > 0)


def test_cluster_statistics_are_consistent():
    L = 10
    lattice = build_lattice(L)
    rng = np.random.default_rng(3)
    J = couplings_from_disorder(*rbim.generate_disorder(L, 0.15, rng))
    spins = rng.integers(0, 2, size=lattice.n_sites) * 2 - 1
    active = activate_bonds(spins, lattice, J, 0.5, rng)
    n, labels = label_clusters(active, lattice)
    sizes, mags = cluster_statistics(labels, spins, n)
    assert sizes.sum() == lattice.n_sites
    assert mags.sum() == pytest.approx(spins.sum())
    assert np.all(np.abs(mags) <= sizes + 1e-9)
    assert np.array_equal(sizes, np.bincount(labels, minlength=n))
    assert np.array_equal(mags, np.bincount(labels, weights=spins.astype(float), minlength=n))


def test_spanning_and_wrapping_are_distinct():
    L = 6
    lattice = build_lattice(L)
    active = np.zeros(lattice.n_bonds, bool)
    for y in range(L - 1):
        active[2 * y + 1] = True
    assert spans_along(active, lattice, "y") and not wraps_along(active, lattice, "y")
    active[2 * (L - 1) + 1] = True
    assert wraps_along(active, lattice, "y")
    with pytest.raises(ValueError):
        wraps_along(active, lattice, "z")


# ------------------------------------------------ test_wrapping.py:67-120
def _sample(sizes, disorders, temperatures, n_samples, seed=0):
    rng = np.random.default_rng(seed)
    for L in sizes:
        lattice = build_lattice(L)
        for p in disorders:
            J = couplings_from_disorder(*rbim.generate_disorder(L, p, rng))
            for T in temperatures:
                spins = random_spins(lattice.n_sites, rng)
                for _ in range(n_samples):
                    sw_sweep(spins, lattice, J, 1.0 / T, rng, measure=False)
                    yield lattice, activate_bonds(spins, lattice, J, 1.0 / T, rng)


def test_wrapping_matches_brute_force(oracle):
    checked = wrapping = 0
    for lattice, active in _sample((3, 6, 10), (0.0, 0.1, 0.5), (1.0, 2.27, 4.0), 4):
        expected = oracle.brute_force_winding(active, lattice.L)
        _, _, wx, wy = label_and_wrap(active, lattice)
        assert (wx, wy) == expected
        checked += 1
        wrapping += int(any(expected))
    assert checked > 50 and wrapping > 5


def test_wrapping_matches_brute_force_on_larger_lattices(oracle):
    for lattice, active in _sample((16, 24), (0.0, 0.2), (1.2, 2.27), 6, seed=17):
        assert label_and_wrap(active, lattice)[2:] == oracle.brute_force_winding(active, lattice.L)


def test_labels_partition_the_lattice():
    for lattice, active in _sample((6, 10), (0.1,), (1.5,), 3, seed=5):
        n, labels, _, _ = label_and_wrap(active, lattice)
        assert labels.min() == 0 and labels.max() == n - 1
        assert np.bincount(labels, minlength=n).sum() == lattice.n_sites
        ends = lattice.bonds[active]
        assert np.all(labels[ends[:, 0]] == labels[ends[:, 1]])


def test_sw_sweep_updates_in_place_and_reports():
    L = 12
    lattice = build_lattice(L)
    rng = np.random.default_rng(1)
    spins = random_spins(lattice.n_sites, rng)
    before = spins.copy()
    J = ferro(L)
    res = sw_sweep(spins, lattice, J, 1 / 2.0, rng)
    assert spins.dtype == np.int64 and not np.array_equal(spins, before)
    assert res.energy == rbim.swendsen_wang.energy(spins, lattice, J)
    assert res.magnetization == spins.sum()
    assert res.largest_size >= res.second_size and res.wraps_x in (True, False)
    assert sw_sweep(spins, lattice, J, 0.5, rng, measure=False) is None
    quiet = sw_sweep(spins, lattice, J, 0.5, rng, measure_percolation=False)
    assert quiet.wraps_x is None and quiet.wraps_y is None


@pytest.mark.parametrize("engine", [_hip.ENGINE_PLAIN, _hip.ENGINE_BATCHED])
def test_sw_sweep_bit_exact_against_the_oracle(coracle, engine):
    """The shim's one sweep replayed on the oracle: sw_sweep draws its Philox key
    from the caller's generator and runs lattice id 0, sweep 0, so the oracle's
    activation with that key, and its sweep given those bits and flip bits,
    must give the same spins and every field of the SweepResult."""
    from rbim.clusters import activation_threshold
    from rbim.lattice import split_couplings

    for L, T, p, seed in ((12, 2.3, 0.1, 1), (16, 1.5, 0.05, 2), (33, 2.27, 0.1, 3),
                          (64, 3.5, 0.0, 4), (40, 1.2, 0.12, 5)):
        lattice = build_lattice(L)
        rng = np.random.default_rng(seed)
        J_h = np.where(rng.random((L, L)) < p, -1, 1).astype(np.int8)
        J_v = np.where(rng.random((L, L)) < p, -1, 1).astype(np.int8)
        J = couplings_from_disorder(J_h, J_v)
        spins = random_spins(lattice.n_sites, rng)
        replay = np.random.default_rng(seed + 1000)
        key = int(replay.integers(0, 2**64, dtype=np.uint64))
        before = spins.copy()
        res = sw_sweep(spins, lattice, J, 1.0 / T, np.random.default_rng(seed + 1000),
                       engine=engine)
        jh, jv = split_couplings(J)
        active, flips = coracle.activate(before, jh, jv, activation_threshold(1.0 / T), key, 0, 0)
        want, st, _, _ = coracle.sweep_given_bits(before, jh, jv, active, flips)
        case = f"L={L} T={T} engine={engine}"
        assert np.array_equal(spins, want.astype(np.int64)), case
        assert (res.energy, res.magnetization, res.n_clusters) == \
            (st["energy"], st["magnetization"], st["n_clusters"]), case
        assert (res.largest_size, res.largest_abs_magnetization) == \
            (st["largest_size"], st["largest_abs_mag"]), case
        assert (res.second_size, res.second_abs_magnetization) == \
            (st["second_size"], st["second_abs_mag"]), case
        assert (res.wraps_x, res.wraps_y) == (st["wrap_x"], st["wrap_y"]), case


# ---------------------------------------------------- test_physics.py
def _observables(L, T, **kw):
    return derive_observables(run_single_realization(L, ferro(L), T, **kw).primitives, T, L * L)


def test_repeated_runs_with_the_same_seed_agree():
    kw = dict(n_sweeps=200, n_therm=50, n_bins=4, seed=1234)
    a = run_single_realization(8, ferro(8), 2.5, **kw)
    b = run_single_realization(8, ferro(8), 2.5, **kw)
    assert a.primitives == b.primitives


def test_different_seeds_disagree():
    kw = dict(n_sweeps=200, n_therm=50, n_bins=4)
    a = run_single_realization(8, ferro(8), 2.5, seed=1, **kw)
    b = run_single_realization(8, ferro(8), 2.5, seed=2, **kw)
    assert a.primitives["energy"] != pytest.approx(b.primitives["energy"])


def test_binning_does_not_change_central_values():
    reference = None
    for n_bins in (2, 8, 40):
        d = _observables(12, 2.3, n_sweeps=800, n_therm=200, n_bins=n_bins, seed=77)
        if reference is None:
            reference = d
        else:
            for name, value in reference.items():
                assert d[name] == pytest.approx(value, rel=1e-12, abs=1e-12)


def test_percolation_off_leaves_the_rest_unchanged():
    """measure_percolation=False (simulation.py:229-236): the wrap observables
    are NaN; every other observable is exactly what the same seed gives with
    the measurement on (the flag changes no draw)."""
    temps = [1.6, 2.3, 3.0]
    on = run_temperature_scan(12, 0.1, temps, n_realizations=3, n_sweeps=200, n_therm=50,
                              n_bins=4, seed=21)
    off = run_temperature_scan(12, 0.1, temps, n_realizations=3, n_sweeps=200, n_therm=50,
                               n_bins=4, seed=21, measure_percolation=False)
    assert off["metadata"]["measure_percolation"] is False
    for name, entry in off["observables"].items():
        for key in ("value", "error", "thermal_error"):
            if name.startswith("wrap"):
                assert np.all(np.isnan(entry[key])), (name, key)
            else:
                assert np.array_equal(entry[key], on["observables"][name][key]), (name, key)


@pytest.mark.slow
def test_energy_stays_within_bounds():
    for T in (1.0, TC_EXACT, 5.0):
        assert -2.0 <= _observables(12, T, n_sweeps=400, n_therm=200, n_bins=4, seed=5)["energy"] <= 0.0


@pytest.mark.slow
def test_spontaneous_magnetization_matches_onsager():
    d = _observables(16, 2.0, n_sweeps=4000, n_therm=1000, n_bins=10, seed=2024)
    exact = (1.0 - np.sinh(2.0 / 2.0) ** -4) ** 0.125
    assert d["abs_magnetization"] == pytest.approx(exact, abs=0.04)


@pytest.mark.slow
def test_disordered_phase_has_small_magnetization():
    assert _observables(16, 4.0, n_sweeps=2000, n_therm=500, n_bins=10, seed=11)["abs_magnetization"] < 0.15


@pytest.mark.slow
def test_percolation_strength_follows_magnetization_in_the_pure_model():
    d = _observables(16, 2.0, n_sweeps=4000, n_therm=1000, n_bins=10, seed=99)
    assert d["p_infinity"] == pytest.approx(d["abs_magnetization"], abs=0.05)
    assert d["cluster_magnetization_ratio"] > 0.9


@pytest.mark.slow
def test_binder_cumulant_brackets_its_limits():
    kw = dict(n_realizations=1, n_sweeps=3000, n_therm=600, n_bins=10, seed=4, n_workers=1)
    ordered = run_temperature_scan(12, 0.0, [1.2], **kw)
    disordered = run_temperature_scan(12, 0.0, [5.0], **kw)
    assert ordered["observables"]["binder"]["value"][0] == pytest.approx(2.0 / 3.0, abs=0.03)
    assert abs(disordered["observables"]["binder"]["value"][0]) < 0.1


@pytest.mark.parametrize("engine", [0, 1, 3])
def test_allocation_stays_within_footprint(engine):
    """The admission figure covers what a batch really holds, for a ragged mix
    across all layouts (LDS, edge-slot, multi-kernel tiled), with a series ring."""
    from rbim.engine import Batch, LatticeSpec, footprint

    rng = np.random.default_rng(4)
    specs = []
    for k, L in enumerate((16, 32, 100, 128, 160, 256, 320, 336, 384)):
        J_h = np.where(rng.random((L, L)) < 0.1, -1, 1).astype(np.int8)
        specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_h.T.copy(), beta=0.45, seed=3, lattice_id=k))
    need = footprint(specs, n_bins=4, engine=engine, series=16)
    with Batch(specs, n_therm=2, n_bins=4, sweeps_per_bin=3, engine=engine, series=16) as batch:
        batch.run()
        info = batch.info()
        assert batch.series().shape == (len(specs), 12)
    assert 0 < info.device_bytes <= need


def test_auto_engine_choice_follows_the_stated_rule():
    """ENGINE_AUTO puts a side on the LDS kernel up to L = 64, and up to L = 128
    while its lattices hold fewer than 2^23 sites; the edge-slot sweep otherwise.
    rbim.engine.auto_uses_lds states the rule for the autotuner and bench.py:
    the batch's own grouping must agree with it on both sides of each limit."""
    from rbim.engine import Batch, LatticeSpec, auto_uses_lds

    for L, n in ((64, 2100), (80, 8), (80, 1400), (128, 8), (128, 511), (128, 512), (129, 4)):
        J = np.ones((L, L), np.int8)
        specs = [LatticeSpec(L=L, J_h=J, J_v=J, beta=0.4, seed=1, lattice_id=i) for i in range(n)]
        with Batch(specs, n_therm=0, n_bins=1, sweeps_per_bin=1) as batch:
            info = batch.info()
        lds = auto_uses_lds(L, n)
        assert (info.lds_groups > 0, info.tiled_groups > 0) == (lds, not lds), (L, n)
    assert auto_uses_lds(128, 511) and not auto_uses_lds(128, 512) and not auto_uses_lds(129, 1)


def test_descriptor_validation():
    from rbim.engine import Batch, LatticeSpec

    J = np.ones((32, 32), np.int8)
    spec = [LatticeSpec(L=32, J_h=J, J_v=J, beta=0.4, seed=1, lattice_id=0)]
    with pytest.raises(ValueError):
        Batch(spec, n_therm=1, n_bins=1, sweeps_per_bin=1, streams=17)
    with pytest.raises(ValueError):
        Batch(spec, n_therm=1, n_bins=1, sweeps_per_bin=1, series=-1)
    with pytest.raises(ValueError):
        Batch(spec, n_therm=1, n_bins=1, sweeps_per_bin=1, engine=4)
    with Batch(spec, n_therm=1, n_bins=1, sweeps_per_bin=1) as batch:
        batch.run()
        with pytest.raises(RuntimeError):
            batch.series()
