"""Host-side parts of the drop-in package (no GPU): geometry, disorder, the
estimators, the moment decoding and the results format.

The geometry and disorder checks restate the reference's tests/test_lattice.py;
the estimator checks pin derive_observables / the jackknife / the thermal
error against closed forms on hand-built primitive tables.
"""

import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from rbim.disorder import (
    default_disorder_filename,
    frustrated_plaquettes,
    frustration_density,
    generate_disorder,
    load_disorder,
    save_disorder,
)
from rbim.engine import (
    PRIMITIVE_SLOTS,
    engine_seed,
    moments_to_primitive_arrays,
    moments_to_primitives,
)
from rbim.lattice import build_lattice, couplings_from_disorder, site_index, split_couplings
from rbim.simulation import (
    OBSERVABLES,
    PRIMITIVES,
    RealizationResult,
    _assemble_arrays,
    _combine_realizations,
    derive_observables,
    load_results,
    save_results,
)


# ------------------------------------------------------------ geometry
@pytest.mark.parametrize("L", [3, 4, 8])
def test_bond_count_and_degree(L):
    lattice = build_lattice(L)
    assert lattice.n_sites == L * L and lattice.bonds.shape == (2 * L * L, 2)
    assert np.all(np.bincount(lattice.bonds.reshape(-1), minlength=L * L) == 4)


@pytest.mark.parametrize("L", [3, 5, 8])
def test_wrap_masks(L):
    lattice = build_lattice(L)
    assert lattice.wraps_x.sum() == L and lattice.wraps_y.sum() == L
    assert not np.any(lattice.wraps_x & lattice.wraps_y)
    assert np.all(lattice.bonds[lattice.wraps_x, 0] // L == L - 1)
    assert np.all(lattice.bonds[lattice.wraps_y, 0] % L == L - 1)


def test_geometry_matches_oracle(oracle):
    for L in (3, 4, 7, 16):
        lattice = build_lattice(L)
        bonds, dx, dy, wx, wy = oracle.geometry(L)
        assert np.array_equal(lattice.bonds, bonds)
        assert np.array_equal(lattice.bond_dx, dx) and np.array_equal(lattice.bond_dy, dy)
        assert np.array_equal(lattice.wraps_x, wx) and np.array_equal(lattice.wraps_y, wy)


def test_bond_ordering_matches_coupling_layout():
    L = 4
    lattice = build_lattice(L)
    J_h = np.arange(L * L, dtype=np.int8).reshape(L, L)
    J_v = -np.arange(L * L, dtype=np.int8).reshape(L, L)
    J = couplings_from_disorder(J_h, J_v)
    for x in range(L):
        for y in range(L):
            n = site_index(x, y, L)
            assert J[2 * n] == J_h[x, y] and J[2 * n + 1] == J_v[x, y]
            assert lattice.bonds[2 * n, 0] == n and lattice.bonds[2 * n + 1, 0] == n
    back_h, back_v = split_couplings(J)
    assert np.array_equal(back_h, J_h) and np.array_equal(back_v, J_v)


def test_small_lattice_and_bad_couplings_rejected():
    with pytest.raises(ValueError):
        build_lattice(2)
    with pytest.raises(ValueError):
        couplings_from_disorder(np.ones((3, 3), np.int8), np.ones((4, 4), np.int8))
    with pytest.raises(ValueError):
        build_lattice(4).wrap_mask("z")


# ------------------------------------------------------------- disorder
def test_disorder_draws_match_the_reference_sequence():
    """Same generator calls as disorder.py:59-60: J_h then J_v from one rng."""
    a = generate_disorder(16, 0.1, np.random.default_rng(3))
    rng = np.random.default_rng(3)
    u_h, u_v = rng.random((16, 16)), rng.random((16, 16))
    assert np.array_equal(a[0], np.where(u_h < 0.1, -1, 1))
    assert np.array_equal(a[1], np.where(u_v < 0.1, -1, 1))
    assert a[0].dtype == np.int8
    with pytest.raises(ValueError):
        generate_disorder(4, 1.5)


def test_frustration():
    L = 6
    ones = np.ones((L, L), np.int8)
    assert frustration_density(ones, ones) == 0.0
    J_h = ones.copy()
    J_h[2, 3] = -1
    f = frustrated_plaquettes(J_h, ones)
    assert f.sum() == 2 and f[2, 3] and f[2, 2]
    p = 0.3
    J_h, J_v = generate_disorder(128, p, np.random.default_rng(12345))
    assert abs(frustration_density(J_h, J_v) - 4 * p * (1 - p) * ((1 - p) ** 2 + p**2)) < 0.01


def test_disorder_file_roundtrip(tmp_path):
    J_h, J_v = generate_disorder(8, 0.2, np.random.default_rng(1))
    path = tmp_path / default_disorder_filename(8, 0.2, 3)
    assert path.name == "disorder_L=8_p=0.200_r=0003.npz"
    save_disorder(str(path), J_h, J_v, p=0.2, seed=9)
    got_h, got_v = load_disorder(str(path))
    assert np.array_equal(got_h, J_h) and np.array_equal(got_v, J_v) and got_h.dtype == np.int8
    with np.load(str(path)) as data:
        assert float(data["p"]) == 0.2 and int(data["seed"]) == 9
    bad = tmp_path / "bad.npz"
    np.savez(str(bad), A=np.zeros(3))
    with pytest.raises(KeyError):
        load_disorder(str(bad))


# ----------------------------------------------------------- estimators
def _prims(**over):
    base = {name: 0.0 for name in PRIMITIVES}
    base.update(over)
    return base


def test_derive_observables_closed_forms():
    N, T = 16, 2.0
    P = _prims(energy=-20.0, energy_sq=420.0, magnetization=4.0, abs_magnetization=10.0,
               magnetization_sq=120.0, magnetization_4=20000.0, largest_size=12.0,
               largest_abs_mag=9.0, second_size=3.0, mag_ratio=0.75, n_clusters=5.0,
               wrap_x=0.25, wrap_y=0.5, wrap_any=0.6)
    d = derive_observables(P, T, N)
    assert d["energy"] == -20.0 / 16
    assert d["specific_heat"] == pytest.approx((420.0 - 400.0) / (4.0 * 16))
    assert d["susceptibility"] == pytest.approx((120.0 - 100.0) / (2.0 * 16))
    m2, m4 = 120.0 / 256, 20000.0 / 65536
    assert d["binder"] == pytest.approx(1 - m4 / (3 * m2**2))
    assert d["p_infinity"] == 12.0 / 16 and d["largest_size"] == 12.0
    assert d["cluster_density_gap"] == (12.0 - 3.0) / 16
    assert d["n_clusters_per_site"] == 5.0 / 16
    assert (d["wrap_x"], d["wrap_y"], d["wrap_any"]) == (0.25, 0.5, 0.6)
    assert derive_observables(_prims(), T, N)["binder"] == 0.0
    assert set(d) == set(OBSERVABLES)


def _realizations(table, bins=None, n_bins=2):
    out = []
    for i, row in enumerate(table):
        prims = _prims(**row)
        b = {name: np.full(n_bins, prims[name]) for name in PRIMITIVES}
        if bins is not None:
            for name, vals in bins[i].items():
                b[name] = np.asarray(vals, float)
        out.append(RealizationResult(primitives=prims, bin_primitives=b))
    return out


def test_jackknife_of_linear_observables_is_the_standard_error():
    x = np.array([-30.0, -26.0, -31.0, -25.0, -29.5])
    res = _realizations([{"energy": v, "largest_size": -v / 2} for v in x])
    c = _combine_realizations(res, 2.0, 16)
    sem = np.sqrt(np.sum((x - x.mean()) ** 2) / (len(x) * (len(x) - 1)))
    assert c["energy"]["value"] == pytest.approx(x.mean() / 16)
    assert c["energy"]["error"] == pytest.approx(sem / 16, rel=1e-12)
    assert c["p_infinity"]["error"] == pytest.approx(sem / 2 / 16, rel=1e-12)


def test_jackknife_of_binder_by_brute_force():
    rows = [{"magnetization_sq": m2, "magnetization_4": m4}
            for m2, m4 in [(100.0, 15000.0), (80.0, 9000.0), (120.0, 20000.0)]]
    c = _combine_realizations(_realizations(rows), 2.0, 16)

    def binder(m2, m4):
        return 1 - (m4 / 16**4) / (3 * (m2 / 16**2) ** 2)

    m2 = np.array([r["magnetization_sq"] for r in rows])
    m4 = np.array([r["magnetization_4"] for r in rows])
    loo = np.array([binder(np.delete(m2, i).mean(), np.delete(m4, i).mean()) for i in range(3)])
    assert c["binder"]["value"] == pytest.approx(binder(m2.mean(), m4.mean()))
    assert c["binder"]["error"] == pytest.approx(np.sqrt(2 / 3 * np.sum((loo - loo.mean()) ** 2)))


def test_thermal_error_and_single_realization():
    bins = [{"energy": [-10.0, -12.0, -11.0, -13.0]}, {"energy": [-9.0, -9.5, -10.0, -8.5]}]
    res = _realizations([{"energy": -11.5}, {"energy": -9.25}], bins=bins, n_bins=4)
    c = _combine_realizations(res, 2.0, 4)
    spreads = [np.std(np.array(b["energy"]) / 4, ddof=1) / 2 for b in bins]
    assert c["energy"]["thermal_error"] == pytest.approx(np.sqrt(np.mean(np.square(spreads)) / 2))
    one = _combine_realizations(res[:1], 2.0, 4)
    assert one["energy"]["error"] == one["energy"]["thermal_error"]
    assert one["energy"]["error"] == pytest.approx(spreads[0])


# ------------------------------------------------------ moment decoding
def test_moments_to_primitives_exact_means():
    acc = np.zeros((2, 20), np.uint64)
    rows = [(-30, 900 + 2**64 * 3, 6, 6, 2**63, 7, 0.5), (-10, 100, -2, 2, 4, 16, 0.25)]
    for b, (e, e2, m, am, m2, m4, ratio) in enumerate(rows):
        acc[b, 0] = np.array([e], np.int64).view(np.uint64)[0]
        acc[b, 1], acc[b, 2] = e2 & (2**64 - 1), e2 >> 64
        acc[b, 3] = np.array([m], np.int64).view(np.uint64)[0]
        acc[b, 4] = am
        acc[b, 5], acc[b, 6] = m2 & (2**64 - 1), m2 >> 64
        acc[b, 7], acc[b, 8] = m4 & (2**64 - 1), m4 >> 64
        acc[b, 12] = np.array([ratio], np.float64).view(np.uint64)[0]
        acc[b, 17] = 2
    prims, bins = moments_to_primitives(acc)
    assert prims["energy"] == (-30 - 10) / 4
    assert prims["energy_sq"] == (900 + 2**64 * 3 + 100) / 4
    assert prims["magnetization"] == (6 - 2) / 4
    assert prims["magnetization_sq"] == (2**63 + 4) / 4
    assert prims["mag_ratio"] == 0.75 / 4
    assert np.array_equal(bins["energy"], [-15.0, -5.0])
    assert set(prims) == set(PRIMITIVES) == set(PRIMITIVE_SLOTS)
    nopc, _ = moments_to_primitives(acc, measure_percolation=False)
    assert np.isnan(nopc["wrap_x"]) and np.isnan(nopc["wrap_any"])


def test_moment_arrays_exact_beyond_float_range():
    """Batch decoding: sums below 2^53 take the float path, larger ones (here
    M^4 sums past 2^64) the exact integer path; every mean is the correctly
    rounded quotient, i.e. Python's int / int, lattice by lattice."""
    rng = np.random.default_rng(5)
    n, nb = 3, 4
    acc = np.zeros((n, nb, 20), np.uint64)
    big = []
    for i in range(n):
        for b in range(nb):
            acc[i, b, 17] = 1 + b
            acc[i, b, 0] = np.array([int(rng.integers(-2**40, 2**40))], np.int64).view(np.uint64)[0]
            m4 = int(rng.integers(1, 2**62)) * (3**20 if i == 1 else 1)  # lattice 1 past 2^64
            acc[i, b, 7], acc[i, b, 8] = m4 & (2**64 - 1), m4 >> 64
            big.append(m4)
    prims, bins = moments_to_primitive_arrays(acc)
    j4 = PRIMITIVES.index("magnetization_4")
    for i in range(n):
        m4s = big[i * nb:(i + 1) * nb]
        assert prims[i, j4] == sum(m4s) / sum(range(1, nb + 1))
        assert list(bins[i, :, j4]) == [m / (b + 1) for b, m in enumerate(m4s)]
        one_p, one_b = moments_to_primitives(acc[i])
        for j, name in enumerate(PRIMITIVES):
            assert one_p[name] == prims[i, j] and np.array_equal(one_b[name], bins[i, :, j])
    assert prims.shape == (n, len(PRIMITIVES)) and bins.shape == (n, nb, len(PRIMITIVES))
    with pytest.raises(ZeroDivisionError):
        moments_to_primitive_arrays(np.zeros((1, 2, 20), np.uint64))


def test_scan_estimators_equal_per_temperature_combination():
    """The whole-scan estimator (all temperatures at once) gives, at every
    temperature, what the one-temperature combination of the same
    realizations gives, for R = 1 (error := thermal error) and R > 1."""
    rng = np.random.default_rng(9)
    for R, nb in ((1, 5), (6, 1), (5, 8)):
        temps = [1.5, 2.0, 2.7]
        prims = rng.normal(size=(len(temps) * R, len(PRIMITIVES))) + 3.0
        bins = rng.normal(size=(len(temps) * R, nb, len(PRIMITIVES))) + 3.0
        res = _assemble_arrays(8, temps, prims, bins, R, {})
        for t, T in enumerate(temps):
            raw = [RealizationResult(primitives={k: prims[t * R + r, j] for j, k in enumerate(PRIMITIVES)},
                                     bin_primitives={k: bins[t * R + r, :, j] for j, k in enumerate(PRIMITIVES)})
                   for r in range(R)]
            one = _combine_realizations(raw, T, 64)
            for name in OBSERVABLES:
                for key in ("value", "error", "thermal_error"):
                    a, b = res["observables"][name][key][t], one[name][key]
                    assert a == b or (np.isnan(a) and np.isnan(b)), (R, nb, name, key, a, b)


def test_engine_seed_is_deterministic():
    assert engine_seed(5) == engine_seed(np.random.SeedSequence(5))
    assert engine_seed(5) != engine_seed(6)
    assert 0 <= engine_seed(0) < 2**64


# ------------------------------------------------------- results format
def test_results_roundtrip_and_reference_fixture(tmp_path):
    ref = load_results(os.path.join(GOLDEN, "survey_L16_p0.00.npz"))
    assert ref["metadata"]["L"] == 16 and len(ref["temperatures"]) == 24
    assert set(ref["observables"]) == set(OBSERVABLES)
    for entry in ref["observables"].values():
        assert set(entry) == {"value", "error", "thermal_error"}
    path = tmp_path / "out.npz"
    save_results(str(path), ref)
    again = load_results(str(path))
    assert again["metadata"] == ref["metadata"]
    for name in OBSERVABLES:
        for key in ("value", "error", "thermal_error"):
            assert np.array_equal(again["observables"][name][key], ref["observables"][name][key])
    with np.load(str(path), allow_pickle=False) as data:
        assert len(data.files) == 2 + 3 * len(OBSERVABLES)
        json.loads(str(data["metadata_json"]))


def test_integrated_autocorrelation_ar1_closed_form():
    """AR(1) x_t = r x_{t-1} + e_t has rho(t) = r^t, so tau_int = (1+r)/(2(1-r));
    white noise gives 1/2, a constant series (0.5, 0)."""
    from rbim.timeseries import integrated_autocorrelation

    rng = np.random.default_rng(11)
    for r in (0.0, 0.5, 0.9):
        n = 400_000
        e = rng.standard_normal(n)
        x = np.empty(n)
        x[0] = e[0] / np.sqrt(1 - r * r)
        for t in range(1, n):
            x[t] = r * x[t - 1] + e[t]
        tau, window = integrated_autocorrelation(x)
        exact = (1 + r) / (2 * (1 - r))
        assert abs(tau - exact) < 0.05 * exact + 0.02, (r, tau, exact)
        assert window >= 6 * tau - 1
    assert integrated_autocorrelation(np.ones(100)) == (0.5, 0)
    assert integrated_autocorrelation(np.arange(10.0)) == (0.5, 0)


def test_crossing_interpolation():
    """Linear crossing of two lines and of sampled curves with a known
    intersection; no crossing -> NaN; a grid point where they meet; the first
    of two crossings wins; a meeting at the last grid point is not an event;
    all size pairs at once equal pair by pair."""
    from rbim.crossings import first_crossing, mean_crossing, size_pair_crossings

    t = np.linspace(2.0, 2.5, 11)
    assert abs(first_crossing(t, 2.0 * (t - 2.2691), -0.5 * (t - 2.2691)) - 2.2691) < 1e-12
    assert np.isnan(first_crossing(t, t, t + 1.0))
    assert first_crossing(t, np.where(t == t[3], 0.0, t - t[3]), np.zeros_like(t)) == t[3]
    # (t - 2.12)(t - 2.38): zeros at 2.12 and 2.38 -> the first; the secant through
    # the bracketing grid points 2.10 (+0.0056) and 2.15 (-0.0069) crosses at
    # 2.10 + 0.05 * 0.0056 / 0.0125
    quad = (t - 2.12) * (t - 2.38)
    assert abs(first_crossing(t, quad, 0.0 * t) - (2.10 + 0.05 * 0.0056 / 0.0125)) < 1e-12
    assert np.isnan(first_crossing(t, t - t[-1], 0.0 * t))
    curves = {16: 1.0 - 2.0 * (t - 2.27), 8: 1.0 - (t - 2.27), 32: 1.0 - 4.0 * (t - 2.27),
              64: 5.0 + 0.0 * t}
    found = size_pair_crossings(t, curves)
    assert [(int(a), int(b)) for a, b, _ in found] == [(8, 16), (8, 32), (16, 32)]
    assert all(abs(x - 2.27) < 1e-12 for x in found["T"])
    assert abs(mean_crossing(found) - 2.27) < 1e-12
    rng = np.random.default_rng(5)
    noisy = {L: rng.normal(size=t.size) for L in (4, 8, 12, 20)}
    pairs = size_pair_crossings(t, noisy)
    for a, b, x in pairs:
        assert first_crossing(t, noisy[int(a)], noisy[int(b)]) == x
    assert mean_crossing(size_pair_crossings(t, {8: t})) is None


def test_agreement_rule_on_synthetic_scores():
    from agreement import agreement

    """The acceptance rule itself: N(0, 1) scores pass nearly always, a shift
    of one observable, a few wild points or inflated scatter fail."""
    rng = np.random.default_rng(8)
    names = [f"o{i}" for i in range(16)]

    def rows(z):
        return [(names[i % 16], float(i // 16), 0.0, 0.0, float(v)) for i, v in enumerate(z)]

    passes = sum(all(p for p, _, _ in agreement(rows(rng.normal(size=384))).values())
                 for _ in range(200))
    assert passes >= 198
    shifted = rng.normal(size=384)
    shifted[::16] += 2.5  # one observable off by 2.5 sigma at every temperature
    assert not agreement(rows(shifted))["chi2"][0]
    wild = rng.normal(size=384)
    wild[:7] = 3.4
    assert not agreement(rows(wild))["beyond3"][0]
    assert not agreement(rows(np.concatenate([rng.normal(size=383), [6.0]])))["max"][0]
