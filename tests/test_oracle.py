"""Pins the oracle (oracle/sw_oracle.py and oracle/sw_oracle.c) before it is
trusted as the checker of the GPU engine.

Each restated function is checked against something that does not share its
code path: published known-answer vectors, an independent library routine,
a brute-force method, a closed form, or the known answers of the reference's
own tests.
"""

import json
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from conftest import GOLDEN, random_disorder
from samples import column_case, diagonal_loop, double_winding, sample_configurations

# Random123 kat_vectors, philox4x32 with 10 rounds
RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


# ------------------------------------------------------------------ Philox
@pytest.mark.parametrize("ctr,key,expected", RANDOM123_KAT)
def test_philox_random123_kat(oracle, coracle, ctr, key, expected):
    assert tuple(int(v) for v in oracle.philox4x32_10([ctr], key)[0]) == expected
    import ctypes

    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    out = (ctypes.c_uint32 * 4)()
    coracle.load_c().oracle_philox4x32_10(c, k, out)
    assert tuple(out) == expected


def test_philox_matches_rocrand_engine(oracle):
    with open(os.path.join(GOLDEN, "philox_vectors.json")) as fh:
        vectors = json.load(fh)["vectors"]
    assert len(vectors) >= 9
    for v in vectors:
        got = [int(x) for x in oracle.philox4x32_10([v["counter"]], v["key"])[0]]
        assert got == v["output"], v


def test_init_spins_python_matches_c(oracle, coracle):
    for L, seed, lid in [(3, 1, 0), (16, 2**40 + 7, 5), (33, 99, 123456)]:
        a = oracle.init_spins(L, seed, lid, "hot")
        b = coracle.c_init_spins(L, seed, lid, "hot")
        assert np.array_equal(a, b)
        assert set(np.unique(a)) <= {-1, 1}
    assert np.all(coracle.c_init_spins(8, 1, 1, "cold") == 1)


def test_threshold_limits(oracle):
    # p_add limits of test_clusters.py:24-27 carried to the integer threshold
    assert oracle.threshold31(0.0) == 0
    assert oracle.threshold31(100.0) == 2**31
    assert oracle.threshold31(0.5) == int(np.floor((1 - np.exp(-1.0)) * 2**31 + 0.5))


# ------------------------------------------------------------- activation
def test_activation_python_matches_c_and_only_satisfied(oracle, coracle):
    rng = np.random.default_rng(7)
    for L in (3, 5, 12, 17):
        J_h, J_v = random_disorder(rng, L, 0.3)
        spins = np.where(rng.random(L * L) < 0.5, -1, 1).astype(np.int8)
        for beta in (0.0, 0.6, 50.0):
            thr = oracle.threshold31(beta)
            a_py, f_py = oracle.activate(spins, J_h, J_v, thr, 1234, 3, 17)
            a_c, f_c = coracle.c_activate(spins, J_h, J_v, thr, 1234, 3, 17)
            assert np.array_equal(a_py, a_c)
            assert np.array_equal(f_py, f_c)
            bonds = oracle.geometry(L)[0]
            J = np.stack([J_h.reshape(-1), J_v.reshape(-1)], 1).reshape(-1).astype(np.int64)
            sat = J * spins[bonds[:, 0]] * spins[bonds[:, 1]] > 0
            assert not np.any(a_c & ~sat)  # test_clusters.py:62-73
            if beta == 0.0:
                assert not a_c.any()  # test_clusters.py:30-43
            if beta == 50.0:
                assert np.array_equal(a_c, sat)


def test_activation_rate_matches_probability(oracle, coracle):
    L = 64
    J_h = np.ones((L, L), np.int8)
    J_v = np.ones((L, L), np.int8)
    spins = np.ones(L * L, np.int8)
    beta = 0.3
    p_add = -np.expm1(-2 * beta)
    hits = 0
    for sweep in range(8):
        a, _ = coracle.c_activate(spins, J_h, J_v, oracle.threshold31(beta), 99, 0, sweep)
        hits += int(a.sum())
    n = 8 * 2 * L * L
    assert abs(hits / n - p_add) < 4 * np.sqrt(p_add * (1 - p_add) / n)


# ------------------------------------------------------------- union-find
def _scipy_canonical(active, L, oracle):
    bonds = oracle.geometry(L)[0][np.asarray(active, bool)]
    g = csr_matrix((np.ones(len(bonds), np.int8), (bonds[:, 0], bonds[:, 1])), shape=(L * L, L * L))
    n, lab = connected_components(g, directed=False)
    return n, oracle.canonical(lab)


def test_union_find_restatements_agree_with_independent_methods(oracle, coracle):
    checked = wrapped = 0
    for s in sample_configurations(oracle, (3, 4, 5, 6, 10), (0.0, 0.1, 0.5), (1.0, 2.27, 4.0), 3,
                                   seed=11):
        L, active = s["L"], s["active"]
        n_py, lab_py, wx_py, wy_py = oracle.label_and_wrap(active, L)
        n_c, lab_c, wx_c, wy_c = coracle.c_label_and_wrap(active, L)
        assert n_py == n_c and np.array_equal(lab_py, lab_c) and (wx_py, wy_py) == (wx_c, wy_c)
        # labels: unique min-site-ordered numbering of scipy's components
        n_sp, canon_sp = _scipy_canonical(active, L, oracle)
        assert n_sp == n_c
        assert np.array_equal(oracle.canonical(lab_c), canon_sp)
        first = np.full(n_c, -1)
        for site, lab in enumerate(lab_c):
            if first[lab] < 0:
                first[lab] = site
        assert np.all(np.diff(first) > 0)  # numbered in increasing min-site order
        # wraps: the reference test's BFS oracle and the covering-space check
        assert oracle.brute_force_winding(active, L) == (wx_c, wy_c)
        assert oracle.cover_winding(active, L) == (wx_c, wy_c)
        checked += 1
        wrapped += int(wx_c or wy_c)
    assert checked > 100 and wrapped > 10


def test_union_find_known_answers(oracle, coracle):
    L = 8
    none = np.zeros(2 * L * L, bool)
    n, lab, wx, wy = coracle.c_label_and_wrap(none, L)
    assert n == L * L and not wx and not wy
    n, lab, wx, wy = coracle.c_label_and_wrap(~none, L)
    assert n == 1 and wx and wy
    L = 6
    open_col, closed_col = column_case(L)
    assert coracle.c_label_and_wrap(open_col, L)[2:] == (False, False)
    assert coracle.c_label_and_wrap(closed_col, L)[2:] == (False, True)
    for L in (3, 4, 7):
        assert coracle.c_label_and_wrap(diagonal_loop(L), L)[2:] == (True, True)
        assert oracle.brute_force_winding(diagonal_loop(L), L) == (True, True)
    assert coracle.c_label_and_wrap(double_winding(6), 6)[2:] == (False, True)


# --------------------------------------------------- sweep given the bits
def test_sweep_given_bits_c_matches_numpy(oracle, coracle):
    n = 0
    for s in sample_configurations(oracle, (3, 6, 9, 16), (0.0, 0.15), (1.2, 2.27, 3.5), 2, seed=3):
        sp_np, st_np, lab_np, _ = oracle.sweep_given_bits(s["spins"], s["J_h"], s["J_v"],
                                                           s["active"], s["flips"])
        sp_c, st_c, lab_c = coracle.c_sweep_given_bits(s["spins"], s["J_h"], s["J_v"],
                                                        s["active"], s["flips"])
        assert np.array_equal(sp_np, sp_c)
        assert np.array_equal(lab_np, lab_c)
        assert st_np == st_c
        # invariants of test_clusters.py:76-91 and the flip rule
        L = s["L"]
        sizes = np.bincount(lab_c)
        mags = np.bincount(lab_c, weights=sp_c.astype(float))
        assert sizes.sum() == L * L and mags.sum() == sp_c.sum()
        assert np.all(np.abs(mags) <= sizes)
        assert st_c["largest_size"] == sizes.max()
        assert st_c["magnetization"] == int(sp_c.sum())
        canon = oracle.canonical(lab_c)
        flipped = s["flips"][canon].astype(bool)
        assert np.array_equal(sp_c, np.where(flipped, -s["spins"], s["spins"]))
        n += 1
    assert n > 20


def test_run_composes_activation_and_sweeps(oracle, coracle):
    """oracle_run == sweep-by-sweep activation (Philox sweep index g) + update,
    with unmeasured thermalisation and per-bin sums."""
    rng = np.random.default_rng(5)
    L, key, lid = 10, 0xABCDEF12345, 7
    J_h, J_v = random_disorder(rng, L, 0.1)
    thr = oracle.threshold31(1 / 2.1)
    s0 = coracle.c_init_spins(L, key, lid)
    n_therm, spb, n_bins = 3, 4, 3
    s_run, acc = coracle.c_run(L, s0, J_h, J_v, thr, key, lid, 0, n_therm + spb * n_bins,
                               n_therm, spb, n_bins)
    s = s0.copy()
    energies = np.zeros(n_bins, dtype=np.int64)
    for g in range(n_therm + spb * n_bins):
        active, flips = coracle.c_activate(s, J_h, J_v, thr, key, lid, g)
        s, st, _ = coracle.c_sweep_given_bits(s, J_h, J_v, active, flips)
        if g >= n_therm:
            energies[(g - n_therm) // spb] += st["energy"]
    assert np.array_equal(s, s_run)
    assert np.array_equal(acc[:, 0].view(np.int64), energies)
    assert np.all(acc[:, 17] == spb)
    # a run split in two calls accumulates the same sums
    s_a, acc_a = coracle.c_run(L, s0, J_h, J_v, thr, key, lid, 0, 7, n_therm, spb, n_bins)
    s_b, acc_b = coracle.c_run(L, s_a, J_h, J_v, thr, key, lid, 7, 100, n_therm, spb, n_bins, acc_a)
    assert np.array_equal(s_b, s_run) and np.array_equal(acc_b, acc)


# ----------------------------------------------------------- exact energy
@pytest.mark.parametrize("L,T,expected", [
    (16, 2.0, -1.7455307), (16, 2.269, -1.4533425), (16, 2.5, -1.1313180),
    (32, 2.269, -1.4340006), (32, 2.5, -1.1072917),
    (256, 2.269, -1.4171780), (1024, 2.269, -1.4154795),
])
def test_kaufman_energy_table(oracle, L, T, expected):
    assert oracle.kaufman_energy(L, T) == pytest.approx(expected, abs=2e-7)


def test_kaufman_tends_to_onsager(oracle):
    for T in (1.5, 2.0, 3.0, 4.0):
        assert oracle.kaufman_energy(128, T) == pytest.approx(oracle.onsager_energy(T), abs=1e-7)
    # high-temperature series: E/N = -2 tanh(K) - O(K^3) ... at K -> 0 the energy vanishes
    assert abs(oracle.kaufman_energy(64, 200.0)) < 0.011
