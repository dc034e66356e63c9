"""Per-sweep time series (device stats ring): bit-exact against the oracle's
sweep-by-sweep stats, consistent with the moment sums, ring order on
wrap-around, identical across engines, and the autocorrelation / equilibration
helpers and script on top of it."""

import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, random_disorder

pytestmark = pytest.mark.gpu

PLAIN, BATCHED = 1, 2


def _specs(Ls, seed=5, p=0.1, temps=(1.5, 2.27, 3.0)):
    from rbim.engine import LatticeSpec

    rng = np.random.default_rng(seed)
    specs, k = [], 0
    for L in Ls:
        J_h, J_v = random_disorder(rng, L, p)
        for T in temps:
            specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / T, seed=977 + L, lattice_id=k))
            k += 1
    return specs


def _oracle_series(coracle, spec, n_therm, n_measured):
    """Stats of each measured sweep from one-sweep oracle runs (a one-sweep bin
    holds that sweep's values)."""
    from rbim.clusters import activation_threshold

    s = coracle.c_init_spins(spec.L, spec.seed, spec.lattice_id, "hot")
    thr = activation_threshold(spec.beta)
    rows = []
    for g in range(n_therm + n_measured):
        s, acc = coracle.c_run(spec.L, s, spec.J_h, spec.J_v, thr, spec.seed, spec.lattice_id, g, 1,
                               g if g >= n_therm else g + 1, 1, 1)
        if g >= n_therm:
            a = acc[0]
            rows.append({
                "energy": int(a[0:1].view(np.int64)[0]), "magnetization": int(a[3:4].view(np.int64)[0]),
                "largest_size": int(a[9]), "largest_abs_mag": int(a[10]), "second_size": int(a[11]),
                "n_clusters": int(a[13]), "wrap_x": int(a[14]), "wrap_y": int(a[15]),
            })
    return rows


@pytest.mark.parametrize("engine,Ls", [(BATCHED, (5, 16, 32)), (PLAIN, (16, 70))])
def test_series_bit_exact_against_oracle(coracle, engine, Ls):
    from rbim.engine import Batch

    specs = _specs(Ls)
    n_therm, n_meas = 4, 9
    with Batch(specs, n_therm=n_therm, n_bins=3, sweeps_per_bin=3, engine=engine, series=n_meas,
               sweeps_per_launch=4) as batch:
        batch.run(6)
        batch.run()
        got = batch.series()
    assert got.shape == (len(specs), n_meas)
    for i, spec in enumerate(specs):
        ref = _oracle_series(coracle, spec, n_therm, n_meas)
        for j, row in enumerate(ref):
            for name, value in row.items():
                assert int(got[i, j][name]) == value, (spec.L, j, name)


def test_series_sums_equal_moment_sums_and_ring_order():
    from rbim.engine import Batch

    specs = _specs((24, 160), temps=(2.0, 2.6))
    n_therm, spb, n_bins = 10, 7, 4
    measured = spb * n_bins
    full = {}
    for cap in (measured, 5):
        with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, series=cap) as batch:
            batch.run()
            full[cap] = batch.series()
            acc = batch.moments()
    s = full[measured]
    assert np.array_equal(full[5], s[:, -5:])  # the ring keeps the newest, oldest first
    for i in range(len(specs)):
        assert int(acc[i, :, 0].view(np.int64).sum()) == int(s[i]["energy"].sum())
        assert int(acc[i, :, 3].view(np.int64).sum()) == int(s[i]["magnetization"].sum())
        assert int(acc[i, :, 9].sum()) == int(s[i]["largest_size"].sum())
        assert int(acc[i, :, 13].sum()) == int(s[i]["n_clusters"].sum())
        assert int(acc[i, :, 15].sum()) == int(s[i]["wrap_y"].sum())


def test_collect_series_engines_agree_and_tau():
    from rbim.timeseries import TRACKED, collect_series, integrated_autocorrelation

    rng = np.random.default_rng(3)
    couplings = [random_disorder(rng, 32, 0.05) for _ in range(2)]
    temps = [1.5, 2.3, 3.0]
    a = collect_series(32, couplings, temps, 400, 50, seed=1, engine=BATCHED)
    b = collect_series(32, couplings, temps, 400, 50, seed=1, engine=PLAIN)
    for name in TRACKED:
        assert a[name].shape == (3, 2, 400)
        assert np.array_equal(a[name], b[name]), name
    tau = [integrated_autocorrelation(a["energy"][t, 0])[0] for t in range(3)]
    assert all(0.5 <= x < 50 for x in tau)


def test_equilibration_check_and_script():
    from rbim.timeseries import equilibration_check

    rng = np.random.default_rng(9)
    J_h, J_v = random_disorder(rng, 32, 0.05)
    out = equilibration_check(32, J_h, J_v, 2.0, 2000, n_blocks=10, seed=4)
    assert out["hot"].shape == (10,) and np.all(np.isfinite(out["z"]))
    # far from the start both starts sample the same distribution
    assert np.max(np.abs(out["z"][5:])) < 6.0
    script = os.path.join(ROOT, "rbim-fk-percolation_amd", "scripts", "measure_autocorrelation.py")
    res = subprocess.run([sys.executable, script, "--L", "16", "--p", "0.05", "--temperatures", "2.6",
                          "1.5", "--sweeps", "500", "--therm", "50", "--realizations", "2"],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert "tau_int" in res.stdout and "2.60" in res.stdout


def test_device_moments_match_reference_float_means():
    """The reference forms 14 float64 series per lattice and takes np.mean over
    the run and over each bin (simulation.py:206-241).  Rebuilt from the
    device's per-sweep ring, those means agree with the device's bin sums
    (exact integers and one FP64 sum) to 1e-12 relative, for every primitive."""
    from rbim.engine import PRIMITIVES, Batch, moments_to_primitives

    specs = _specs((16, 32, 64), temps=(1.6, 2.3, 3.0))
    n_therm, spb, n_bins = 30, 40, 5
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, series=spb * n_bins) as b:
        b.run()
        acc = b.moments()
        ser = b.series()
    for i in range(len(specs)):
        st = ser[i]
        E = st["energy"].astype(np.float64)
        M = st["magnetization"].astype(np.float64)
        S = st["largest_size"].astype(np.float64)
        MC = st["largest_abs_mag"].astype(np.float64)
        series = {
            "energy": E, "energy_sq": E * E, "magnetization": M, "abs_magnetization": np.abs(M),
            "magnetization_sq": M ** 2, "magnetization_4": M ** 4, "largest_size": S,
            "largest_abs_mag": MC, "second_size": st["second_size"].astype(np.float64),
            "mag_ratio": np.divide(MC, S, out=np.zeros_like(S), where=S > 0),
            "n_clusters": st["n_clusters"].astype(np.float64),
            "wrap_x": st["wrap_x"].astype(np.float64), "wrap_y": st["wrap_y"].astype(np.float64),
            "wrap_any": ((st["wrap_x"] != 0) | (st["wrap_y"] != 0)).astype(np.float64),
        }
        prim, bins = moments_to_primitives(acc[i])
        for name in PRIMITIVES:
            ref = float(np.mean(series[name]))
            ref_bins = series[name].reshape(n_bins, spb).mean(axis=1)
            assert prim[name] == pytest.approx(ref, rel=1e-12, abs=1e-12), (i, name)
            np.testing.assert_allclose(bins[name], ref_bins, rtol=1e-12, atol=1e-12, err_msg=name)
