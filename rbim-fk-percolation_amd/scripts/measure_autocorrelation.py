#!/usr/bin/env python3
"""Integrated autocorrelation times of the cluster observables across the
transition, on the GPU engine.

Same flags and table as the reference's scripts/measure_autocorrelation.py:
for each p, every (temperature x realization) lattice runs as one device batch
that keeps the per-sweep stats of all measured sweeps; tau_int is estimated
with Sokal's self-consistent window (rbim.timeseries.integrated_autocorrelation)
and averaged over realizations.  In parentheses: sweeps / (2 tau_int), the
number of effectively independent measurements.

    python rbim-fk-percolation_amd/scripts/measure_autocorrelation.py --L 32 --p 0.05 0.20 --sweeps 20000
"""

from __future__ import annotations

import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from rbim.disorder import generate_disorder  # noqa: E402
from rbim.timeseries import TRACKED, collect_series, integrated_autocorrelation  # noqa: E402


def scan(L, p, temperatures, n_sweeps, n_therm, n_realizations, seed):
    """Print the tau_int table of one disorder strength (reference scan(), :93-130)."""
    print(f"\n=== L = {L}, p = {p} ===")
    header = f"{'T':>6} " + " ".join(f"{name[:16]:>17}" for name in TRACKED)
    print(header)
    print("-" * len(header))
    disorder_seed, run_seed = np.random.SeedSequence(seed).spawn(2)
    rng = np.random.default_rng(disorder_seed)
    couplings = [generate_disorder(L, p, rng) for _ in range(n_realizations)]
    series = collect_series(L, couplings, temperatures, n_sweeps, n_therm, run_seed)
    for t, temperature in enumerate(temperatures):
        row = f"{temperature:6.2f} "
        for name in TRACKED:
            tau = float(np.mean([integrated_autocorrelation(series[name][t, r])[0]
                                 for r in range(n_realizations)]))
            row += f" {tau:7.1f} ({n_sweeps / (2.0 * tau):6.0f})"
        print(row)
    print("  values are tau_int (effectively independent measurements in parentheses)")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--p", type=float, nargs="+", default=[0.05, 0.20])
    ap.add_argument("--temperatures", type=float, nargs="+",
                    default=[2.6, 2.3, 2.1, 1.9, 1.7, 1.5, 1.3, 1.1])
    ap.add_argument("--sweeps", type=int, default=20000)
    ap.add_argument("--therm", type=int, default=2000)
    ap.add_argument("--realizations", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    for p in args.p:
        scan(args.L, p, args.temperatures, args.sweeps, args.therm, args.realizations, args.seed)


if __name__ == "__main__":
    main()
