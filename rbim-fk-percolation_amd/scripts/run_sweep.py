#!/usr/bin/env python3
"""Run a temperature scan of the 2D +-J random-bond Ising model on the GPU.

Drop-in for the reference's scripts/run_sweep.py (same flags, same output
file).  Example:

    python rbim-fk-percolation_amd/scripts/run_sweep.py --L 32 --p 0.05 \\
        --t-min 1.0 --t-max 3.0 --n-temps 40 --realizations 32 --sweeps 20000 \\
        --out results/L32_p0.05.npz
"""

from __future__ import annotations

import argparse
import glob
import os
import sys
import time
from typing import List, Tuple

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from rbim.disorder import load_disorder  # noqa: E402
from rbim.simulation import run_temperature_scan, save_results  # noqa: E402


def load_disorder_directory(pattern: str, n_realizations: int) -> List[Tuple[np.ndarray, np.ndarray]]:
    """The first ``n_realizations`` files of the sorted glob (run_sweep.py:23-31)."""
    paths = sorted(glob.glob(pattern))
    if not paths:
        raise FileNotFoundError(f"no disorder files matched {pattern!r}")
    if len(paths) < n_realizations:
        raise ValueError(
            f"{len(paths)} disorder files matched {pattern!r} but --realizations={n_realizations}"
        )
    return [load_disorder(path) for path in paths[:n_realizations]]


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--L", type=int, required=True, help="linear system size")
    ap.add_argument("--p", type=float, required=True, help="antiferromagnetic bond concentration")
    ap.add_argument("--t-min", type=float, default=1.0, help="lowest temperature")
    ap.add_argument("--t-max", type=float, default=3.0, help="highest temperature")
    ap.add_argument("--n-temps", type=int, default=30, help="number of temperatures")
    ap.add_argument("--realizations", type=int, default=1, help="number of disorder realizations")
    ap.add_argument("--sweeps", type=int, default=10_000, help="measured sweeps per realization")
    ap.add_argument("--therm", type=int, default=None, help="thermalisation sweeps (default: sweeps // 10)")
    ap.add_argument("--bins", type=int, default=10, help="number of bins")
    ap.add_argument("--seed", type=int, default=0, help="master seed")
    ap.add_argument("--workers", type=int, default=None,
                    help="accepted for compatibility; all lattices run in one GPU batch")
    ap.add_argument("--no-percolation", action="store_true", help="skip the wrapping observables")
    ap.add_argument("--start", choices=["hot", "cold"], default="hot",
                    help="initial configuration; run both and compare to test equilibration")
    ap.add_argument("--disorder-glob", type=str, default=None,
                    help="optional glob of stored .npz realizations to use instead of generating new ones")
    ap.add_argument("--out", type=str, required=True, help="output .npz path")
    return ap


def main(argv=None) -> None:
    args = build_parser().parse_args(argv)
    temperatures = np.linspace(args.t_min, args.t_max, args.n_temps)
    disorder = None
    if args.disorder_glob is not None:
        disorder = load_disorder_directory(args.disorder_glob, args.realizations)
    print(f"L={args.L}  p={args.p}  {args.n_temps} temperatures in [{args.t_min}, {args.t_max}]  "
          f"{args.realizations} realization(s)  {args.sweeps} sweeps")
    start = time.time()
    results = run_temperature_scan(
        args.L, args.p, temperatures,
        n_realizations=args.realizations, n_sweeps=args.sweeps, n_therm=args.therm,
        n_bins=args.bins, seed=args.seed, n_workers=args.workers,
        measure_percolation=not args.no_percolation, start=args.start, disorder=disorder,
    )
    elapsed = time.time() - start
    save_results(args.out, results)
    print(f"finished in {elapsed:.1f} s; wrote {args.out}")
    if args.realizations == 1:
        print("warning: a single disorder realization was used, so the reported error is "
              "thermal only and does not include sample-to-sample fluctuation.")


if __name__ == "__main__":
    main()
