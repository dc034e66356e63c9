#!/usr/bin/env python3
"""Generate and store +-J disorder realizations (host only).

Same files as the reference's scripts/generate_disorder.py: realization i is
drawn from ``default_rng(SeedSequence(seed).spawn(R)[i])``.

    python rbim-fk-percolation_amd/scripts/generate_disorder.py --L 32 --p 0.05 \\
        --realizations 64 --out data/disorder
"""

from __future__ import annotations

import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from rbim.disorder import (  # noqa: E402
    default_disorder_filename,
    frustration_density,
    generate_disorder,
    save_disorder,
)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--L", type=int, required=True, help="linear system size")
    ap.add_argument("--p", type=float, required=True, help="antiferromagnetic bond concentration")
    ap.add_argument("--realizations", type=int, default=1, help="number of realizations to generate")
    ap.add_argument("--out", type=str, default="data/disorder", help="output directory")
    ap.add_argument("--seed", type=int, default=0, help="master seed")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    densities = []
    for index, child in enumerate(np.random.SeedSequence(args.seed).spawn(args.realizations)):
        J_h, J_v = generate_disorder(args.L, args.p, np.random.default_rng(child))
        densities.append(frustration_density(J_h, J_v))
        save_disorder(os.path.join(args.out, default_disorder_filename(args.L, args.p, index)),
                      J_h, J_v, p=args.p, seed=args.seed)
    expected = 4 * args.p * (1 - args.p) * ((1 - args.p) ** 2 + args.p**2)
    print(f"wrote {args.realizations} realization(s) for L={args.L}, p={args.p} to {args.out}")
    print(f"frustrated plaquette density: {np.mean(densities):.4f} (expected {expected:.4f})")


if __name__ == "__main__":
    main()
