"""Launch-shape autotuner for the GPU engine.

For a lattice size L it times short runs of a synthetic batch (p = 0.05
disorder, temperatures across [1, 3], hot start) over the candidate shapes and
keeps the fastest:

* LDS-resident kernel (where the auto engine picks it: L <= 64, or L <= 128 on
  sets of fewer than 2^23 sites): workgroup size, i.e. wavefronts per
  lattice (64 ... 1024 threads), and sweeps per launch;
* tiled kernels (otherwise): tile side (32 or 64, the last tile row and
  column cut when it does not divide L) and sweeps per captured hipGraph.

Results go to a JSON cache keyed by GPU architecture, engine kind (LDS or
tiled), L and a lattice-count bucket (``$RBIM_AUTOTUNE_CACHE``, default ``~/.cache/rbim-hip/autotune.json``);
:class:`rbim.engine.Batch` consults it when it is given no explicit shape.
Every shape gives bit-identical results (the random streams depend only on
seed, lattice, sweep and site), so tuning never changes a result.

    python -m rbim.autotune --L 16 32 64 128 256
"""

from __future__ import annotations

import argparse
import json
import os
import threading
from typing import Dict, Optional

import numpy as np

__all__ = ["tune", "lookup", "cache_path"]

LDS_THREADS = (64, 128, 256, 512, 1024)
LDS_SWEEPS = (250, 1000, 2000)
GRAPH_SWEEPS = (4, 16, 64)
_lock = threading.Lock()


def cache_path() -> str:
    return os.environ.get("RBIM_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "rbim-hip", "autotune.json"))


def _bucket(n_lattices: int) -> int:
    b = 1
    while b < n_lattices:
        b *= 4
    return b


def _kind(L: int, engine: int = 0, n_lattices: int = 1) -> Optional[str]:
    """Engine a tuned shape applies to: "lds" (the LDS-resident kernel the auto
    engine picks for n_lattices of side L, or BATCHED), "tiled" (the edge-slot
    sweep it picks otherwise, or PLAIN), or None for an engine choice the tuner
    never timed."""
    from rbim import _hip
    from rbim.engine import auto_uses_lds

    if engine == _hip.ENGINE_AUTO:
        return "lds" if auto_uses_lds(L, n_lattices) else "tiled"
    if engine == _hip.ENGINE_BATCHED and L <= 128:
        return "lds"
    return "tiled" if engine == _hip.ENGINE_PLAIN else None


def _key(L: int, n_lattices: int, kind: str, arch: str = "gfx950") -> str:
    return f"{arch}/{kind}/L{L}/n{_bucket(n_lattices)}"


def _load() -> Dict[str, dict]:
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def lookup(L: int, n_lattices: int, engine: int = 0) -> Optional[dict]:
    """Tuned {"threads", "sweeps_per_launch", "tile"} for this shape on this
    engine choice, or None."""
    kind = _kind(L, engine, n_lattices)
    if kind is None:
        return None
    with _lock:
        return _load().get(_key(L, n_lattices, kind))


def _store(L: int, n_lattices: int, entry: dict) -> None:
    with _lock:
        data = _load()
        data[_key(L, n_lattices, _kind(L, 0, n_lattices))] = entry
        os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
        tmp = cache_path() + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, cache_path())


def _synthetic(L: int, n_lattices: int):
    from rbim.engine import LatticeSpec

    rng = np.random.default_rng(L)
    temps = np.linspace(1.0, 3.0, n_lattices)
    specs = []
    for i in range(n_lattices):
        J_h = np.where(rng.random((L, L)) < 0.05, -1, 1).astype(np.int8)
        J_v = np.where(rng.random((L, L)) < 0.05, -1, 1).astype(np.int8)
        specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / temps[i], seed=7, lattice_id=i))
    return specs


def tune(L: int, n_lattices: int = 1280, sweeps: int = 2000, store: bool = True) -> dict:
    """Time every candidate shape for L and return (and cache) the fastest."""
    from rbim import _hip
    from rbim.engine import Batch

    specs = _synthetic(L, n_lattices)
    sites = L * L * n_lattices
    lds = _kind(L, 0, n_lattices) == "lds"  # the engine the auto choice runs
    if lds:
        shapes = [(t, k, 0) for t in LDS_THREADS for k in LDS_SWEEPS if k <= sweeps]
    else:
        tiles = [T for T in (32, 64) if L >= T]  # cut tiles where T does not divide L
        shapes = [(0, k, T) for T in tiles for k in GRAPH_SWEEPS if k <= sweeps]
    trials = []
    for threads, per_launch, tile in shapes:
        with Batch(specs, n_therm=sweeps // 10, n_bins=1, sweeps_per_bin=sweeps - sweeps // 10,
                   engine=_hip.ENGINE_AUTO, threads=threads, sweeps_per_launch=per_launch,
                   tile=tile) as batch:
            warm = min(50, sweeps // 4)
            batch.run_timed(warm)
            ms = batch.run_timed(sweeps - warm)
        rate = sites * (sweeps - warm) / (max(ms, 1e-6) / 1e3)
        trials.append({"threads": threads, "sweeps_per_launch": per_launch, "tile": tile,
                       "site_updates_per_s": rate})
    best = max(trials, key=lambda t: t["site_updates_per_s"])
    entry = {"threads": best["threads"], "sweeps_per_launch": best["sweeps_per_launch"],
             "tile": best["tile"], "site_updates_per_s": best["site_updates_per_s"],
             "trials": trials}
    if store:
        _store(L, n_lattices, entry)
    return entry


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--L", type=int, nargs="+", required=True)
    ap.add_argument("--lattices", type=int, default=1280)
    ap.add_argument("--sweeps", type=int, default=2000)
    args = ap.parse_args(argv)
    for L in args.L:
        entry = tune(L, args.lattices, args.sweeps)
        print(json.dumps({"L": L, "threads": entry["threads"],
                          "sweeps_per_launch": entry["sweeps_per_launch"], "tile": entry["tile"],
                          "site_updates_per_s": entry["site_updates_per_s"]}))


if __name__ == "__main__":
    main()
