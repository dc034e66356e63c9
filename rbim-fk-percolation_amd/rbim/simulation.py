"""Temperature scans: device-resident sweeps, host-side estimators.

Drop-in for the reference's ``src/rbim/simulation.py``.  The call signatures,
the primitive and observable names, the results dict, the seeding of the
disorder and the ``.npz`` results format are the same.  What moved: the
per-task sweep loop (simulation.py:196-236) and the process pool (:395-399) are
replaced by one :class:`rbim.engine.Batch` holding every (temperature x
realization) lattice on the GPU; per-bin primitive sums accumulate on the
device.  The estimators -- derived observables, disorder average, jackknife
over realizations, thermal error from bins -- run here on the host, with the
reference's formulas (simulation.py:108-145, 261-333).

Random streams: the disorder comes from ``SeedSequence(seed).spawn(2)[0]``
exactly as in the reference, so the same seed gives the same couplings.  The
thermal stream is Philox4x32-10 keyed by a 64-bit key drawn from the second
child, with the task index ``k = t * R + r`` as the lattice counter word; the
result of a lattice does not depend on what else shares its batch.
"""

from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from rbim import _hip
from rbim.disorder import generate_disorder
from rbim.engine import Batch, LatticeSpec, engine_seed, moments_to_primitive_arrays
from rbim.lattice import split_couplings

__all__ = [
    "OBSERVABLES",
    "PRIMITIVES",
    "RealizationResult",
    "derive_observables",
    "run_single_realization",
    "run_temperature_scan",
    "save_results",
    "load_results",
]

PRIMITIVES = (
    "energy",
    "energy_sq",
    "magnetization",
    "abs_magnetization",
    "magnetization_sq",
    "magnetization_4",
    "largest_size",
    "largest_abs_mag",
    "second_size",
    "mag_ratio",
    "n_clusters",
    "wrap_x",
    "wrap_y",
    "wrap_any",
)

OBSERVABLES = (
    "energy",
    "specific_heat",
    "magnetization",
    "abs_magnetization",
    "susceptibility",
    "binder",
    "p_infinity",
    "largest_size",
    "largest_abs_magnetization",
    "cluster_magnetization_ratio",
    "second_size_fraction",
    "cluster_density_gap",
    "n_clusters_per_site",
    "wrap_x",
    "wrap_y",
    "wrap_any",
)

METADATA_CONVENTIONS = {
    "algorithm": "Swendsen-Wang cluster update on satisfied bonds",
    "estimator": (
        "primitive moments accumulated over the full run of each realization; "
        "derived observables formed after disorder averaging; "
        "uncertainties by jackknife over realizations"
    ),
    "binder_convention": "1 - [<m^4>] / (3 [<m^2>]^2), from disorder-averaged moments",
    "specific_heat_convention": "[<E^2> - <E>^2] / (T^2 N)",
    "susceptibility_convention": "[<M^2> - <|M|>^2] / (T N)",
    "error_convention": (
        "error = jackknife over disorder realizations; "
        "thermal_error = mean within-realization standard error across bins"
    ),
}


@dataclass
class RealizationResult:
    """Full-run primitive means of one realization, plus per-bin means."""

    primitives: Dict[str, float]
    bin_primitives: Dict[str, np.ndarray] = field(default_factory=dict)


def _derive(P, temperature, n_sites: int):
    """derive_observables on floats or on numpy arrays of any common shape
    (elementwise; squares are written as products)."""
    N = n_sites
    m2 = P["magnetization_sq"] / N**2
    m4 = P["magnetization_4"] / N**4
    with np.errstate(divide="ignore", invalid="ignore"):
        binder = np.where(m2 > 0.0, 1.0 - m4 / (3.0 * (m2 * m2)), 0.0)
    return {
        "energy": P["energy"] / N,
        "specific_heat": (P["energy_sq"] - P["energy"] * P["energy"]) / (temperature * temperature * N),
        "magnetization": P["magnetization"] / N,
        "abs_magnetization": P["abs_magnetization"] / N,
        "susceptibility": (P["magnetization_sq"] - P["abs_magnetization"] * P["abs_magnetization"])
        / (temperature * N),
        "binder": binder,
        "p_infinity": P["largest_size"] / N,
        "largest_size": P["largest_size"],
        "largest_abs_magnetization": P["largest_abs_mag"],
        "cluster_magnetization_ratio": P["mag_ratio"],
        "second_size_fraction": P["second_size"] / N,
        "cluster_density_gap": (P["largest_size"] - P["second_size"]) / N,
        "n_clusters_per_site": P["n_clusters"] / N,
        "wrap_x": P["wrap_x"],
        "wrap_y": P["wrap_y"],
        "wrap_any": P["wrap_any"],
    }


def derive_observables(primitives: Dict[str, float], temperature: float, n_sites: int
                       ) -> Dict[str, float]:
    """Physical observables from (disorder-averaged) primitive moments
    (simulation.py:108-145): C = var(E)/(T^2 N), chi = var(|M|)/(T N),
    Binder = 1 - m4 / (3 m2^2) (0 when m2 = 0), P_inf = S_max / N, ..."""
    P = {k: np.float64(primitives[k]) for k in PRIMITIVES}
    return {k: float(v) for k, v in _derive(P, np.float64(temperature), n_sites).items()}


def _check_bins(n_sweeps: int, n_bins: int) -> int:
    if n_bins < 1:
        raise ValueError(f"n_bins must be positive, got {n_bins}")
    per_bin = n_sweeps // n_bins
    if per_bin < 1:
        raise ValueError(
            f"n_sweeps ({n_sweeps}) must be at least n_bins ({n_bins}) so every bin "
            "contains at least one sweep."
        )
    return per_bin


def _arrays_from_batch(batch: Batch, measure_percolation: bool):
    """(prims (n, 14), bins (n, n_bins, 14)) of every lattice of the batch."""
    return moments_to_primitive_arrays(batch.moments(), measure_percolation)


def _results_from_batch(batch: Batch, measure_percolation: bool) -> List[RealizationResult]:
    prims, bins = _arrays_from_batch(batch, measure_percolation)
    return [RealizationResult(primitives={k: float(prims[i, j]) for j, k in enumerate(PRIMITIVES)},
                              bin_primitives={k: bins[i, :, j].copy() for j, k in enumerate(PRIMITIVES)})
            for i in range(batch.n)]


def run_single_realization(
    L: int,
    J: np.ndarray,
    temperature: float,
    *,
    n_sweeps: int,
    n_therm: int,
    n_bins: int = 10,
    seed=None,
    measure_percolation: bool = True,
    start: str = "hot",
    engine: int = _hip.ENGINE_AUTO,
) -> RealizationResult:
    """One realization at one temperature (simulation.py:148-242), on the GPU.

    ``n_therm`` unmeasured sweeps, then ``n_bins * (n_sweeps // n_bins)``
    measured ones (any remainder is not run, as in the reference).
    """
    per_bin = _check_bins(n_sweeps, n_bins)
    if start not in ("hot", "cold"):
        raise ValueError(f"start must be 'hot' or 'cold', got {start!r}")
    J_h, J_v = split_couplings(J)
    if J_h.shape[0] != L:
        raise ValueError(f"couplings are for L={J_h.shape[0]}, not L={L}")
    spec = LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / temperature,
                       seed=engine_seed(seed), lattice_id=0)
    with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=per_bin, start=start,
               engine=engine) as batch:
        batch.run()
        return _results_from_batch(batch, measure_percolation)[0]


def _combine_arrays(prims: np.ndarray, bins: np.ndarray, temperatures: np.ndarray,
                    n_sites: int) -> Dict[str, Dict[str, np.ndarray]]:
    """Disorder average, jackknife over realizations and thermal error from
    bins (simulation.py:261-333) for every temperature at once.

    ``prims`` is (n_T, R, 14) and ``bins`` (n_T, R, n_bins, 14) of primitive
    means in PRIMITIVES order; returns name -> {"value", "error",
    "thermal_error"} arrays of length n_T.  Every reduction runs along a
    contiguous last axis, i.e. as numpy reduces one temperature's vector.
    """
    nT, R = prims.shape[0], prims.shape[1]
    n_bins = bins.shape[2]
    T = np.asarray(temperatures, dtype=np.float64)
    table = {k: np.ascontiguousarray(prims[:, :, j]) for j, k in enumerate(PRIMITIVES)}
    central = _derive({k: v.mean(axis=-1) for k, v in table.items()}, T, n_sites)
    out = {name: {"value": np.asarray(central[name], dtype=np.float64) * np.ones(nT)}
           for name in OBSERVABLES}
    if R > 1:
        sums = {k: v.sum(axis=-1) for k, v in table.items()}
        reduced = {k: (sums[k][:, None] - table[k]) / (R - 1) for k in PRIMITIVES}
        loo = _derive(reduced, T[:, None], n_sites)
        for name in OBSERVABLES:
            x = np.asarray(loo[name], dtype=np.float64) * np.ones((nT, R))
            spread = x - x.mean(axis=-1)[:, None]
            out[name]["error"] = np.sqrt((R - 1) / R * np.sum(spread**2, axis=-1))
    else:
        for name in OBSERVABLES:
            out[name]["error"] = np.full(nT, np.nan)
    if n_bins > 1:
        per_bin = _derive({k: np.ascontiguousarray(bins[:, :, :, j]) for j, k in enumerate(PRIMITIVES)},
                          T[:, None, None], n_sites)
        for name in OBSERVABLES:
            x = np.asarray(per_bin[name], dtype=np.float64) * np.ones((nT, R, n_bins))
            spread = np.std(x, axis=-1, ddof=1) / np.sqrt(n_bins)
            out[name]["thermal_error"] = np.sqrt(np.mean(spread**2, axis=-1) / R)
    else:
        for name in OBSERVABLES:
            out[name]["thermal_error"] = np.full(nT, np.nan)
    if R == 1:
        for name in OBSERVABLES:
            out[name]["error"] = out[name]["thermal_error"].copy()
    return out


def _stack(results: Sequence[RealizationResult]):
    prims = np.array([[r.primitives[k] for k in PRIMITIVES] for r in results], dtype=np.float64)
    bins = np.stack([np.stack([np.asarray(r.bin_primitives[k], dtype=np.float64) for k in PRIMITIVES],
                              axis=-1) for r in results])
    return prims, bins


def _combine_realizations(results: Sequence[RealizationResult], temperature: float,
                          n_sites: int) -> Dict[str, Dict[str, float]]:
    """Disorder average, jackknife over realizations, thermal error from bins
    (simulation.py:261-333) at one temperature."""
    prims, bins = _stack(results)
    c = _combine_arrays(prims[None], bins[None], np.array([temperature]), n_sites)
    return {name: {key: float(v[0]) for key, v in entry.items()} for name, entry in c.items()}


def scan_lattices(L: int, realizations: Sequence[tuple], temperatures: Sequence[float],
                  key: int) -> List[LatticeSpec]:
    """Lattices of a scan in task order: temperature outer, realization inner
    (simulation.py:385-393); lattice k = t * R + r."""
    R = len(realizations)
    specs = []
    for t, temperature in enumerate(temperatures):
        for r, (J_h, J_v) in enumerate(realizations):
            specs.append(LatticeSpec(L=L, J_h=np.asarray(J_h, np.int8), J_v=np.asarray(J_v, np.int8),
                                     beta=1.0 / float(temperature), seed=key, lattice_id=t * R + r))
    return specs


def _assemble_arrays(L: int, temperatures: Sequence[float], prims: np.ndarray, bins: np.ndarray,
                     R: int, metadata: Dict[str, object]) -> Dict[str, object]:
    """The results dict from the primitive means of a scan's lattices in task
    order (lattice k = t * R + r)."""
    nT = len(temperatures)
    c = _combine_arrays(prims.reshape(nT, R, -1), bins.reshape(nT, R, bins.shape[1], -1),
                        np.asarray(temperatures, dtype=np.float64), L * L)
    observables = {name: {key: np.array(c[name][key], dtype=np.float64)
                          for key in ("value", "error", "thermal_error")} for name in OBSERVABLES}
    return {"temperatures": np.asarray(temperatures), "observables": observables, "metadata": metadata}


def assemble_results(L: int, p: float, temperatures: Sequence[float], raw: Sequence[RealizationResult],
                     R: int, metadata: Dict[str, object]) -> Dict[str, object]:
    prims, bins = _stack(raw)
    return _assemble_arrays(L, temperatures, prims, bins, R, metadata)


def run_temperature_scan(
    L: int,
    p: float,
    temperatures: Sequence[float],
    *,
    n_realizations: int = 1,
    n_sweeps: int = 10_000,
    n_therm: Optional[int] = None,
    n_bins: int = 10,
    seed: int = 0,
    n_workers: Optional[int] = None,
    measure_percolation: bool = True,
    start: str = "hot",
    disorder: Optional[Iterable[tuple]] = None,
    engine: int = _hip.ENGINE_AUTO,
) -> Dict[str, object]:
    """Scan temperatures, averaging over disorder realizations
    (simulation.py:336-444).  ``n_workers`` is accepted for compatibility and
    ignored: every (temperature x realization) lattice runs in one GPU batch."""
    temperatures = list(temperatures)
    if n_therm is None:
        n_therm = max(1, n_sweeps // 10)
    per_bin = _check_bins(n_sweeps, n_bins)
    if start not in ("hot", "cold"):
        raise ValueError(f"start must be 'hot' or 'cold', got {start!r}")
    disorder_seeds, run_seeds = np.random.SeedSequence(seed).spawn(2)
    if disorder is None:
        rng = np.random.default_rng(disorder_seeds)
        realizations = [generate_disorder(L, p, rng) for _ in range(n_realizations)]
    else:
        realizations = list(disorder)[:n_realizations]
        if len(realizations) < n_realizations:
            raise ValueError(
                f"only {len(realizations)} disorder realizations supplied but "
                f"n_realizations={n_realizations}"
            )
    R = len(realizations)
    specs = scan_lattices(L, realizations, temperatures, engine_seed(run_seeds))
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=per_bin, start=start,
               engine=engine) as batch:
        batch.run()
        prims, bins = _arrays_from_batch(batch, measure_percolation)
    metadata = {
        "L": L,
        "p": p,
        "n_realizations": R,
        "n_sweeps": n_sweeps,
        "n_therm": n_therm,
        "n_bins": n_bins,
        "seed": seed,
        "measure_percolation": measure_percolation,
        "start": start,
        **METADATA_CONVENTIONS,
        "engine": "rbim-hip gfx950",
        "rng": "Philox4x32-10 keyed by (seed, lattice, sweep, site pair)",
    }
    return _assemble_arrays(L, temperatures, prims, bins, R, metadata)


def save_results(path: str, results: Dict[str, object]) -> None:
    """Flat compressed ``.npz``: temperatures, ``{name}__{field}``, metadata_json
    (simulation.py:447-458)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    flat = {"temperatures": results["temperatures"]}
    for name, entry in results["observables"].items():
        for key, array in entry.items():
            flat[f"{name}__{key}"] = array
    flat["metadata_json"] = np.asarray(json.dumps(results["metadata"]))
    np.savez_compressed(path, **flat)


def load_results(path: str) -> Dict[str, object]:
    """Read a file written by :func:`save_results` (or by the reference)."""
    with np.load(path, allow_pickle=False) as data:
        observables: Dict[str, Dict[str, np.ndarray]] = {}
        for key in data.files:
            if "__" in key:
                name, field_name = key.split("__", 1)
                observables.setdefault(name, {})[field_name] = data[key]
        return {
            "temperatures": data["temperatures"],
            "observables": observables,
            "metadata": json.loads(str(data["metadata_json"])),
        }
