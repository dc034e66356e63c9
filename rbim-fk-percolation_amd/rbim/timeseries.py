"""Per-sweep time series from the GPU engine: autocorrelation times and
equilibration checks.

The reference measures integrated autocorrelation times by stepping
``sw_sweep`` in a Python loop per (temperature, realization) and storing each
sweep's observables (``scripts/measure_autocorrelation.py:69-90``).  Here every
(temperature x realization) lattice of a scan runs as one device batch with a
per-sweep stats ring (``rbim_batch_series``), and the series come back in one
copy.  The random stream is the engine's Philox stream, so the series agree
with the reference statistically, not sweep by sweep.

Series fields: energy, magnetization, n_clusters, largest_size,
largest_abs_mag, second_size, second_abs_mag, wrap_x, wrap_y, plus the
derived mag_ratio = largest_abs_mag / largest_size (0 when there is no
cluster), as the reference script tracks it.
"""

from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from rbim import _hip
from rbim.engine import Batch, LatticeSpec, engine_seed

__all__ = [
    "TRACKED",
    "collect_series",
    "integrated_autocorrelation",
    "equilibration_check",
]

# Observables the reference script reports (measure_autocorrelation.py:36).
TRACKED = ("energy", "largest_size", "largest_abs_mag", "mag_ratio", "wrap_y")


def _fields(stats: np.ndarray) -> Dict[str, np.ndarray]:
    out = {name: stats[name].astype(np.float64) for name in stats.dtype.names}
    size = out["largest_size"]
    out["mag_ratio"] = np.divide(out["largest_abs_mag"], size, out=np.zeros_like(size),
                                 where=size > 0)
    return out


def collect_series(L: int, couplings: Sequence[Tuple[np.ndarray, np.ndarray]],
                   temperatures: Sequence[float], n_sweeps: int, n_therm: int, seed,
                   start: str = "hot", engine: int = _hip.ENGINE_AUTO
                   ) -> Dict[str, np.ndarray]:
    """Time series of every (temperature, realization) lattice, run as one batch.

    ``couplings`` is a list of (J_h, J_v) planes; lattice k = t * R + r (the
    task order of run_temperature_scan).  Returns name -> (n_T, R, n_sweeps)
    float64 arrays of the measured sweeps (after ``n_therm`` unmeasured ones).
    """
    temps = [float(t) for t in temperatures]
    R = len(couplings)
    if n_sweeps < 1:
        raise ValueError("n_sweeps must be >= 1")
    key = engine_seed(seed)
    specs = []
    for t, T in enumerate(temps):
        for r, (J_h, J_v) in enumerate(couplings):
            specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / T, seed=key,
                                     lattice_id=t * R + r))
    with Batch(specs, n_therm=n_therm, n_bins=1, sweeps_per_bin=n_sweeps, start=start,
               engine=engine, series=n_sweeps) as batch:
        batch.run()
        stats = batch.series()
    fields = _fields(stats)
    return {name: arr.reshape(len(temps), R, -1) for name, arr in fields.items()}


def integrated_autocorrelation(series: np.ndarray, c: float = 6.0) -> Tuple[float, int]:
    """``tau_int = 1/2 + sum_{t>=1} rho(t)`` with Sokal's automatic window.

    rho is the normalised autocovariance of the centred series (FFT with
    zero padding to at least twice the length, so it is not circular).  The
    window W is the first lag with W >= c * tau_int(W), where tau_int(W) =
    1/2 + sum_{t=1..W} rho(t) (A. D. Sokal, "Monte Carlo methods in
    statistical mechanics", 1996, sec. 3).  Constant or very short series give
    (0.5, 0).  Same estimator as scripts/measure_autocorrelation.py:39-66.
    """
    x = np.asarray(series, dtype=np.float64)
    n = x.size
    x = x - x.mean() if n else x
    if n < 16 or not np.any(x):
        return 0.5, 0
    size = 1 << int(np.ceil(np.log2(2 * n)))
    f = np.fft.rfft(x, size)
    acov = np.fft.irfft(f * np.conj(f), size)[:n]
    if acov[0] <= 0:
        return 0.5, 0
    rho = acov / acov[0]
    tau = 0.5 + np.cumsum(rho[1:])          # tau[W - 1] = tau_int(W)
    lags = np.arange(1, n)
    hit = np.nonzero(lags >= c * tau)[0]
    if hit.size:
        W = int(lags[hit[0]])
        return float(max(tau[hit[0]], 0.5)), W
    return float(max(tau[-1], 0.5)), n - 1


def equilibration_check(L: int, J_h: np.ndarray, J_v: np.ndarray, temperature: float,
                        n_sweeps: int, n_blocks: int = 10, seed=0,
                        observable: str = "energy", engine: int = _hip.ENGINE_AUTO
                        ) -> Dict[str, np.ndarray]:
    """Hot and cold starts of one lattice, run side by side with the same
    random stream; returns per-block means of ``observable`` for both starts
    and their difference in units of the combined block standard error.

    An equilibrated run shows the two block-mean curves merging early; a
    persistent gap in the later blocks means ``n_therm`` is too short."""
    if n_sweeps < n_blocks:
        raise ValueError("n_sweeps must be >= n_blocks")
    key = engine_seed(seed)
    out = {}
    for start in ("hot", "cold"):
        spec = LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / float(temperature), seed=key,
                           lattice_id=0)
        with Batch([spec], n_therm=0, n_bins=1, sweeps_per_bin=n_sweeps, start=start,
                   engine=engine, series=n_sweeps) as batch:
            batch.run()
            series = _fields(batch.series())[observable][0]
        blocks = series[: n_blocks * (n_sweeps // n_blocks)].reshape(n_blocks, -1)
        out[start] = blocks.mean(axis=1)
        out[start + "_sem"] = blocks.std(axis=1, ddof=1) / np.sqrt(blocks.shape[1])
    err = np.sqrt(out["hot_sem"] ** 2 + out["cold_sem"] ** 2)
    out["z"] = np.divide(out["hot"] - out["cold"], err, out=np.zeros_like(err), where=err > 0)
    return out
