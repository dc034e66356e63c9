"""Finite-size crossings of observable curves (Binder cumulant, wrapping
probability) on a common temperature grid.

The reference locates them one size pair at a time in
``scripts/validate_pure_ising.py:38-80`` (and ``scripts/analyze_crossings.py``):
the first grid interval over which the difference of two curves changes
sign, linearly interpolated, or the first grid point where the difference is
exactly zero.  Here every size pair is handled at once with array operations.
"""

from __future__ import annotations

from typing import Dict, Optional

import numpy as np

__all__ = ["first_crossing", "size_pair_crossings", "mean_crossing"]

PAIR_DTYPE = np.dtype([("L_small", np.int64), ("L_large", np.int64), ("T", np.float64)])


def first_crossing(temperatures, a, b) -> np.ndarray:
    """Temperature of the first crossing of ``a`` and ``b`` along the last axis.

    ``a`` and ``b`` broadcast to (..., n_T).  An event at grid index i < n_T - 1
    is either ``a[i] == b[i]`` (the crossing is T[i]) or a strict sign change
    of ``a - b`` between i and i + 1 (the crossing is the zero of the linear
    interpolant).  The earliest event wins; NaN where there is none.
    """
    t = np.asarray(temperatures, dtype=float)
    d = np.asarray(a, dtype=float) - np.asarray(b, dtype=float)
    if t.size < 2:
        return np.full(d.shape[:-1], np.nan)
    lo, hi = d[..., :-1], d[..., 1:]
    touch = lo == 0.0
    event = touch | (np.sign(lo) * np.sign(hi) < 0.0)
    i = np.argmax(event, axis=-1)[..., None]
    d0 = np.take_along_axis(lo, i, -1)[..., 0]
    d1 = np.take_along_axis(hi, i, -1)[..., 0]
    t0, t1 = t[i[..., 0]], t[i[..., 0] + 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        interp = t0 + (t1 - t0) * d0 / (d0 - d1)
    x = np.where(np.take_along_axis(touch, i, -1)[..., 0], t0, interp)
    return np.where(event.any(axis=-1), x, np.nan)


def size_pair_crossings(temperatures, curves: Dict[int, np.ndarray]) -> np.ndarray:
    """Crossings of every pair of sizes (smaller first, in increasing order of
    both), as a structured array with fields L_small, L_large, T; pairs whose
    curves do not cross on the grid are left out."""
    sizes = np.array(sorted(curves), dtype=np.int64)
    if sizes.size < 2:
        return np.zeros(0, dtype=PAIR_DTYPE)
    stack = np.stack([np.asarray(curves[int(L)], dtype=float) for L in sizes])
    ia, ib = np.triu_indices(sizes.size, k=1)
    where = first_crossing(temperatures, stack[ia], stack[ib])
    keep = np.isfinite(where)
    out = np.zeros(int(keep.sum()), dtype=PAIR_DTYPE)
    out["L_small"], out["L_large"], out["T"] = sizes[ia[keep]], sizes[ib[keep]], where[keep]
    return out


def mean_crossing(pairs: np.ndarray) -> Optional[float]:
    """Mean crossing temperature over the size pairs (None if there are none)."""
    return float(np.mean(pairs["T"])) if pairs.size else None
