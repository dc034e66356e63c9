"""One Swendsen-Wang update and its measurement, on the GPU.

Mirrors ``src/rbim/swendsen_wang.py``: :func:`sw_sweep` updates ``spins`` in
place and returns a :class:`SweepResult` (``None`` for thermalisation sweeps).
The whole update -- activation, labelling with exact wrapping, the cluster
flip, cluster statistics, energy and magnetization -- is one call into the
gfx950 engine.  Its random draws are Philox words keyed by one 64-bit value
taken from ``rng`` per call.

On exact size ties the "largest" cluster is the one with the smaller minimum
site index (the reference leaves that choice to numpy's introselect,
``swendsen_wang.py:92-100``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from rbim import _hip
from rbim.clusters import activation_threshold
from rbim.lattice import Lattice, split_couplings

__all__ = [
    "SweepResult",
    "energy",
    "magnetization",
    "random_spins",
    "aligned_spins",
    "initial_spins",
    "sw_sweep",
]


@dataclass(frozen=True)
class SweepResult:
    """Observables of one sweep (swendsen_wang.py:24-53)."""

    energy: float
    magnetization: float
    n_clusters: int
    largest_size: int
    largest_abs_magnetization: float
    second_size: int
    second_abs_magnetization: float
    wraps_x: Optional[bool]
    wraps_y: Optional[bool]


def random_spins(n_sites: int, rng: np.random.Generator) -> np.ndarray:
    return rng.integers(0, 2, size=n_sites, dtype=np.int64) * 2 - 1


def aligned_spins(n_sites: int) -> np.ndarray:
    return np.ones(n_sites, dtype=np.int64)


def initial_spins(n_sites: int, rng: np.random.Generator, start: str = "hot") -> np.ndarray:
    if start == "hot":
        return random_spins(n_sites, rng)
    if start == "cold":
        return aligned_spins(n_sites)
    raise ValueError(f"start must be 'hot' or 'cold', got {start!r}")


def _em(spins: np.ndarray, lattice: Lattice, J: np.ndarray):
    J_h, J_v = split_couplings(J)
    s = np.ascontiguousarray(spins, dtype=np.int8)
    e, m = ctypes.c_int64(), ctypes.c_int64()
    ctx = _hip.context()
    _hip.check(
        _hip.lib().rbim_energy_magnetization(ctx, lattice.L, _hip.ptr(s), _hip.ptr(J_h),
                                             _hip.ptr(J_v), ctypes.byref(e), ctypes.byref(m)),
        ctx,
    )
    return e.value, m.value


def energy(spins: np.ndarray, lattice: Lattice, J: np.ndarray) -> float:
    """``-sum_<ij> J_ij s_i s_j`` over the 2N bonds (swendsen_wang.py:81-84)."""
    return float(_em(spins, lattice, J)[0])


def magnetization(spins: np.ndarray) -> float:
    """``sum_i s_i`` (swendsen_wang.py:87-89)."""
    return float(np.sum(spins))


def _result(st: _hip.SweepStats, measure_percolation: bool) -> SweepResult:
    return SweepResult(
        energy=float(st.energy),
        magnetization=float(st.magnetization),
        n_clusters=int(st.n_clusters),
        largest_size=int(st.largest_size),
        largest_abs_magnetization=float(st.largest_abs_mag),
        second_size=int(st.second_size),
        second_abs_magnetization=float(st.second_abs_mag),
        wraps_x=bool(st.wrap_x) if measure_percolation else None,
        wraps_y=bool(st.wrap_y) if measure_percolation else None,
    )


def sw_sweep(
    spins: np.ndarray,
    lattice: Lattice,
    J: np.ndarray,
    beta: float,
    rng: np.random.Generator,
    *,
    measure: bool = True,
    measure_percolation: bool = True,
    engine: int = _hip.ENGINE_PLAIN,
) -> Optional[SweepResult]:
    """One SW update of ``spins`` in place (swendsen_wang.py:103-172)."""
    J_h, J_v = split_couplings(J)
    s = np.ascontiguousarray(spins, dtype=np.int8).copy()
    key = int(rng.integers(0, 2**64, dtype=np.uint64))
    st = _hip.SweepStats()
    ctx = _hip.context()
    _hip.check(
        _hip.lib().rbim_sweep(ctx, lattice.L, _hip.ptr(s), _hip.ptr(J_h), _hip.ptr(J_v),
                              activation_threshold(beta), key, 0, 0, 1 if measure else 0,
                              engine, ctypes.byref(st)),
        ctx,
    )
    spins[...] = s
    if not measure:
        return None
    return _result(st, measure_percolation)


def sweep_given_bits(spins: np.ndarray, lattice: Lattice, J: np.ndarray, active: np.ndarray,
                     flip_by_minsite: np.ndarray, *, engine: int = _hip.ENGINE_PLAIN,
                     return_labels: bool = False):
    """One SW update with given occupation bits and per-cluster flip decisions
    (``flip_by_minsite[m]`` for the cluster whose minimum site is m).  Updates
    ``spins`` in place; returns the :class:`SweepResult` (and labels)."""
    J_h, J_v = split_couplings(J)
    L = lattice.L
    s = np.ascontiguousarray(spins, dtype=np.int8).copy()
    act = np.ascontiguousarray(active, dtype=np.uint8)
    flips = np.ascontiguousarray(flip_by_minsite, dtype=np.uint8)
    labels = np.empty(L * L, dtype=np.int32) if return_labels else None
    st = _hip.SweepStats()
    ctx = _hip.context()
    _hip.check(
        _hip.lib().rbim_sweep_given_bits(ctx, L, _hip.ptr(s), _hip.ptr(J_h), _hip.ptr(J_v),
                                         _hip.ptr(act), _hip.ptr(flips), engine,
                                         _hip.ptr(labels), ctypes.byref(st)),
        ctx,
    )
    spins[...] = s
    result = _result(st, True)
    if return_labels:
        return result, labels.astype(np.int64)
    return result
