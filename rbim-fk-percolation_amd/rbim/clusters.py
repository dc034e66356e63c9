"""FK bond activation, cluster labelling and wrapping -- on the GPU.

Same functions and return conventions as the reference's
``src/rbim/clusters.py``; every one of them calls the gfx950 kernels of
``librbim_hip.so``.  Random draws come from the engine's Philox4x32-10 stream;
functions that take a numpy ``Generator`` (as the reference's do) draw one
64-bit key from it per call, so a seeded generator still gives a reproducible
result.
"""

from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from rbim import _hip
from rbim.lattice import Lattice, split_couplings

__all__ = [
    "bond_activation_probability",
    "activation_threshold",
    "activate_bonds",
    "label_clusters",
    "label_and_wrap",
    "canonical_labels",
    "cluster_statistics",
    "wraps_along",
    "spans_along",
]


def bond_activation_probability(beta: float) -> float:
    """``1 - exp(-2 beta)`` via expm1 (clusters.py:40-45)."""
    return float(-np.expm1(-2.0 * beta))


def activation_threshold(beta: float) -> int:
    """31-bit integer threshold the kernels compare draws with:
    ``round(p_add * 2^31)``; a bond is occupied when ``draw >> 1 < threshold``."""
    return int(np.floor(bond_activation_probability(beta) * 2.0**31 + 0.5))


def _key_from(rng: np.random.Generator) -> int:
    return int(rng.integers(0, 2**64, dtype=np.uint64))


def _spins_i8(spins: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(spins, dtype=np.int8)


def activate_bonds(spins: np.ndarray, lattice: Lattice, J: np.ndarray, beta: float,
                   rng: np.random.Generator) -> np.ndarray:
    """Occupied FK bonds, (2N,) bool in bond order (clusters.py:48-79).

    Only satisfied bonds (``J s_i s_j > 0``) can be occupied, each with
    probability ``1 - exp(-2 beta)``; the draws are Philox words keyed by one
    64-bit value taken from ``rng``.
    """
    J_h, J_v = split_couplings(J)
    L = lattice.L
    s = _spins_i8(spins)
    active = np.empty(2 * L * L, dtype=np.uint8)
    ctx = _hip.context()
    _hip.check(
        _hip.lib().rbim_activate(ctx, L, _hip.ptr(s), _hip.ptr(J_h), _hip.ptr(J_v),
                                 activation_threshold(beta), _key_from(rng), 0, 0,
                                 _hip.ptr(active), None),
        ctx,
    )
    return active.astype(bool)


def _label(active: np.ndarray, L: int, engine: int = _hip.ENGINE_PLAIN):
    act = np.ascontiguousarray(active, dtype=np.uint8)
    if act.size != 2 * L * L:
        raise ValueError(f"active has {act.size} entries, expected {2 * L * L}")
    labels = np.empty(L * L, dtype=np.int32)
    canon = np.empty(L * L, dtype=np.int32)
    n = ctypes.c_int32()
    wraps = np.zeros(2, dtype=np.int32)
    ctx = _hip.context()
    _hip.check(
        _hip.lib().rbim_label_and_wrap(ctx, L, _hip.ptr(act), engine, _hip.ptr(labels),
                                       _hip.ptr(canon), ctypes.byref(n), _hip.ptr(wraps)),
        ctx,
    )
    return int(n.value), labels.astype(np.int64), canon.astype(np.int64), bool(wraps[0]), bool(wraps[1])


def label_and_wrap(active: np.ndarray, lattice: Lattice) -> Tuple[int, np.ndarray, bool, bool]:
    """(n_clusters, labels, wraps_x, wraps_y) (clusters.py:144-159).

    Labels are numbered in increasing min-site order, as the reference's
    union-find numbers them; wrapping is decided exactly from the winding of
    every non-tree occupied bond.
    """
    n, labels, _, wx, wy = _label(active, lattice.L)
    return n, labels, wx, wy


def canonical_labels(active: np.ndarray, lattice: Lattice) -> np.ndarray:
    """Per-site minimum site index of its cluster (the engine's native label)."""
    return _label(active, lattice.L)[2]


def label_clusters(active: np.ndarray, lattice: Lattice) -> Tuple[int, np.ndarray]:
    """(n_clusters, labels) of the active-bond graph (clusters.py:82-95)."""
    n, labels, _, _, _ = _label(active, lattice.L)
    return n, labels


def cluster_statistics(labels: np.ndarray, spins: np.ndarray, n_clusters: int
                       ) -> Tuple[np.ndarray, np.ndarray]:
    """Sizes (int64) and signed magnetizations (float64) per cluster
    (clusters.py:98-118)."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    s = _spins_i8(spins)
    sizes = np.zeros(int(n_clusters), dtype=np.int64)
    mags = np.zeros(int(n_clusters), dtype=np.int64)
    ctx = _hip.context()
    _hip.check(
        _hip.lib().rbim_cluster_statistics(ctx, lab.size, _hip.ptr(lab), _hip.ptr(s),
                                           int(n_clusters), _hip.ptr(sizes), _hip.ptr(mags)),
        ctx,
    )
    return sizes, mags.astype(np.float64)


def wraps_along(active: np.ndarray, lattice: Lattice, axis: str) -> bool:
    """Whether some cluster winds around the torus along ``axis`` (exact)."""
    _, _, wx, wy = label_and_wrap(active, lattice)
    if axis == "x":
        return bool(wx)
    if axis == "y":
        return bool(wy)
    raise ValueError(f"axis must be 'x' or 'y', got {axis!r}")


def spans_along(active: np.ndarray, lattice: Lattice, axis: str) -> bool:
    """Whether, with the boundary bonds along ``axis`` removed, some cluster
    touches both open faces (clusters.py:162-180).  Labels come from the GPU."""
    mask = lattice.wrap_mask(axis)
    _, labels = label_clusters(np.asarray(active, dtype=bool) & ~mask, lattice)
    L = lattice.L
    sites = np.arange(lattice.n_sites)
    coord = sites // L if axis == "x" else sites % L
    return bool(np.intersect1d(labels[coord == 0], labels[coord == L - 1]).size > 0)
