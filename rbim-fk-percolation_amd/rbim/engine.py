"""Batches of lattices on the GPU: the device-resident replacement for the
reference's per-task sweep loop (``src/rbim/simulation.py:196-241``).

A :class:`Batch` uploads couplings once, keeps spins and per-bin moment sums in
HBM, and advances every lattice by whole blocks of sweeps per kernel launch.
Moments come back as exact integer sums (and one FP64 sum, |M_C|/S_max) in the
20-slot layout of include/rbim_hip.h; :func:`moments_to_primitives` turns
them into the reference's full-run and per-bin primitive means.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from rbim import _hip

__all__ = ["Batch", "LatticeSpec", "moments_to_primitives", "moments_to_primitive_arrays",
           "engine_seed", "PRIMITIVE_SLOTS"]

PRIMITIVES = (
    "energy", "energy_sq", "magnetization", "abs_magnetization", "magnetization_sq",
    "magnetization_4", "largest_size", "largest_abs_mag", "second_size", "mag_ratio",
    "n_clusters", "wrap_x", "wrap_y", "wrap_any",
)

# primitive -> (slot, kind): "i" signed int64, "u" uint64, "w" uint128 (lo, hi), "d" double
PRIMITIVE_SLOTS = {
    "energy": (0, "i"),
    "energy_sq": (1, "w"),
    "magnetization": (3, "i"),
    "abs_magnetization": (4, "u"),
    "magnetization_sq": (5, "w"),
    "magnetization_4": (7, "w"),
    "largest_size": (9, "u"),
    "largest_abs_mag": (10, "u"),
    "second_size": (11, "u"),
    "mag_ratio": (12, "d"),
    "n_clusters": (13, "u"),
    "wrap_x": (14, "u"),
    "wrap_y": (15, "u"),
    "wrap_any": (16, "u"),
}
COUNT_SLOT = 17


def engine_seed(seed) -> int:
    """64-bit Philox key from an int or ``numpy.random.SeedSequence``."""
    seq = seed if isinstance(seed, np.random.SeedSequence) else np.random.SeedSequence(seed)
    words = seq.generate_state(2, dtype=np.uint32)
    return int(words[0]) | (int(words[1]) << 32)


def _slot_value(row: np.ndarray, slot: int, kind: str):
    if kind == "i":
        return int(row[slot:slot + 1].view(np.int64)[0])
    if kind == "u":
        return int(row[slot])
    if kind == "w":
        return int(row[slot]) | (int(row[slot + 1]) << 64)
    return float(row[slot:slot + 1].view(np.float64)[0])


_EXACT = float(2**53)  # integers below this convert to float64 exactly


def moments_to_primitive_arrays(acc: np.ndarray, measure_percolation: bool = True
                                ) -> Tuple[np.ndarray, np.ndarray]:
    """Full-run and per-bin primitive means of many lattices at once.

    ``acc`` is (n, n_bins, 20) uint64; returns ``(prims, bins)`` of shapes
    (n, 14) and (n, n_bins, 14), columns in PRIMITIVES order.  Integer sums are
    exact, so each mean is the correctly rounded quotient of exact integers:
    float64 division of exactly converted operands while the sums stay below
    2^53, Python integer division beyond.  The one FP64 slot is summed over the
    bins in bin order.  (The reference's ``np.mean`` over a float series,
    simulation.py:238-241, agrees to ~1e-16.)
    """
    acc = np.asarray(acc, dtype=np.uint64)
    n, n_bins = acc.shape[0], acc.shape[1]
    counts = acc[:, :, COUNT_SLOT].astype(np.int64)
    if np.any(counts <= 0):
        raise ZeroDivisionError("a bin holds no measured sweep")
    total = counts.sum(axis=1)
    prims = np.empty((n, len(PRIMITIVES)))
    bins = np.empty((n, n_bins, len(PRIMITIVES)))
    for j, name in enumerate(PRIMITIVES):
        slot, kind = PRIMITIVE_SLOTS[name]
        if not measure_percolation and name.startswith("wrap"):
            prims[:, j] = np.nan
            bins[:, :, j] = np.nan
            continue
        if kind == "d":
            v = np.ascontiguousarray(acc[:, :, slot]).view(np.float64)
            bins[:, :, j] = v / counts
            s = np.zeros(n)
            for b in range(n_bins):  # in bin order, as a Python sum of the bins
                s = s + v[:, b]
            prims[:, j] = s / total
            continue
        if kind == "i":
            v = np.ascontiguousarray(acc[:, :, slot]).view(np.int64)
            fast = np.abs(v).max(axis=1) < _EXACT
            sums = v.sum(axis=1)  # |v| < 2^53 per bin: no int64 overflow
        else:
            v = acc[:, :, slot]
            fast = v.max(axis=1) < _EXACT
            if kind == "w":
                fast &= acc[:, :, slot + 1].max(axis=1) == 0
            sums = v.sum(axis=1)  # bins below 2^53: no uint64 overflow
        fast &= np.abs(sums.astype(np.float64)) < _EXACT
        bins[fast, :, j] = v[fast].astype(np.float64) / counts[fast]
        prims[fast, j] = sums[fast].astype(np.float64) / total[fast]
        for i in np.nonzero(~fast)[0]:
            per_bin = [_slot_value(acc[i, b], slot, kind) for b in range(n_bins)]
            bins[i, :, j] = [x / int(c) for x, c in zip(per_bin, counts[i])]
            prims[i, j] = sum(per_bin) / int(total[i])
    return prims, bins


def moments_to_primitives(acc: np.ndarray, measure_percolation: bool = True
                          ) -> Tuple[Dict[str, float], Dict[str, np.ndarray]]:
    """Full-run and per-bin primitive means from one lattice's bin sums.

    ``acc`` is (n_bins, 20) uint64; the values are those of
    :func:`moments_to_primitive_arrays`.
    """
    p, b = moments_to_primitive_arrays(np.asarray(acc)[None], measure_percolation)
    primitives = {name: float(p[0, j]) for j, name in enumerate(PRIMITIVES)}
    bins = {name: b[0, :, j].copy() for j, name in enumerate(PRIMITIVES)}
    return primitives, bins


@dataclass
class LatticeSpec:
    """One lattice of a batch."""

    L: int
    J_h: np.ndarray
    J_v: np.ndarray
    beta: float
    seed: int
    lattice_id: int
    spins0: Optional[np.ndarray] = None


# Layout of rbim_sweep_stats (one measured sweep of one lattice).
SWEEP_STATS_DTYPE = np.dtype([
    ("energy", np.int64), ("magnetization", np.int64), ("n_clusters", np.int64),
    ("largest_size", np.int64), ("largest_abs_mag", np.int64), ("second_size", np.int64),
    ("second_abs_mag", np.int64), ("wrap_x", np.int32), ("wrap_y", np.int32),
])
assert SWEEP_STATS_DTYPE.itemsize == ctypes.sizeof(_hip.SweepStats)


class Batch:
    """A set of independent lattices advanced together on the device."""

    def __init__(self, lattices: Sequence[LatticeSpec], *, n_therm: int, n_bins: int,
                 sweeps_per_bin: int, start: str = "hot", engine: int = _hip.ENGINE_AUTO,
                 sweeps_per_launch: int = 0, threads: int = 0, streams: int = 0,
                 series: int = 0, tile: int = 0, single_launch: bool = False):
        from rbim.clusters import activation_threshold

        if start not in ("hot", "cold"):
            raise ValueError(f"start must be 'hot' or 'cold', got {start!r}")
        if len(lattices) == 0:
            raise ValueError("a batch needs at least one lattice")
        self.n = len(lattices)
        self.Ls = np.array([spec.L for spec in lattices], dtype=np.int32)
        self.sizes = self.Ls.astype(np.int64) ** 2
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.n_bins = int(n_bins)
        self.sweeps_per_bin = int(sweeps_per_bin)
        self.n_therm = int(n_therm)
        self.total_sweeps = self.n_therm + self.n_bins * self.sweeps_per_bin
        for spec in lattices:
            if spec.J_h.shape != (spec.L, spec.L) or spec.J_v.shape != (spec.L, spec.L):
                raise ValueError(f"couplings must have shape ({spec.L}, {spec.L})")
        self._Jh = np.ascontiguousarray(np.concatenate([np.asarray(s.J_h, np.int8).reshape(-1) for s in lattices]))
        self._Jv = np.ascontiguousarray(np.concatenate([np.asarray(s.J_v, np.int8).reshape(-1) for s in lattices]))
        self._thr = np.array([activation_threshold(s.beta) for s in lattices], dtype=np.uint32)
        self._seed = np.array([s.seed & (2**64 - 1) for s in lattices], dtype=np.uint64)
        self._lid = np.array([s.lattice_id & 0xFFFFFFFF for s in lattices], dtype=np.uint32)
        self._spins0 = None
        if any(s.spins0 is not None for s in lattices):
            self._spins0 = np.ascontiguousarray(np.concatenate([
                (np.asarray(s.spins0, np.int8).reshape(-1) if s.spins0 is not None
                 else np.ones(s.L * s.L, np.int8)) for s in lattices]))
        if (threads == 0 and sweeps_per_launch == 0 and tile == 0 and not single_launch
                and len(set(self.Ls.tolist())) == 1):
            from rbim import autotune

            tuned = autotune.lookup(int(self.Ls[0]), self.n, int(engine))
            if tuned:
                threads, sweeps_per_launch = int(tuned["threads"]), int(tuned["sweeps_per_launch"])
                tile = int(tuned.get("tile", 0))
        self.threads, self.sweeps_per_launch = int(threads), int(sweeps_per_launch)
        desc = _hip.BatchDesc()
        desc.n_lattices = self.n
        desc.L = _hip.ptr(self.Ls)
        desc.Jh = _hip.ptr(self._Jh)
        desc.Jv = _hip.ptr(self._Jv)
        desc.thr31 = _hip.ptr(self._thr)
        desc.seed = _hip.ptr(self._seed)
        desc.lattice_id = _hip.ptr(self._lid)
        desc.spins0 = _hip.ptr(self._spins0)
        desc.start = _hip.START_COLD if start == "cold" else _hip.START_HOT
        desc.n_bins = self.n_bins
        desc.sweeps_per_bin = self.sweeps_per_bin
        desc.n_therm = self.n_therm
        desc.engine = int(engine)
        desc.sweeps_per_launch = int(sweeps_per_launch)
        desc.threads = int(threads)
        desc.streams = int(streams)
        desc.tile = int(tile)
        desc.single_launch = 1 if single_launch else 0
        if series < 0:
            raise ValueError("series must be >= 0")
        self.series_capacity = int(series)
        desc.series = self.series_capacity
        self._desc = desc
        self._ctx = _hip.context()
        handle = ctypes.c_void_p()
        _hip.check(_hip.lib().rbim_batch_create(self._ctx, ctypes.byref(desc), ctypes.byref(handle)), self._ctx)
        self._h = handle

    # -------------------------------------------------------------- control
    def run(self, n_sweeps: Optional[int] = None) -> None:
        """Queue sweeps (all remaining by default); returns immediately."""
        if n_sweeps is None:
            n_sweeps = self.total_sweeps - self.info().sweeps_done
        _hip.check(_hip.lib().rbim_batch_run(self._h, int(n_sweeps)), self._ctx)

    def run_timed(self, n_sweeps: int) -> float:
        """Run and wait; milliseconds between HIP events on the batch's stream."""
        ms = ctypes.c_double()
        _hip.check(_hip.lib().rbim_batch_run_timed(self._h, int(n_sweeps), ctypes.byref(ms)), self._ctx)
        return ms.value

    def done(self) -> bool:
        flag = ctypes.c_int32()
        _hip.check(_hip.lib().rbim_batch_query(self._h, ctypes.byref(flag)), self._ctx)
        return bool(flag.value)

    def sync(self) -> None:
        _hip.check(_hip.lib().rbim_batch_sync(self._h), self._ctx)

    def info(self) -> _hip.BatchInfo:
        out = _hip.BatchInfo()
        _hip.check(_hip.lib().rbim_batch_info_get(self._h, ctypes.byref(out)), self._ctx)
        return out

    # -------------------------------------------------------------- results
    def moments(self) -> np.ndarray:
        out = np.zeros((self.n, self.n_bins, _hip.ACC_SLOTS), dtype=np.uint64)
        _hip.check(_hip.lib().rbim_batch_moments(self._h, _hip.ptr(out)), self._ctx)
        return out

    def spins(self) -> list:
        flat = np.empty(int(self.offsets[-1]), dtype=np.int8)
        _hip.check(_hip.lib().rbim_batch_spins(self._h, _hip.ptr(flat)), self._ctx)
        return [flat[self.offsets[i]:self.offsets[i + 1]].copy() for i in range(self.n)]

    def series(self) -> np.ndarray:
        """Per-sweep stats of the measured sweeps kept in the device ring, in
        order: a (n_lattices, k) structured array (fields of SWEEP_STATS_DTYPE),
        k = min(measured sweeps so far, ring capacity), oldest first."""
        if self.series_capacity <= 0:
            raise RuntimeError("batch was created without series=")
        cap = self.series_capacity
        ring = np.zeros((self.n, cap), dtype=SWEEP_STATS_DTYPE)
        recorded = ctypes.c_int64()
        _hip.check(_hip.lib().rbim_batch_series(self._h, _hip.ptr(ring), ctypes.byref(recorded)),
                   self._ctx)
        n = int(recorded.value)
        if n <= cap:
            return ring[:, :n].copy()
        start = n % cap
        return np.concatenate([ring[:, start:], ring[:, :start]], axis=1)

    def site_updates(self, n_sweeps: int) -> int:
        return int(self.sizes.sum()) * int(n_sweeps)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _hip.lib().rbim_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# The auto engine's choice for a set of lattices of one side (the C library's
# rule, restated for the autotuner and callers that want to know): the LDS
# kernel up to L = 64, and up to L = 128 while the set holds fewer than 2^23
# sites; the edge-slot sweep above.
AUTO_LDS_ALWAYS_L = 64
AUTO_LDS_MAX_SITES = 1 << 23


def auto_uses_lds(L: int, n_lattices: int) -> bool:
    """Whether a batch's ``n_lattices`` lattices of side ``L`` run on the
    LDS-resident kernel under ENGINE_AUTO."""
    return L <= AUTO_LDS_ALWAYS_L or (L <= 128 and n_lattices * L * L < AUTO_LDS_MAX_SITES)


def footprint(lattices: Sequence[LatticeSpec], *, n_bins: int, engine: int = _hip.ENGINE_AUTO,
              series: int = 0, tile: int = 0) -> int:
    """Device bytes a batch of these lattices would hold (admission control)."""
    Ls = np.array([s.L for s in lattices], dtype=np.int32)
    n = int(Ls.size)
    total = int((Ls.astype(np.int64) ** 2).sum())
    zeros8 = np.zeros(total, np.int8)
    desc = _hip.BatchDesc()
    desc.n_lattices = n
    desc.L = _hip.ptr(Ls)
    desc.Jh = _hip.ptr(zeros8)
    desc.Jv = _hip.ptr(zeros8)
    thr = np.zeros(n, np.uint32)
    seed = np.zeros(n, np.uint64)
    lid = np.zeros(n, np.uint32)
    desc.thr31 = _hip.ptr(thr)
    desc.seed = _hip.ptr(seed)
    desc.lattice_id = _hip.ptr(lid)
    desc.n_bins = int(n_bins)
    desc.sweeps_per_bin = 1
    desc.engine = int(engine)
    desc.series = int(series)
    desc.tile = int(tile)
    out = ctypes.c_uint64()
    _hip.check(_hip.lib().rbim_batch_footprint(ctypes.byref(desc), ctypes.byref(out)))
    return int(out.value)
