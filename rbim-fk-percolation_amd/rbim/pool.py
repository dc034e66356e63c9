"""Resident lattice slots of one L on the device (``rbim_pool_*`` of
include/rbim_hip.h): the unit of continuous batching in the scan server.

A pool holds ``n_slots`` lattices of side L, each with its own schedule
(n_therm, bins, sweeps per bin).  Lattices of several requests are loaded into
free slots and advanced together by shared launches; every slot stops at its
own total.  A lattice's moment sums depend only on its own seed, lattice id,
couplings, start and schedule, so they equal those of the same lattice in a
:class:`rbim.engine.Batch`, bit for bit.  This is the serving-side
replacement of the reference's process pool (``src/rbim/simulation.py:395-399``).
"""

from __future__ import annotations

import ctypes
from typing import Sequence, Tuple

import numpy as np

from rbim import _hip
from rbim.engine import LatticeSpec

__all__ = ["SlotPool", "pool_footprint"]


def _desc(L: int, n_slots: int, max_bins: int, engine: int, threads: int,
          sweeps_per_launch: int, tile: int) -> _hip.PoolDesc:
    d = _hip.PoolDesc()
    d.L, d.n_slots, d.max_bins, d.engine = int(L), int(n_slots), int(max_bins), int(engine)
    d.threads, d.sweeps_per_launch, d.tile = int(threads), int(sweeps_per_launch), int(tile)
    return d


def pool_footprint(L: int, n_slots: int, max_bins: int = 10, engine: int = _hip.ENGINE_AUTO,
                   threads: int = 0, sweeps_per_launch: int = 0, tile: int = 0) -> int:
    """Device bytes a pool of these dimensions holds."""
    d = _desc(L, n_slots, max_bins, engine, threads, sweeps_per_launch, tile)
    out = ctypes.c_uint64()
    _hip.check(_hip.lib().rbim_pool_footprint(ctypes.byref(d), ctypes.byref(out)))
    return int(out.value)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class SlotPool:
    """``n_slots`` resident lattice slots of side ``L``."""

    def __init__(self, L: int, n_slots: int, max_bins: int = 10, engine: int = _hip.ENGINE_AUTO,
                 threads: int = 0, sweeps_per_launch: int = 0, tile: int = 0):
        self.L, self.n_slots, self.max_bins = int(L), int(n_slots), int(max_bins)
        self.nbytes = pool_footprint(L, n_slots, max_bins, engine, threads, sweeps_per_launch, tile)
        self._ctx = _hip.context()
        d = _desc(L, n_slots, max_bins, engine, threads, sweeps_per_launch, tile)
        handle = ctypes.c_void_p()
        _hip.check(_hip.lib().rbim_pool_create(self._ctx, ctypes.byref(d), ctypes.byref(handle)),
                   self._ctx)
        self._h = handle
        self._free = np.ones(self.n_slots, dtype=bool)

    # ------------------------------------------------------------- slots
    @property
    def free_slots(self) -> int:
        return int(self._free.sum())

    @property
    def in_use(self) -> int:
        return self.n_slots - self.free_slots

    @staticmethod
    def pack(specs: Sequence[LatticeSpec]) -> dict:
        """Host arrays of a load (couplings, thresholds, keys, initial spins),
        built once so a later :meth:`load` only copies them to the device."""
        from rbim.clusters import activation_threshold

        if len(specs) == 0:
            raise ValueError("nothing to load")
        L = specs[0].L
        for s in specs:
            if s.L != L or s.J_h.shape != (L, L) or s.J_v.shape != (L, L):
                raise ValueError(f"a load holds lattices of one L with ({L}, {L}) couplings")
        packed = {
            "L": L,
            "n": len(specs),
            "Jh": np.ascontiguousarray(np.concatenate([np.asarray(s.J_h, np.int8).reshape(-1) for s in specs])),
            "Jv": np.ascontiguousarray(np.concatenate([np.asarray(s.J_v, np.int8).reshape(-1) for s in specs])),
            "thr": np.array([activation_threshold(s.beta) for s in specs], dtype=np.uint32),
            "seed": np.array([s.seed & (2**64 - 1) for s in specs], dtype=np.uint64),
            "lid": np.array([s.lattice_id & 0xFFFFFFFF for s in specs], dtype=np.uint32),
            "spins0": None,
        }
        if any(s.spins0 is not None for s in specs):
            packed["spins0"] = np.ascontiguousarray(np.concatenate([
                (np.asarray(s.spins0, np.int8).reshape(-1) if s.spins0 is not None
                 else np.ones(L * L, np.int8)) for s in specs]))
        return packed

    def load(self, specs, *, n_therm: int, n_bins: int, sweeps_per_bin: int,
             start: str = "hot") -> np.ndarray:
        """Load lattices (LatticeSpecs, or the result of :meth:`pack`) into the
        lowest free slots; returns their slot indices."""
        packed = specs if isinstance(specs, dict) else self.pack(specs)
        if start not in ("hot", "cold"):
            raise ValueError(f"start must be 'hot' or 'cold', got {start!r}")
        n = packed["n"]
        if packed["L"] != self.L:
            raise ValueError(f"this pool holds L={self.L} lattices, not L={packed['L']}")
        if n > self.free_slots:
            raise RuntimeError(f"{n} lattices for {self.free_slots} free slots")
        slots = _i32(np.flatnonzero(self._free)[:n])
        code = _hip.START_COLD if start == "cold" else _hip.START_HOT
        _hip.check(_hip.lib().rbim_pool_load(
            self._h, n, _hip.ptr(slots), _hip.ptr(packed["Jh"]), _hip.ptr(packed["Jv"]),
            _hip.ptr(packed["thr"]), _hip.ptr(packed["seed"]), _hip.ptr(packed["lid"]),
            _hip.ptr(packed["spins0"]), code, int(n_therm), int(n_bins), int(sweeps_per_bin)),
            self._ctx)
        self._free[slots] = False
        return slots

    def release(self, slots) -> None:
        slots = _i32(slots)
        _hip.check(_hip.lib().rbim_pool_release(self._h, int(slots.size), _hip.ptr(slots)), self._ctx)
        self._free[slots] = True

    # ------------------------------------------------------------ running
    def run(self, n_sweeps: int) -> None:
        """Queue up to ``n_sweeps`` sweeps of every loaded, unfinished slot."""
        _hip.check(_hip.lib().rbim_pool_run(self._h, int(n_sweeps)), self._ctx)

    def idle(self) -> bool:
        flag = ctypes.c_int32()
        _hip.check(_hip.lib().rbim_pool_query(self._h, ctypes.byref(flag)), self._ctx)
        return bool(flag.value)

    def sync(self) -> None:
        _hip.check(_hip.lib().rbim_pool_sync(self._h), self._ctx)

    def progress(self, slots) -> Tuple[np.ndarray, np.ndarray]:
        """(sweeps queued, total sweeps) of each slot; -1 for slots not loaded."""
        slots = _i32(slots)
        done = np.empty(slots.size, np.int64)
        total = np.empty(slots.size, np.int64)
        _hip.check(_hip.lib().rbim_pool_progress(self._h, int(slots.size), _hip.ptr(slots),
                                                 _hip.ptr(done), _hip.ptr(total)), self._ctx)
        return done, total

    # ------------------------------------------------------------ results
    def moments(self, slots, n_bins: int) -> np.ndarray:
        """Moment sums (len(slots), n_bins, 20) of the slots (waits for the pool)."""
        slots = _i32(slots)
        out = np.zeros((slots.size, self.max_bins, _hip.ACC_SLOTS), dtype=np.uint64)
        _hip.check(_hip.lib().rbim_pool_read(self._h, int(slots.size), _hip.ptr(slots),
                                             _hip.ptr(out), None), self._ctx)
        return np.ascontiguousarray(out[:, :n_bins])

    def spins(self, slots) -> np.ndarray:
        slots = _i32(slots)
        out = np.empty((slots.size, self.L * self.L), dtype=np.int8)
        _hip.check(_hip.lib().rbim_pool_read(self._h, int(slots.size), _hip.ptr(slots), None,
                                             _hip.ptr(out)), self._ctx)
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            _hip.lib().rbim_pool_destroy(self._h)
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
