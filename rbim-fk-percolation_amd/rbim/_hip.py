"""ctypes binding of ``librbim_hip.so`` (C ABI declared in include/rbim_hip.h).

This replaces the reference's JIT boundary (``src/rbim/_unionfind.py:39-53``,
numba or plain Python): every sweep, label and measurement of this package runs
in the gfx950 kernels behind this library.  There is no CPU fallback -- if the
library or a GPU is missing, the first call raises :class:`EngineUnavailable`.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import numpy as np

__all__ = [
    "ACC_SLOTS",
    "ENGINE_AUTO",
    "ENGINE_PLAIN",
    "ENGINE_BATCHED",
    "ENGINE_TILED",
    "EngineError",
    "EngineUnavailable",
    "SweepStats",
    "BatchDesc",
    "BatchInfo",
    "PoolDesc",
    "library_path",
    "lib",
    "context",
    "check",
    "device_memory",
]

ACC_SLOTS = 20
ENGINE_AUTO, ENGINE_PLAIN, ENGINE_BATCHED, ENGINE_TILED = 0, 1, 2, 3
START_HOT, START_COLD = 0, 1

RBIM_OK = 0
RBIM_E_INVAL = -1
RBIM_E_HIP = -2
RBIM_E_OOM = -3
RBIM_E_ADMIT = -4
RBIM_E_STATE = -5

_HERE = os.path.dirname(os.path.abspath(__file__))


class EngineError(RuntimeError):
    """A call into librbim_hip.so failed; ``code`` is the RBIM_E_* status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class EngineUnavailable(EngineError):
    """The HIP library is not built or no GPU is visible."""


class SweepStats(ctypes.Structure):
    _fields_ = [
        ("energy", ctypes.c_int64),
        ("magnetization", ctypes.c_int64),
        ("n_clusters", ctypes.c_int64),
        ("largest_size", ctypes.c_int64),
        ("largest_abs_mag", ctypes.c_int64),
        ("second_size", ctypes.c_int64),
        ("second_abs_mag", ctypes.c_int64),
        ("wrap_x", ctypes.c_int32),
        ("wrap_y", ctypes.c_int32),
    ]


_P = ctypes.c_void_p


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("n_lattices", ctypes.c_int32),
        ("L", _P),
        ("Jh", _P),
        ("Jv", _P),
        ("thr31", _P),
        ("seed", _P),
        ("lattice_id", _P),
        ("spins0", _P),
        ("start", ctypes.c_int32),
        ("n_bins", ctypes.c_int32),
        ("sweeps_per_bin", ctypes.c_int64),
        ("n_therm", ctypes.c_int64),
        ("engine", ctypes.c_int32),
        ("sweeps_per_launch", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("streams", ctypes.c_int32),
        ("series", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("single_launch", ctypes.c_int32),
    ]


class PoolDesc(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_int32),
        ("n_slots", ctypes.c_int32),
        ("max_bins", ctypes.c_int32),
        ("engine", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("sweeps_per_launch", ctypes.c_int32),
        ("tile", ctypes.c_int32),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("n_groups", ctypes.c_int32),
        ("lds_groups", ctypes.c_int32),
        ("tiled_groups", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("sweeps_per_launch", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64),
        ("sweeps_done", ctypes.c_int64),
        ("total_sweeps", ctypes.c_int64),
        ("kernel_launches", ctypes.c_int64),
    ]


# Exported symbols and their signatures, in the order of include/rbim_hip.h.
_i32, _i64, _u32, _u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64
SIGNATURES = {
    "rbim_abi_version": ([], ctypes.c_int),
    "rbim_open": ([ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "rbim_close": ([_P], ctypes.c_int),
    "rbim_last_error": ([_P], ctypes.c_char_p),
    "rbim_device_memory": ([_P, _P, _P], ctypes.c_int),
    "rbim_label_and_wrap": ([_P, _i32, _P, _i32, _P, _P, _P, _P], ctypes.c_int),
    "rbim_activate": ([_P, _i32, _P, _P, _P, _u32, _u64, _u32, _u32, _P, _P], ctypes.c_int),
    "rbim_sweep_given_bits": ([_P, _i32, _P, _P, _P, _P, _P, _i32, _P, _P], ctypes.c_int),
    "rbim_sweep": ([_P, _i32, _P, _P, _P, _u32, _u64, _u32, _u32, _i32, _i32, _P], ctypes.c_int),
    "rbim_sweep_noalloc": ([_P, _i32, _P, _P, _P, _u32, _u64, _u32, _u32, _i32, _i32, _P],
                           ctypes.c_int),
    "rbim_cluster_statistics": ([_P, _i64, _P, _P, _i32, _P, _P], ctypes.c_int),
    "rbim_energy_magnetization": ([_P, _i32, _P, _P, _P, _P, _P], ctypes.c_int),
    "rbim_batch_footprint": ([_P, _P], ctypes.c_int),
    "rbim_batch_create": ([_P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rbim_batch_run": ([_P, _i64], ctypes.c_int),
    "rbim_batch_run_timed": ([_P, _i64, _P], ctypes.c_int),
    "rbim_batch_query": ([_P, _P], ctypes.c_int),
    "rbim_batch_sync": ([_P], ctypes.c_int),
    "rbim_batch_moments": ([_P, _P], ctypes.c_int),
    "rbim_batch_spins": ([_P, _P], ctypes.c_int),
    "rbim_batch_info_get": ([_P, _P], ctypes.c_int),
    "rbim_batch_series": ([_P, _P, _P], ctypes.c_int),
    "rbim_batch_destroy": ([_P], ctypes.c_int),
    "rbim_pool_footprint": ([_P, _P], ctypes.c_int),
    "rbim_pool_create": ([_P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rbim_pool_load": ([_P, _i32, _P, _P, _P, _P, _P, _P, _P, _i32, _i64, _i32, _i64],
                       ctypes.c_int),
    "rbim_pool_run": ([_P, _i64], ctypes.c_int),
    "rbim_pool_query": ([_P, _P], ctypes.c_int),
    "rbim_pool_sync": ([_P], ctypes.c_int),
    "rbim_pool_progress": ([_P, _i32, _P, _P, _P], ctypes.c_int),
    "rbim_pool_read": ([_P, _i32, _P, _P, _P], ctypes.c_int),
    "rbim_pool_release": ([_P, _i32, _P], ctypes.c_int),
    "rbim_pool_destroy": ([_P], ctypes.c_int),
    "rbim_inject_alloc_failure": ([_i64], ctypes.c_int),
}

_lib = None
_ctx = None
_lock = threading.Lock()


def library_path() -> str:
    return os.environ.get("RBIM_HIP_LIBRARY", os.path.join(_HERE, "librbim_hip.so"))


def lib():
    """Load librbim_hip.so (no GPU needed to load it)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise EngineUnavailable(
            RBIM_E_HIP,
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`",
        )
    handle = ctypes.CDLL(path)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = handle
    return handle


def _message(ctx) -> str:
    raw = lib().rbim_last_error(ctx)
    return raw.decode() if raw else "unknown error"


def check(status: int, ctx=None) -> None:
    """Raise the Python exception the reference would raise for this status."""
    if status == RBIM_OK:
        return
    message = _message(ctx)
    if status == RBIM_E_INVAL:
        raise ValueError(message)
    if status == RBIM_E_OOM:
        raise MemoryError(message)
    raise EngineError(status, message)


def context():
    """Process-wide device context on GPU 0 (opened on first use)."""
    global _ctx
    with _lock:
        if _ctx is None:
            handle = lib()
            ctx = _P()
            status = handle.rbim_open(int(os.environ.get("RBIM_DEVICE", "0")), ctypes.byref(ctx))
            if status != RBIM_OK:
                raise EngineUnavailable(status, _message(None))
            _ctx = ctx
        return _ctx


def device_memory():
    """(free, total) bytes of the context's GPU."""
    ctx = context()
    free, total = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().rbim_device_memory(ctx, ctypes.byref(free), ctypes.byref(total)), ctx)
    return int(free.value), int(total.value)


def ptr(array: Optional[np.ndarray]):
    if array is None:
        return None
    return array.ctypes.data_as(ctypes.c_void_p)
