"""+-J bond disorder: generation, ``.npz`` storage, frustration (host side).

Byte-compatible with the reference's files (``src/rbim/disorder.py``): a
realization is a pair of (L, L) int8 planes ``J_h``, ``J_v``, each -1 with
probability p; the draw order (``J_h`` then ``J_v`` from one generator) is kept
so that the same seed yields the same couplings as the reference.
"""

from __future__ import annotations

import os
from typing import Optional, Tuple

import numpy as np

__all__ = [
    "generate_disorder",
    "save_disorder",
    "load_disorder",
    "frustrated_plaquettes",
    "frustration_density",
    "default_disorder_filename",
]


def generate_disorder(
    L: int, p: float, rng: Optional[np.random.Generator] = None
) -> Tuple[np.ndarray, np.ndarray]:
    """One realization; ``rng.random((L, L)) < p`` marks antiferromagnetic bonds
    (disorder.py:32-61)."""
    if not 0.0 <= p <= 1.0:
        raise ValueError(f"p must lie in [0, 1], got {p}")
    generator = np.random.default_rng() if rng is None else rng
    planes = []
    for _ in range(2):  # J_h first, then J_v
        planes.append((1 - 2 * (generator.random((L, L)) < p)).astype(np.int8))
    return planes[0], planes[1]


def default_disorder_filename(L: int, p: float, index: int = 0) -> str:
    return f"disorder_L={L}_p={p:.3f}_r={index:04d}.npz"


def save_disorder(path: str, J_h: np.ndarray, J_v: np.ndarray, *, p: Optional[float] = None,
                  seed: Optional[int] = None) -> None:
    """Compressed ``.npz`` with ``J_h``, ``J_v`` and optional ``p`` / ``seed``."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    extra = {}
    if p is not None:
        extra["p"] = np.asarray(p)
    if seed is not None:
        extra["seed"] = np.asarray(seed)
    np.savez_compressed(path, J_h=J_h, J_v=J_v, **extra)


def load_disorder(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """Read ``J_h`` / ``J_v`` as int8; a file without them raises ``KeyError``."""
    with np.load(path) as data:
        if "J_h" not in data or "J_v" not in data:
            raise KeyError(
                f"{path} does not contain 'J_h' and 'J_v' arrays; found {list(data.keys())}"
            )
        return data["J_h"].astype(np.int8), data["J_v"].astype(np.int8)


def frustrated_plaquettes(J_h: np.ndarray, J_v: np.ndarray) -> np.ndarray:
    """Plaquette (x, y) is bounded by J_h[x, y], J_h[x, y+1], J_v[x, y],
    J_v[x+1, y] (periodic); it is frustrated when their product is negative."""
    product = (
        J_h.astype(np.int64)
        * np.roll(J_h, -1, axis=1).astype(np.int64)
        * J_v.astype(np.int64)
        * np.roll(J_v, -1, axis=0).astype(np.int64)
    )
    return product < 0


def frustration_density(J_h: np.ndarray, J_v: np.ndarray) -> float:
    """Fraction of frustrated plaquettes; 4p(1-p)[(1-p)^2 + p^2] on average."""
    return float(frustrated_plaquettes(J_h, J_v).mean())
