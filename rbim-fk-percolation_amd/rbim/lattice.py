"""Periodic L x L square-lattice geometry (host side).

Same conventions as the reference (``src/rbim/lattice.py:1-24``): site
``n = x * L + y`` with x the slow index; bond ``2n`` joins n to (x+1, y) and
bond ``2n+1`` joins n to (x, y+1); ``J_h[x, y]`` sits on bond 2n and
``J_v[x, y]`` on bond 2n+1.  The GPU engine never materialises these bond
arrays -- its kernels derive neighbours arithmetically (``n + L`` / ``n + 1``
with the periodic wrap) and keep ``J_h`` / ``J_v`` as two coalesced int8 planes.
The arrays here serve callers of the host API and the tests.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Lattice", "build_lattice", "site_index", "couplings_from_disorder"]


@dataclass(frozen=True)
class Lattice:
    """Bond structure of a periodic L x L lattice (lattice.py:35-81).

    ``bonds`` (2N, 2) holds bond endpoints, ``wraps_x`` / ``wraps_y`` mark the
    bonds closing the boundary along each axis, ``bond_dx`` / ``bond_dy`` are
    the unwrapped displacements (1, 0) / (0, 1) of every bond.
    """

    L: int
    bonds: np.ndarray
    wraps_x: np.ndarray
    wraps_y: np.ndarray
    bond_dx: np.ndarray
    bond_dy: np.ndarray

    @property
    def n_sites(self) -> int:
        return self.L * self.L

    @property
    def n_bonds(self) -> int:
        return 2 * self.L * self.L

    def wrap_mask(self, axis: str) -> np.ndarray:
        if axis not in ("x", "y"):
            raise ValueError(f"axis must be 'x' or 'y', got {axis!r}")
        return self.wraps_x if axis == "x" else self.wraps_y


def site_index(x, y, L: int):
    """Flat site index ``x * L + y``."""
    return x * L + y


def build_lattice(L: int) -> Lattice:
    """Geometry of the periodic L x L lattice; L < 3 is rejected (lattice.py:89-135)."""
    if L < 3:
        raise ValueError(
            f"L must be at least 3 (got {L}); smaller lattices produce duplicate "
            "bonds under periodic boundary conditions."
        )
    sites = np.arange(L * L, dtype=np.int64)
    x, y = np.divmod(sites, L)
    ends = np.empty((L * L, 2, 2), dtype=np.int64)
    ends[:, :, 0] = sites[:, None]
    ends[:, 0, 1] = np.where(x == L - 1, y, sites + L)
    ends[:, 1, 1] = np.where(y == L - 1, sites - y, sites + 1)
    bonds = ends.reshape(2 * L * L, 2)
    kind = np.tile(np.array([0, 1]), L * L)  # 0 horizontal, 1 vertical
    wraps_x = (kind == 0) & np.repeat(x == L - 1, 2)
    wraps_y = (kind == 1) & np.repeat(y == L - 1, 2)
    return Lattice(
        L=L,
        bonds=bonds,
        wraps_x=wraps_x,
        wraps_y=wraps_y,
        bond_dx=(kind == 0).astype(np.int64),
        bond_dy=(kind == 1).astype(np.int64),
    )


def couplings_from_disorder(J_h: np.ndarray, J_v: np.ndarray) -> np.ndarray:
    """Interleave ``J_h``, ``J_v`` into the 2N bond order (lattice.py:138-161)."""
    if J_h.shape != J_v.shape:
        raise ValueError(f"J_h and J_v must have the same shape, got {J_h.shape} and {J_v.shape}")
    if J_h.ndim != 2 or J_h.shape[0] != J_h.shape[1]:
        raise ValueError(f"couplings must be square 2D arrays, got shape {J_h.shape}")
    return np.stack([J_h.reshape(-1), J_v.reshape(-1)], axis=1).reshape(-1).astype(np.int8)


def split_couplings(J: np.ndarray):
    """Inverse of :func:`couplings_from_disorder`: (J_h, J_v) planes of a 2N vector."""
    J = np.asarray(J)
    L = int(round(np.sqrt(J.size // 2)))
    if 2 * L * L != J.size:
        raise ValueError(f"coupling vector of length {J.size} is not 2 L^2")
    pairs = J.reshape(L * L, 2)
    return (np.ascontiguousarray(pairs[:, 0], dtype=np.int8).reshape(L, L),
            np.ascontiguousarray(pairs[:, 1], dtype=np.int8).reshape(L, L))
