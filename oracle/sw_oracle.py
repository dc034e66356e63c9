"""CPU restatement of the reference's Swendsen-Wang path, in numpy / Python.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP engine; the
shipped ``rbim`` package never imports it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may use it.

Functions and the reference text each one follows:

* :func:`philox4x32_10` -- Philox4x32-10 (Salmon et al., SC'11), vectorised.
  Pinned by Random123 known-answer vectors and by rocRAND's host engine
  (``tests/golden/philox_vectors.json``).
* :func:`geometry` -- ``src/rbim/lattice.py:105-126`` (bond ends, wrap masks,
  unwrapped bond displacements).
* :func:`label_and_wrap` -- ``src/rbim/_unionfind.py:59-136`` line for line
  (pure Python; small lattices only).
* :func:`brute_force_winding` -- ``tests/test_wrapping.py:22-64`` (BFS spanning
  forest, every non-tree bond).
* :func:`cover_winding` -- independent of both: a cluster winds along an axis
  iff some site is connected to one of its own translates in the m-fold cover of
  the torus along that axis (checked for m = 2, 3, 5 with scipy's
  ``connected_components``).
* :func:`activate` -- ``src/rbim/clusters.py:48-79`` with the engine's Philox
  draws in place of PCG64.
* :func:`sweep_given_bits` -- ``src/rbim/swendsen_wang.py:103-172`` given the
  occupation and flip bits; :func:`two_largest` fixes the tie rule that the
  reference leaves to numpy's introselect (``swendsen_wang.py:92-100``): equal
  sizes order by the smaller minimum site.
* :func:`kaufman_energy` -- exact finite-L torus energy (Kaufman 1949, Ferdinand
  and Fisher 1969); :func:`onsager_energy` -- the L -> infinity limit.
* :func:`load_c` -- ctypes view of ``oracle/_build/libsw_oracle.so``, the C
  restatement of the same sweep (``oracle/sw_oracle.c``).
"""

from __future__ import annotations

import ctypes
import os
from collections import deque

import numpy as np

SLOTS = 20
_HERE = os.path.dirname(os.path.abspath(__file__))

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = np.uint64(0xFFFFFFFF)


# --------------------------------------------------------------------- Philox
def philox4x32_10(ctr, key):
    """Philox4x32-10 on arrays of counters.

    ``ctr`` is ``(..., 4)`` uint32-valued, ``key`` is ``(k0, k1)``.  Returns a
    ``(..., 4)`` uint32 array.
    """
    c = np.asarray(ctr, dtype=np.uint64) & MASK32
    c0, c1, c2, c3 = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n1 = p1 & MASK32
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        n3 = p0 & MASK32
        c0, c1, c2, c3 = n0, n1, n2, n3
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def sweep_draws(seed: int, lattice_id: int, sweep: int, n_sites: int):
    """Per-site (horizontal draw word, vertical draw word) for one sweep."""
    n_pairs = (n_sites + 1) // 2
    ctr = np.zeros((n_pairs, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(n_pairs)
    ctr[:, 1] = sweep
    ctr[:, 2] = lattice_id
    out = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    dh = np.empty(2 * n_pairs, dtype=np.uint32)
    dv = np.empty(2 * n_pairs, dtype=np.uint32)
    dh[0::2], dv[0::2], dh[1::2], dv[1::2] = out[:, 0], out[:, 1], out[:, 2], out[:, 3]
    return dh[:n_sites], dv[:n_sites]


def init_spins(L: int, seed: int, lattice_id: int, start: str = "hot") -> np.ndarray:
    n = L * L
    if start == "cold":
        return np.ones(n, dtype=np.int8)
    blocks = (n + 127) // 128
    ctr = np.zeros((blocks, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(blocks)
    ctr[:, 2] = lattice_id
    ctr[:, 3] = 1
    words = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32)).reshape(-1)
    site = np.arange(n)
    bits = (words[site >> 5] >> (site & 31).astype(np.uint32)) & 1
    return np.where(bits == 1, 1, -1).astype(np.int8)


def threshold31(beta: float) -> int:
    """round(p_add * 2^31) with p_add = -expm1(-2 beta) (clusters.py:40-45)."""
    p_add = float(-np.expm1(-2.0 * beta))
    return int(np.floor(p_add * 2.0**31 + 0.5))


# ------------------------------------------------------------------- geometry
def geometry(L: int):
    """Bond ends, wrap masks and unwrapped displacements (lattice.py:105-126)."""
    x = np.repeat(np.arange(L), L)
    y = np.tile(np.arange(L), L)
    n = x * L + y
    bonds = np.empty((2 * L * L, 2), dtype=np.int64)
    bonds[0::2, 0] = n
    bonds[0::2, 1] = ((x + 1) % L) * L + y
    bonds[1::2, 0] = n
    bonds[1::2, 1] = x * L + (y + 1) % L
    dx = np.zeros(2 * L * L, dtype=np.int64)
    dy = np.zeros(2 * L * L, dtype=np.int64)
    dx[0::2] = 1
    dy[1::2] = 1
    wraps_x = np.zeros(2 * L * L, dtype=bool)
    wraps_y = np.zeros(2 * L * L, dtype=bool)
    wraps_x[0::2] = x == L - 1
    wraps_y[1::2] = y == L - 1
    return bonds, dx, dy, wraps_x, wraps_y


# ----------------------------------------------------------------- union-find
def _find(parent, rx, ry, start):
    root, ax, ay = start, 0, 0
    while parent[root] != root:
        ax += rx[root]
        ay += ry[root]
        root = parent[root]
    node, bx, by = start, 0, 0
    while parent[node] != node:
        nxt = parent[node]
        ox, oy = rx[node], ry[node]
        parent[node] = root
        rx[node] = ax - bx
        ry[node] = ay - by
        bx += ox
        by += oy
        node = nxt
    return root, ax, ay


def label_and_wrap(active, L):
    """(n_clusters, labels in min-site order, wrap_x, wrap_y)."""
    bonds, bdx, bdy, _, _ = geometry(L)
    n_sites = L * L
    parent = list(range(n_sites))
    rx = [0] * n_sites
    ry = [0] * n_sites
    wrap_x = wrap_y = False
    for k in np.flatnonzero(active):
        a, b = int(bonds[k, 0]), int(bonds[k, 1])
        ex, ey = int(bdx[k]), int(bdy[k])
        ra, ax, ay = _find(parent, rx, ry, a)
        rb, bx, by = _find(parent, rx, ry, b)
        if ra != rb:
            parent[rb] = ra
            rx[rb] = ax - ex - bx
            ry[rb] = ay - ey - by
        else:
            if ex + bx - ax != 0:
                wrap_x = True
            if ey + by - ay != 0:
                wrap_y = True
    labels = np.empty(n_sites, dtype=np.int64)
    remap = {}
    for site in range(n_sites):
        root, _, _ = _find(parent, rx, ry, site)
        if root not in remap:
            remap[root] = len(remap)
        labels[site] = remap[root]
    return len(remap), labels, wrap_x, wrap_y


def brute_force_winding(active, L):
    """BFS spanning forest, then the winding of every non-tree bond."""
    bonds, bdx, bdy, _, _ = geometry(L)
    n_sites = L * L
    neighbours = [[] for _ in range(n_sites)]
    for k in np.flatnonzero(active):
        a, b = int(bonds[k, 0]), int(bonds[k, 1])
        dx, dy = int(bdx[k]), int(bdy[k])
        neighbours[a].append((b, dx, dy, int(k)))
        neighbours[b].append((a, -dx, -dy, int(k)))
    position = {}
    tree = set()
    for start in range(n_sites):
        if start in position:
            continue
        position[start] = (0, 0)
        queue = deque([start])
        while queue:
            node = queue.popleft()
            px, py = position[node]
            for nb, dx, dy, k in neighbours[node]:
                if nb not in position:
                    position[nb] = (px + dx, py + dy)
                    tree.add(k)
                    queue.append(nb)
    wx = wy = False
    for k in np.flatnonzero(active):
        k = int(k)
        if k in tree:
            continue
        a, b = int(bonds[k, 0]), int(bonds[k, 1])
        if position[a][0] + int(bdx[k]) - position[b][0] != 0:
            wx = True
        if position[a][1] + int(bdy[k]) - position[b][1] != 0:
            wy = True
    return wx, wy


def cover_winding(active, L, moduli=(2, 3, 5)):
    """Wrapping by connectivity in m-fold covers of the torus (independent check).

    In the cover that unrolls axis ``x`` m times, a site ``(x, y)`` has copies
    ``(x + jL, y)``.  A cycle of x-winding w lifts to a path from a copy to the
    copy shifted by ``w mod m``, so some cluster winds along x iff, for some m
    not dividing the gcd of its windings, a site is connected to another copy
    of itself.
    """
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    bonds, _, _, wraps_x, wraps_y = geometry(L)
    act = np.flatnonzero(active)
    result = []
    for axis, wraps in (("x", wraps_x), ("y", wraps_y)):
        found = False
        for m in moduli:
            n = L * L
            a = bonds[act, 0]
            b = bonds[act, 1]
            crosses = wraps[act].astype(np.int64)
            rows, cols = [], []
            for j in range(m):
                rows.append(a + j * n)
                cols.append(b + ((j + crosses) % m) * n)
            rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
            cols = np.concatenate(cols) if cols else np.zeros(0, np.int64)
            g = csr_matrix((np.ones(rows.size, np.int8), (rows, cols)), shape=(m * n, m * n))
            _, lab = connected_components(g, directed=False)
            lab = lab.reshape(m, n)
            if np.any(lab[0] == lab[1:]).item() if m > 1 else False:
                found = True
                break
        result.append(found)
    return tuple(result)


def canonical(labels):
    """Per-site minimum site index of its cluster (the engine's canonical label)."""
    labels = np.asarray(labels)
    n = labels.size
    first = np.full(int(labels.max()) + 1, n, dtype=np.int64)
    np.minimum.at(first, labels, np.arange(n))
    return first[labels]


# ----------------------------------------------------------------- activation
def activate(spins, Jh, Jv, thr31, seed, lattice_id, sweep):
    """(active[2N] bond order, flipbits[N]) -- clusters.py:48-79 with Philox draws."""
    L = Jh.shape[0]
    bonds, _, _, _, _ = geometry(L)
    J = np.empty(2 * L * L, dtype=np.int64)
    J[0::2] = Jh.reshape(-1)
    J[1::2] = Jv.reshape(-1)
    s = np.asarray(spins, dtype=np.int64)
    satisfied = J * s[bonds[:, 0]] * s[bonds[:, 1]] > 0
    dh, dv = sweep_draws(seed, lattice_id, sweep, L * L)
    draws = np.empty(2 * L * L, dtype=np.uint32)
    draws[0::2] = dh
    draws[1::2] = dv
    active = satisfied & ((draws >> 1) < thr31)
    return active, (dh & 1).astype(np.uint8)


# -------------------------------------------------- sweep given the bits
def two_largest(sizes, minsites):
    """Indices of the largest and second-largest cluster; ties -> smaller min site."""
    order = np.lexsort((minsites, -np.asarray(sizes)))
    first = int(order[0])
    second = int(order[1]) if len(order) > 1 else None
    return first, second


def sweep_given_bits(spins, Jh, Jv, active, flip_by_minsite):
    """One SW update given occupation and flip bits (swendsen_wang.py:103-172).

    Returns (new_spins int8, stats dict, labels, canonical labels).
    """
    L = Jh.shape[0]
    n_clusters, labels, wx, wy = label_and_wrap(active, L)
    canon = canonical(labels)
    flip = np.asarray(flip_by_minsite, dtype=bool)[canon]
    s = np.asarray(spins, dtype=np.int64).copy()
    s[flip] *= -1
    sizes = np.bincount(labels, minlength=n_clusters)
    mags = np.bincount(labels, weights=s.astype(np.float64), minlength=n_clusters)
    minsites = np.full(n_clusters, L * L, dtype=np.int64)
    np.minimum.at(minsites, labels, np.arange(L * L))
    first, second = two_largest(sizes, minsites)
    bonds, _, _, _, _ = geometry(L)
    J = np.empty(2 * L * L, dtype=np.int64)
    J[0::2] = Jh.reshape(-1)
    J[1::2] = Jv.reshape(-1)
    stats = {
        "energy": int(-np.sum(J * s[bonds[:, 0]] * s[bonds[:, 1]])),
        "magnetization": int(s.sum()),
        "n_clusters": int(n_clusters),
        "largest_size": int(sizes[first]),
        "largest_abs_mag": int(abs(mags[first])),
        "second_size": 0 if second is None else int(sizes[second]),
        "second_abs_mag": 0 if second is None else int(abs(mags[second])),
        "wrap_x": bool(wx),
        "wrap_y": bool(wy),
    }
    return s.astype(np.int8), stats, labels, canon


# --------------------------------------------------------- exact energies
def kaufman_energy(L: int, T: float) -> float:
    """Exact energy per site of the L x L periodic Ising ferromagnet (J = 1).

    Kaufman's four-term partition function (Ferdinand and Fisher 1969, eq. 2.1):
    ``Z = 1/2 (2 sinh 2K)^{N/2} (Z1 + Z2 + Z3 + Z4)`` with products over
    ``gamma_{2r+1}`` (Z1 cosh, Z2 sinh) and ``gamma_{2r}`` (Z3 cosh, Z4 sinh),
    ``cosh gamma_l = cosh 2K coth 2K - cos(pi l / L)`` and
    ``gamma_0 = 2K + ln tanh K``.  The energy is ``-d ln Z / dK / N``, taken by
    a central difference at 50 significant digits.
    """
    import mpmath as mp

    mp.mp.dps = 50

    def log_z(K):
        n = L
        c = mp.cosh(2 * K) * mp.coth(2 * K)

        def gamma(l):
            if l == 0:
                return 2 * K + mp.log(mp.tanh(K))
            return mp.acosh(c - mp.cos(mp.pi * l / n))

        z = [mp.mpf(1)] * 4
        for r in range(n):
            g_odd = gamma(2 * r + 1)
            g_even = gamma(2 * r)
            z[0] *= 2 * mp.cosh(L * g_odd / 2)
            z[1] *= 2 * mp.sinh(L * g_odd / 2)
            z[2] *= 2 * mp.cosh(L * g_even / 2)
            z[3] *= 2 * mp.sinh(L * g_even / 2)
        total = z[0] + z[1] + z[2] + z[3]
        return mp.log(mp.mpf(1) / 2) + (L * L) / mp.mpf(2) * mp.log(2 * mp.sinh(2 * K)) + mp.log(total)

    K = mp.mpf(1) / mp.mpf(T)
    h = mp.mpf(10) ** -15
    dlnz = (log_z(K + h) - log_z(K - h)) / (2 * h)
    return float(-dlnz / (L * L))


def onsager_energy(T: float) -> float:
    """Infinite-lattice energy per site (Onsager 1944)."""
    from scipy.special import ellipk

    K = 1.0 / T
    k = 2.0 * np.sinh(2 * K) / np.cosh(2 * K) ** 2
    return float(-1.0 / np.tanh(2 * K) * (1 + 2 / np.pi * (2 * np.tanh(2 * K) ** 2 - 1) * ellipk(k * k)))


# ---------------------------------------------------------- C restatement
class OracleStats(ctypes.Structure):
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


_C = None


def load_c():
    """ctypes handle on oracle/_build/libsw_oracle.so (built by ``make -C oracle``)."""
    global _C
    if _C is not None:
        return _C
    path = os.path.join(_HERE, "_build", "libsw_oracle.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing; run `make -C oracle`")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i64, u32, u64 = ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64
    lib.oracle_philox4x32_10.argtypes = [P, P, P]
    lib.oracle_label_and_wrap.argtypes = [i64, P, P, P, P]
    lib.oracle_label_and_wrap.restype = i64
    lib.oracle_activate.argtypes = [i64, P, P, P, u32, u64, u32, u32, P, P]
    lib.oracle_sweep_given_bits.argtypes = [i64, P, P, P, P, P, ctypes.POINTER(OracleStats), P]
    lib.oracle_init_spins.argtypes = [i64, u64, u32, ctypes.c_int, P]
    lib.oracle_run.argtypes = [i64, P, P, P, u32, u64, u32, i64, i64, i64, i64, i64, P]
    _C = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def c_label_and_wrap(active, L):
    lib = load_c()
    act = np.ascontiguousarray(active, dtype=np.uint8)
    labels = np.empty(L * L, dtype=np.int64)
    wx, wy = ctypes.c_int32(), ctypes.c_int32()
    n = lib.oracle_label_and_wrap(L, _ptr(act), _ptr(labels), ctypes.byref(wx), ctypes.byref(wy))
    return int(n), labels, bool(wx.value), bool(wy.value)


def c_activate(spins, Jh, Jv, thr31, seed, lattice_id, sweep):
    lib = load_c()
    L = Jh.shape[0]
    s = np.ascontiguousarray(spins, dtype=np.int8)
    jh = np.ascontiguousarray(Jh, dtype=np.int8)
    jv = np.ascontiguousarray(Jv, dtype=np.int8)
    active = np.empty(2 * L * L, dtype=np.uint8)
    flips = np.empty(L * L, dtype=np.uint8)
    lib.oracle_activate(L, _ptr(s), _ptr(jh), _ptr(jv), thr31, seed, lattice_id, sweep,
                        _ptr(active), _ptr(flips))
    return active.astype(bool), flips


def c_sweep_given_bits(spins, Jh, Jv, active, flip_by_minsite):
    lib = load_c()
    L = Jh.shape[0]
    s = np.array(spins, dtype=np.int8, copy=True)
    jh = np.ascontiguousarray(Jh, dtype=np.int8)
    jv = np.ascontiguousarray(Jv, dtype=np.int8)
    act = np.ascontiguousarray(active, dtype=np.uint8)
    fl = np.ascontiguousarray(flip_by_minsite, dtype=np.uint8)
    st = OracleStats()
    labels = np.empty(L * L, dtype=np.int64)
    lib.oracle_sweep_given_bits(L, _ptr(s), _ptr(jh), _ptr(jv), _ptr(act), _ptr(fl),
                                ctypes.byref(st), _ptr(labels))
    stats = {name: getattr(st, name) for name, _ in OracleStats._fields_}
    stats["wrap_x"] = bool(stats["wrap_x"])
    stats["wrap_y"] = bool(stats["wrap_y"])
    return s, stats, labels


def c_init_spins(L, seed, lattice_id, start="hot"):
    lib = load_c()
    s = np.empty(L * L, dtype=np.int8)
    lib.oracle_init_spins(L, seed, lattice_id, 1 if start == "cold" else 0, _ptr(s))
    return s


def c_run(L, spins, Jh, Jv, thr31, seed, lattice_id, g0, n_sweeps, n_therm, spb, n_bins, acc=None):
    """Run sweeps g0..g0+n_sweeps-1 of one lattice; returns (spins, acc[n_bins, SLOTS])."""
    lib = load_c()
    s = np.array(spins, dtype=np.int8, copy=True)
    jh = np.ascontiguousarray(Jh, dtype=np.int8)
    jv = np.ascontiguousarray(Jv, dtype=np.int8)
    if acc is None:
        acc = np.zeros((n_bins, SLOTS), dtype=np.uint64)
    lib.oracle_run(L, _ptr(s), _ptr(jh), _ptr(jv), thr31, seed, lattice_id, g0, n_sweeps,
                   n_therm, spb, n_bins, _ptr(acc))
    return s, acc
