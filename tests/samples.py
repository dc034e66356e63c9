"""Configuration samplers shared by the oracle and GPU parity tests.

Like the reference's ``tests/test_wrapping.py:67-80``: seeded disorder, random
spins, a few Swendsen-Wang sweeps to reach typical configurations, then one
activation.  The sweeps and the activation run in the C oracle, so the
samples are the same on any machine.
"""

import numpy as np

from conftest import random_disorder


def sample_configurations(oracle, sizes, disorders, temperatures, n_samples, seed=0):
    """Yield dicts with L, p, T, J_h, J_v, spins, active (2N bool), flips (N)."""
    rng = np.random.default_rng(seed)
    for L in sizes:
        for p in disorders:
            J_h, J_v = random_disorder(rng, L, p)
            for T in temperatures:
                thr = oracle.threshold31(1.0 / T)
                key = int(rng.integers(0, 2**63))
                spins = np.where(rng.random(L * L) < 0.5, -1, 1).astype(np.int8)
                for k in range(n_samples):
                    spins, _ = oracle.c_run(L, spins, J_h, J_v, thr, key, 0, k, 1, 10**6, 1, 1)
                    active, flips = oracle.c_activate(spins, J_h, J_v, thr, key, 0, 10**5 + k)
                    yield {"L": L, "p": p, "T": T, "J_h": J_h, "J_v": J_v, "spins": spins.copy(),
                           "active": active, "flips": flips, "thr": thr, "key": key}


def column_case(L):
    """Vertical bonds of column x = 0 except the boundary one (test_clusters.py:94-111)."""
    active = np.zeros(2 * L * L, dtype=bool)
    for y in range(L - 1):
        active[2 * y + 1] = True
    closed = active.copy()
    closed[2 * (L - 1) + 1] = True
    return active, closed


def diagonal_loop(L):
    """Staircase loop (0,0)->(1,0)->(1,1)->...->(L-1,L-1)->(0,L-1)->(0,0): it
    crosses each boundary once and winds (1, 1)."""
    active = np.zeros(2 * L * L, dtype=bool)
    for k in range(L):
        n = k * L + k
        active[2 * n] = True  # (k, k) -> (k+1, k)
        if k < L - 1:
            m = ((k + 1) % L) * L + k
            active[2 * m + 1] = True  # (k+1, k) -> (k+1, k+1)
    active[2 * (L - 1) + 1] = True  # (0, L-1) -> (0, 0)
    return active


def double_winding(L):
    """A single cluster made of two parallel columns joined only through the
    periodic boundary twice: columns x = 0 and x = 2 closed along y, linked by
    horizontal bonds at y = 0 across (0->1->2).  Winds along y; the boundary-
    deletion shortcut also sees it.  Used as a second known-answer case."""
    active = np.zeros(2 * L * L, dtype=bool)
    for col in (0, 2):
        for y in range(L):
            active[2 * (col * L + y) + 1] = True
    active[2 * (0 * L + 0)] = True
    active[2 * (1 * L + 0)] = True
    return active
