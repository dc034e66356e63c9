"""The acceptance rule of a scan against a reference result file: a
family-wise test of the z-scores of all its points (used by
tests/test_gpu_physics.py, pinned by tests/test_host.py)."""

import numpy as np

ALPHA = 1e-3  # family-wise significance of each file's comparison


def agreement(rows, alpha=ALPHA):
    """The three tests a scan's z-scores pass at family-wise significance
    alpha (each z ~ N(0, 1) under agreement, points taken as independent):
      beyond3  the number k of |z| > 3 is not in the upper alpha tail of
               Binomial(n, P(|z| > 3) = 0.0027);
      max      max |z| stays below the Bonferroni bound Phi^-1(1 - alpha / 2n);
      chi2     every observable's sum of z^2 over its temperatures is not in
               the upper alpha / 16 tail of chi^2 with that many degrees of
               freedom (Bonferroni over the 16 observables).
    Returns {test: (passed, statistic, threshold or p-value)}."""
    from scipy import stats

    z = np.array([r[4] for r in rows])
    n = z.size
    p3 = 2.0 * stats.norm.sf(3.0)
    k = int(np.sum(np.abs(z) > 3.0))
    p_k = float(stats.binom.sf(k - 1, n, p3))  # P(X >= k)
    zmax = float(np.abs(z).max())
    bound = float(stats.norm.isf(alpha / (2.0 * n)))
    names = sorted({r[0] for r in rows})
    worst = 1.0
    for name in names:
        zz = np.array([r[4] for r in rows if r[0] == name])
        worst = min(worst, float(stats.chi2.sf(np.sum(zz ** 2), zz.size)))
    return {"beyond3": (p_k >= alpha, k, p_k),
            "max": (zmax < bound, zmax, bound),
            "chi2": (worst >= alpha / len(names), worst, alpha / len(names))}
