"""Shared test helpers: seeded synthetic spectra and norm-wise errors."""
import numpy as np


def rand_cilm(rng, L, slope=-2.0):
    """Real coefficients with Var(C_lm) = (l+1)^slope / (2l+1), S_l0 = 0 and
    zeros for m > l (the synthetic spectra of the benchmark configurations)."""
    c = rng.standard_normal((2, L + 1, L + 1))
    l = np.arange(L + 1)
    c *= np.sqrt((l + 1.0) ** slope / (2 * l + 1.0))[None, :, None]
    c[:, np.triu_indices(L + 1, 1)[0], np.triu_indices(L + 1, 1)[1]] = 0.0
    c[1, :, 0] = 0.0
    return c


def for_norm(c, norm):
    """Re-express 4pi-normalised coefficients for the unnormalised functions
    (norm 3), C_unnorm = C_4pi sqrt((2-d_m0)(2l+1)(l-m)!/(l+m)!), so that the
    field and its grid keep an O(1) range; other norms are returned as is."""
    if norm != 3:
        return c
    from math import lgamma
    L = c.shape[1] - 1
    f = np.zeros((L + 1, L + 1))
    for l in range(L + 1):
        for m in range(l + 1):
            f[l, m] = np.sqrt((1.0 if m == 0 else 2.0) * (2 * l + 1) * np.exp(lgamma(l - m + 1) - lgamma(l + m + 1)))
    return c * f[None]


def grid_err(a, b):
    """max |a - b| / max |b| (the grid parity measure)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def coef_err(a, b):
    """||a - b||_2 / ||b||_2 (the coefficient parity measure)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
