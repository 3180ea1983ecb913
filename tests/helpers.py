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


def real_to_complex(c, norm):
    """Complex coefficients of the real field with real coefficients c:
    c_lm = (C_lm - i S_lm) f, c_l,-m = (-1)^m (C_lm + i S_lm) f for m > 0
    (f = 1/sqrt(2) for the 4pi, Schmidt and orthonormal functions, whose
    complex forms carry no sqrt(2); f = 1/2 unnormalised), c_l0 = C_l0.
    Layout [2][l][m]: [0] orders +m, [1] orders -m."""
    L = c.shape[1] - 1
    f = 0.5 if norm == 3 else 1.0 / np.sqrt(2.0)
    z = np.zeros((2, L + 1, L + 1), dtype=np.complex128)
    sgn = (-1.0) ** np.arange(L + 1)
    z[0] = (c[0] - 1j * c[1]) * f
    z[0][:, 0] = c[0][:, 0]
    z[1] = sgn[None, :] * (c[0] + 1j * c[1]) * f
    z[1][:, 0] = 0.0
    return z * np.tril(np.ones((L + 1, L + 1)))[None]


def rand_cilm_c(rng, L, norm=1, slope=-2.0):
    """Random complex coefficients: the complex forms of two independent real
    fields, a + i b (so every entry is free, entries m > l are zero)."""
    return real_to_complex(for_norm(rand_cilm(rng, L, slope), norm), norm) + \
        1j * real_to_complex(for_norm(rand_cilm(rng, L, slope), norm), norm)
