"""Generate brute_analysis.npz (run from the repo root):

    python tests/golden/make_brute_analysis.py

Analysis of grids that do not come from a synthesis (independent normal
values at every node), by direct quadrature in 40-digit mpmath arithmetic:

    c_lm = (1 / N_lm) sum_i w_i P_lm(cos t_i) (2 pi / nlon) sum_k g_ik {cos, sin}(m p_k)

with P_lm the normalised functions of make_golden.plm_exact (explicit
polynomials; the four normalisations and the Condon-Shortley phase),
N_lm = int Y_lm^2 dOmega of that normalisation (4 pi; 4 pi / (2l+1);
2 pi (1 + d_m0) (l+m)! / ((2l+1)(l-m)!); 1), p_k = 2 pi k / nlon, and the
latitude weights w_i of the grid:

* Driscoll-Healy (n = 2L+2 rows t_i = pi i / n, sampling 1 or 2): the
  weights that integrate sin(t) dt exactly for every Legendre polynomial
  P_l(cos t), l < n -- the n x n linear system sum_i w_i P_l(cos t_i) =
  2 delta_l0 solved in mpmath (not the closed form the transforms use);
* Gauss-Legendre (L+1 rows): roots of P_(L+1) and their weights
  (make_golden.gauss_legendre).

This is the analysis the reference defines (src/SHExpandDH.f95:287-294,
463-702; src/SHExpandGLQ.f95:445-689) written as the quadrature it
approximates, so a wrong weight, factor, phase, normalisation, sign of the
sine terms or ring order fails it.
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import gauss_legendre, plm_exact  # noqa: E402

mp.mp.dps = 40


def dh_weights(n):
    """Weights w_i at t_i = pi i / n (i < n) with sum_i w_i P_l(cos t_i) = 2 d_l0
    for l < n."""
    A = mp.matrix(n, n)
    b = mp.matrix(n, 1)
    for l in range(n):
        for i in range(n):
            A[l, i] = mp.legendre(l, mp.cos(mp.pi * i / n))
    b[0] = 2
    w = mp.lu_solve(A, b)
    return [w[i] for i in range(n)]


def norm_integral(l, m, norm):
    if norm == 1:
        return 4 * mp.pi
    if norm == 2:
        return 4 * mp.pi / (2 * l + 1)
    if norm == 3:
        return 2 * mp.pi * (2 if m == 0 else 1) * mp.factorial(l + m) / ((2 * l + 1) * mp.factorial(l - m))
    return mp.mpf(1)


def analyse(grid, thetas, weights, L, norm, cs):
    nlat, nlon = grid.shape
    # longitude sums per ring and order
    cs_sum = [[None] * (L + 1) for _ in range(nlat)]
    for i in range(nlat):
        row = [mp.mpf(float(x)) for x in grid[i]]
        for m in range(L + 1):
            a = mp.fsum(row[k] * mp.cos(2 * mp.pi * m * k / nlon) for k in range(nlon))
            b = mp.fsum(row[k] * mp.sin(2 * mp.pi * m * k / nlon) for k in range(nlon))
            cs_sum[i][m] = (a * 2 * mp.pi / nlon, b * 2 * mp.pi / nlon)
    out = np.zeros((2, L + 1, L + 1))
    for l in range(L + 1):
        for m in range(l + 1):
            sc = mp.mpf(0)
            ss = mp.mpf(0)
            for i in range(nlat):
                p = plm_exact(l, m, mp.cos(thetas[i]), norm, cs) * weights[i]
                sc += p * cs_sum[i][m][0]
                ss += p * cs_sum[i][m][1]
            nn = norm_integral(l, m, norm)
            out[0, l, m] = float(sc / nn)
            out[1, l, m] = float(ss / nn) if m > 0 else 0.0
    return out


def main():
    rng = np.random.default_rng(20261204)
    out = {}
    for L in (0, 1, 2, 5, 12):
        n = 2 * L + 2
        w_dh = dh_weights(n)
        t_dh = [mp.pi * i / n for i in range(n)]
        xs, w_glq = gauss_legendre(L + 1)
        t_glq = [mp.acos(x) for x in xs]
        for kind, (nlat, nlon, th, w) in {"dh1": (n, n, t_dh, w_dh), "dh2": (n, 2 * n, t_dh, w_dh),
                                          "glq": (L + 1, 2 * L + 1, t_glq, w_glq)}.items():
            g = rng.standard_normal((nlat, nlon))
            out[f"{kind}_L{L}_grid"] = g
            for norm in (1, 2, 3, 4):
                for cs in (1, -1):
                    out[f"{kind}_L{L}_n{norm}_cs{cs}"] = analyse(g, th, w, L, norm, cs)
            print(kind, L, flush=True)
    np.savez_compressed(os.path.join(HERE, "brute_analysis.npz"), **out)
    print("wrote", os.path.join(HERE, "brute_analysis.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
