"""Generate the golden fixtures in this folder (run from the repo root):

    python tests/golden/make_golden.py

Everything here comes from mathematics evaluated in high precision with
mpmath, not from any implementation of the transforms:

* ``plm_*``   associated Legendre functions P_lm(mu) from the explicit
  polynomial P_l(mu) = 2^-l sum_k (-1)^k C(l,k) C(2l-2k,l) mu^(l-2k),
  differentiated m times, times (1-mu^2)^(m/2), with the four SHTOOLS
  normalisations of docs/pages/mydoc/real-spherical-harmonics.md:131-150 and
  the Condon-Shortley phase (-1)^m when csphase = -1;
* ``plmhi``  the same at high degree from the textbook recurrence of the
  unnormalised functions, (l-m) P_l^m = (2l-1) x P_(l-1)^m - (l+m-1) P_(l-2)^m
  from P_m^m = (2m-1)!! (1-x^2)^(m/2), in 60-digit arithmetic with mpmath's
  unbounded exponent range (no scaling), times the normalisation factor;
* ``glq_*``   Gauss-Legendre nodes (roots of P_n) and weights
  2 / ((1-x^2) P_n'(x)^2) by mpmath root finding;
* ``bf_*``    brute-force synthesis sum_lm P_lm(cos t)(C cos m p + S sin m p)
  of small random expansions at every node of DH (sampling 1 and 2, with
  the extra row and column of ``extend``) and GLQ grids.
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def legendre_poly_coeffs(l):
    """Coefficients c[k] of P_l(x) = sum c[k] x^k, exact rationals as mpf."""
    c = [mp.mpf(0)] * (l + 1)
    for k in range(l // 2 + 1):
        c[l - 2 * k] += mp.mpf((-1) ** k) * mp.binomial(l, k) * mp.binomial(2 * l - 2 * k, l) / mp.mpf(2) ** l
    return c


def plm_exact(l, m, x, norm, csphase):
    c = legendre_poly_coeffs(l)
    # m-th derivative
    d = mp.mpf(0)
    for k in range(m, l + 1):
        if c[k] != 0:
            d += c[k] * mp.ff(k, m) * mp.mpf(x) ** (k - m)
    x = mp.mpf(x)
    val = (1 - x * x) ** (mp.mpf(m) / 2) * d
    if csphase == -1:
        val *= (-1) ** m
    return val * norm_factor(l, m, norm)


def norm_factor(l, m, norm):
    r = mp.factorial(l - m) / mp.factorial(l + m)
    two = 1 if m == 0 else 2
    if norm == 1:
        return mp.sqrt(two * (2 * l + 1) * r)
    if norm == 2:
        return mp.sqrt(two * r)
    if norm == 3:
        return mp.mpf(1)
    return mp.sqrt(two * (2 * l + 1) * r / (4 * mp.pi))


def plm_recur(l, m, x):
    """Unnormalised P_l^m(x), no Condon-Shortley phase, by the textbook
    recurrence in l (mpmath's exponent range is unbounded)."""
    x = mp.mpf(x)
    pmm = mp.fac2(2 * m - 1) * (1 - x * x) ** (mp.mpf(m) / 2) if m > 0 else mp.mpf(1)
    if l == m:
        return pmm
    p1 = (2 * m + 1) * x * pmm
    p0 = pmm
    for ll in range(m + 2, l + 1):
        p0, p1 = p1, ((2 * ll - 1) * x * p1 - (ll + m - 1) * p0) / (ll - m)
    return p1


def gauss_legendre(n):
    xs, ws = [], []
    for i in range(1, n + 1):
        x0 = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (n + mp.mpf(1) / 2))
        x = x0
        for _ in range(200):   # Newton on P_n with P_n' = n (x P_n - P_(n-1)) / (x^2 - 1)
            dp = n * (x * mp.legendre(n, x) - mp.legendre(n - 1, x)) / (x * x - 1)
            dx = mp.legendre(n, x) / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-mp.mp.dps + 5):
                break
        dp = n * (x * mp.legendre(n, x) - mp.legendre(n - 1, x)) / (x * x - 1)
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    return xs, ws


def synth_point(cilm, norm, csphase, theta, phi):
    L = cilm.shape[1] - 1
    x = mp.cos(theta)
    s = mp.mpf(0)
    for l in range(L + 1):
        for m in range(l + 1):
            p = plm_exact(l, m, x, norm, csphase)
            s += p * (mp.mpf(cilm[0, l, m]) * mp.cos(m * phi) + mp.mpf(cilm[1, l, m]) * mp.sin(m * phi))
    return s


def random_cilm(rng, L):
    c = rng.standard_normal((2, L + 1, L + 1))
    for l in range(L + 1):
        c[:, l, l + 1:] = 0.0
    c[1, :, 0] = 0.0
    return c


def main():
    out = {}
    # low-degree P_lm on a set of z values, all norms / phases
    L = 24
    zs = [1.0, 0.9, 0.5, 0.1234, 0.0, -0.3, -0.77, -1.0, float(np.cos(np.pi / 64)), float(np.cos(np.pi * 31 / 64))]
    out["plm_z"] = np.array(zs)
    out["plm_lmax"] = np.array(L)
    for norm in (1, 2, 3, 4):
        for cs in (1, -1):
            a = np.zeros((len(zs), L + 1, L + 1))
            for iz, z in enumerate(zs):
                for l in range(L + 1):
                    for m in range(l + 1):
                        a[iz, l, m] = float(plm_exact(l, m, z, norm, cs))
            out[f"plm_n{norm}_cs{cs}"] = a

    # high-degree spot values (4pi and orthonormalised), exercising the
    # extended-range scaling of the recursion
    mp.mp.dps = 60
    for m in range(6):   # the recurrence agrees with the explicit polynomial
        a = plm_exact(7, m, 0.37, 3, 1)
        b = plm_recur(7, m, 0.37)
        assert abs(a - b) < 1e-40 * max(1, abs(a)), (m, a, b)
    hi = []
    for (l, m, z) in [(300, 0, 0.3), (300, 150, 0.3), (300, 299, 0.999), (1000, 1, 0.7),
                      (1000, 500, 0.2), (1000, 999, 0.05), (1000, 700, 0.95), (2047, 1024, 0.5),
                      (2047, 10, 0.999999), (2047, 2000, 0.01), (2799, 1400, 0.6), (2799, 2799, 0.0),
                      (2799, 5, -0.8)]:
        pr = plm_recur(l, m, z)
        for norm in (1, 4):
            hi.append((l, m, z, norm, float(pr * norm_factor(l, m, norm))))
    out["plmhi"] = np.array(hi)
    mp.mp.dps = 40

    # Gauss-Legendre nodes and weights
    for n in (1, 2, 3, 4, 5, 8, 16, 33, 64):
        xs, ws = gauss_legendre(n)
        out[f"glq_x_{n}"] = np.array([float(x) for x in xs])
        out[f"glq_w_{n}"] = np.array([float(w) for w in ws])

    # brute-force synthesis on small grids
    rng = np.random.default_rng(20240611)
    L = 5
    cilm = random_cilm(rng, L)
    out["bf_cilm"] = cilm
    n = 2 * L + 2
    for norm in (1, 2, 3, 4):
        for cs in (1, -1):
            for sampling in (1, 2):
                nlon = sampling * n
                g = np.zeros((n + 1, nlon + 1))
                for i in range(n + 1):
                    th = mp.pi * i / n
                    for k in range(nlon + 1):
                        ph = 2 * mp.pi * k / nlon
                        g[i, k] = float(synth_point(cilm, norm, cs, th, ph))
                out[f"bf_dh_n{norm}_cs{cs}_s{sampling}"] = g
    for Lg in (5, 6):
        cg = cilm[:, :Lg + 1, :Lg + 1] if Lg <= L else None
        if cg is None:
            cg = np.zeros((2, Lg + 1, Lg + 1))
            cg[:, :L + 1, :L + 1] = cilm
            cg[0, 6, 3] = 0.25
            cg[1, 6, 6] = -0.5
        out[f"bf_glq_cilm_{Lg}"] = cg
        xs, _ = gauss_legendre(Lg + 1)
        nlon = 2 * Lg + 1
        for norm in (1, 4):
            for cs in (1, -1):
                g = np.zeros((Lg + 1, nlon + 1))
                for i, x in enumerate(xs):
                    th = mp.acos(x)
                    for k in range(nlon + 1):
                        ph = 2 * mp.pi * k / nlon
                        g[i, k] = float(synth_point(cg, norm, cs, th, ph))
                out[f"bf_glq{Lg}_n{norm}_cs{cs}"] = g
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
