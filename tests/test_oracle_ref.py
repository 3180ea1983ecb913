"""The oracle and the library's host helpers against the reference's own
Fortran helper routines, compiled from their sources in place
(oracle/Makefile, target `ref`; skipped where that build is absent).

DHaj, SHGLQ and GLQGridCoord are pinned bitwise; the Legendre functions of
the transforms' recursion (oracle.plm, and the tables the GPU kernels
recur with, evaluated on the host by shtg_plm) against PlmBar, PlmSchmidt,
PLegendreA and PlmON, and near the poles at high degree against a 50-digit
evaluation, where the reference's double recursion itself is off by about
1e-12.  No GPU."""
import math

import numpy as np
import pytest

import oracle
from shtools_amd import _lib
from shtools_amd import backend as B

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference helpers not built (oracle/_ref)")


def _unpack(p, L):
    out = np.zeros((L + 1, L + 1))
    for l in range(L + 1):
        k = l * (l + 1) // 2
        out[l, : l + 1] = p[k : k + l + 1]
    return out


def test_dhaj_bitwise_equal_to_reference():
    for n in (2, 4, 8, 64, 130, 256, 1024, 4096, 5600):
        for ext in (0, 1):
            ref, st = oracle.ref_DHaj(n, ext)
            assert st == 0
            np.testing.assert_array_equal(oracle.DHaj(n, ext), ref)
            np.testing.assert_array_equal(B.DHaj(n, ext), ref)
    # odd n: the reference's exit status 2, the library's "Improper bounds"
    assert oracle.ref_DHaj(5)[1] == 2
    with pytest.raises(B.SHToolsError, match="Improper bounds"):
        B.DHaj(5)


def test_gauss_legendre_nodes_bitwise_equal_to_reference():
    for L in (0, 1, 2, 5, 63, 200, 1023, 2047):
        z, w, st = oracle.ref_SHGLQ(L)
        assert st == 0
        zo, wo = oracle.SHGLQ(L)
        zb, wb = B.SHGLQ(L)
        for a, b in ((zo, z), (wo, w), (zb, z), (wb, w)):
            np.testing.assert_array_equal(a, b)


def test_glq_coordinates_match_reference():
    for L in (0, 1, 5, 17, 1023):
        for ext in (0, 1):
            lat, lon, nlat, nlong, st = oracle.ref_GLQGridCoord(L, ext)
            # nlong counts the distinct longitudes (the extended column repeats one)
            assert st == 0 and nlat == L + 1 and nlong == 2 * L + 1
            for a, b in (oracle.GLQGridCoord(L, ext), B.GLQGridCoord(L, ext)):
                np.testing.assert_array_equal(a, lat)
                np.testing.assert_array_equal(b, lon)


def test_normal_gravity_matches_reference():
    gm, omega, a, b = 3.986004418e14, 7.292115e-5, 6378137.0, 6356752.314140347
    for lat in (-90.0, -61.3, -12.0, 0.0, 0.5, 33.3, 89.99, 90.0):
        ref = oracle.ref_NormalGravity(lat, gm, omega, a, b)
        assert abs(oracle.NormalGravity(lat, gm, omega, a, b) - ref) <= 2e-15 * ref
        assert abs(B.NormalGravity(lat, gm, omega, a, b) - ref) <= 2e-15 * ref


ZS = (0.0, 0.3, -0.77, 0.999, 1 - 1e-6, -0.99999, 1.0, -1.0, math.cos(math.pi / 5600))
# oracle.plm runs the transforms' recursion (src/MakeGridDH.f95:321-381);
# PlmSchmidt recurs differently, so norm 2 differs by more rounding
ORACLE_TOL = {1: 1e-13, 2: 4e-12, 3: 2e-14, 4: 5e-13}


@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("cs", [1, -1])
def test_legendre_functions_match_reference(norm, cs):
    lib = _lib.lib()
    for L in ((12, 20) if norm == 3 else (100, 1023)):
        for z in ZS:
            p, st = oracle.ref_plm(L, z, norm, cs)
            assert st == 0
            R = _unpack(p, L)
            scale = np.maximum(np.abs(R).max(axis=1), 1e-300)[:, None]
            O = np.tril(oracle.plm(L, z, norm, cs))
            assert (np.abs(O - R) / scale).max() <= ORACLE_TOL[norm], (L, z)
            G = np.zeros((L + 1, L + 1))
            assert lib.shtg_plm(L, norm, cs, float(z), B._dptr(G)) == 0
            # near the poles both carry the recursion's own rounding
            # (about 1e-12 of the degree's largest function at L = 1023)
            assert (np.abs(np.tril(G) - R) / scale).max() <= 1e-11, (L, z)


def _truth(z, m, L):
    """4pi-normalised P_lm(z), l = m..L, by the exact-coefficient recursion
    at 50 digits."""
    import mpmath as mp
    with mp.workdps(50):
        z = mp.mpf(z)
        u = mp.sqrt((1 - z) * (1 + z))
        pmm = mp.mpf(1)
        if m > 0:
            pmm = mp.sqrt(2)
            for k in range(1, m + 1):
                pmm *= mp.sqrt(mp.mpf(2 * k + 1) / (2 * k)) * u
        out = {m: pmm, m + 1: mp.sqrt(2 * m + 3) * z * pmm}
        for l in range(m + 2, L + 1):
            a = mp.sqrt(mp.mpf((2 * l - 1) * (2 * l + 1)) / ((l - m) * (l + m)))
            b = mp.sqrt(mp.mpf((2 * l + 1) * (l - m - 1) * (l + m - 1)) / ((2 * l - 3) * (l + m) * (l - m)))
            out[l] = a * z * out[l - 1] - b * out[l - 2]
        return {l: float(v) for l, v in out.items()}


@pytest.mark.parametrize("z", [1 - 1e-6, math.cos(math.pi / 5600)])
def test_recursion_tables_near_poles_as_accurate_as_reference(z):
    """At degrees 1500..2800 and a few orders, the tables' functions are no
    further from a 50-digit evaluation than the reference's own PlmBar (by
    the rms over the degrees, relative to each degree's largest function).
    Tables rounded along the way in double came out about 2.5x worse."""
    L = 2800
    G = np.zeros((L + 1, L + 1))
    assert _lib.lib().shtg_plm(L, 1, 1, float(z), B._dptr(G)) == 0
    p, st = oracle.ref_plm(L, z, 1, 1)
    assert st == 0
    ours, ref = [], []
    for m in (0, 1, 10, 100):
        t = _truth(z, m, L)
        for l in range(1500, L + 1, 5):
            k = l * (l + 1) // 2
            sc = np.abs(p[k : k + l + 1]).max()
            ours.append((G[l, m] - t[l]) / sc)
            ref.append((p[k + m] - t[l]) / sc)
    rms = lambda e: float(np.sqrt(np.mean(np.square(e))))
    assert rms(ours) <= 1.2 * rms(ref), (rms(ours), rms(ref))
    assert max(np.abs(ours)) <= 1e-11


def test_nglq_counts_equal_to_reference():
    """NGLQ, NGLQSH, NGLQSHN of the library against the reference's own
    (src/PreGLQ.f95:153-233), and the count's meaning: that many
    Gauss-Legendre nodes integrate z^degree exactly, one fewer does not."""
    lib = oracle.ref_lib()
    for d in list(range(0, 200)) + [511, 1023, 2047, 2799, 4000]:
        assert B.NGLQ(d) == lib.ref_nglq(d)
        assert B.NGLQSH(d) == lib.ref_nglqsh(d)
        for n in range(0, 6):
            assert B.NGLQSHN(d, n) == lib.ref_nglqshn(d, n), (d, n)
    for d in range(0, 40):
        for npt, exact in ((B.NGLQ(d), True), (B.NGLQ(d) - 1, False)):
            if npt < 1:
                continue
            x, w = np.polynomial.legendre.leggauss(npt)
            val = float(np.sum(w * x ** d))
            ref = 0.0 if d % 2 else 2.0 / (d + 1)
            if exact:
                assert abs(val - ref) <= 1e-13, (d, npt)
            elif d % 2 == 0:
                assert abs(val - ref) > 1e-12, (d, npt)
    with pytest.raises(B.SHToolsError):
        B.NGLQ(-1)
