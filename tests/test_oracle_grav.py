"""Pins of the oracle's MakeGravGridDH, NormalGravity and SHMultiply restatements
(oracle/sht_oracle.c: or_makegravgriddh, or_normalgravity, or_shmultiply), independent of
the restated code paths:

* NormalGravity: the published WGS84 normal gravity at the equator and the
  poles (9.7803253359 and 9.8321849378 m/s^2) and the closed forms of the
  spherical branches;
* a point mass on a flattened ellipsoid: rad = -GM/r^2, pot = GM/r with the
  ellipsoid radius of each row, theta = phi = 0;
* single zonal and sectoral terms with r0 != a: the closed-form derivatives
  of V = GM/r (r0/r)^l C_lm Pbar_lm(cos theta) cos(m phi);
* the centrifugal terms of a point mass;
* random coefficients on a sphere (f = 0): pot, rad, theta and phi against
  the oracle's MakeGridDH and MakeGradientDH of the continued coefficients
  (r0/a)^l C_lm, which are pinned on their own in test_oracle.py;
* SHMultiply: the square of P_10 in every normalisation (closed form), the
  unit field as identity, exact commutation, the product's DH synthesis
  against the pointwise product of the factors' syntheses, and its table
  branch (precomp = 1) against the recursion branch.
"""
import numpy as np
import pytest

import oracle
from helpers import rand_cilm

GM, A, B_AX, OMEGA = 3.986004418e14, 6378137.0, 6356752.314245, 7.292115e-5


def test_normal_gravity_wgs84():
    assert abs(oracle.NormalGravity(0.0, GM, OMEGA, A, B_AX) - 9.7803253359) < 1e-9
    assert abs(oracle.NormalGravity(90.0, GM, OMEGA, A, B_AX) - 9.8321849378) < 1e-9
    assert abs(oracle.NormalGravity(-90.0, GM, OMEGA, A, B_AX) - 9.8321849378) < 1e-9


def test_normal_gravity_sphere():
    assert oracle.NormalGravity(30.0, GM, 0.0, A, A) == GM / A ** 2
    lat = 37.0
    c = np.cos(np.radians(lat))
    ref = np.sqrt(GM ** 2 / A ** 4 + A * OMEGA ** 2 * c * c * (A * OMEGA ** 2 - 2 * GM / A ** 2))
    assert abs(oracle.NormalGravity(lat, GM, OMEGA, A, A) - ref) <= 1e-15 * ref


def _rows(lmax, extend):
    n = 2 * lmax + 2
    nr = n + extend
    theta = np.pi * np.arange(nr) / n
    return n, theta


def _r_ellipsoid(theta, a, f):
    lat = np.pi / 2 - theta
    r = a / np.sqrt(np.cos(lat) ** 2 + np.sin(lat) ** 2 / (1 - f) ** 2)
    r[0] = a * (1 - f)
    return r


@pytest.mark.parametrize("sampling,extend", [(1, 0), (2, 1)])
def test_point_mass_on_ellipsoid(sampling, extend):
    lmax, f = 15, 1.0 / 298.257
    c = np.zeros((2, lmax + 1, lmax + 1))
    c[0, 0, 0] = 1.0
    rad, th, ph, tot, pot = oracle.MakeGravGridDH(c, GM, A, a=A, f=f, sampling=sampling, extend=extend,
                                                  normal_gravity=0)
    n, theta = _rows(lmax, extend)
    r = _r_ellipsoid(theta, A, f)
    if extend:
        r[-1] = A * (1 - f)
    r[n // 2] = A
    np.testing.assert_allclose(rad, np.broadcast_to((-GM / r ** 2)[:, None], rad.shape), rtol=1e-13)
    np.testing.assert_allclose(pot, np.broadcast_to((GM / r)[:, None], pot.shape), rtol=1e-13)
    assert np.abs(th).max() <= 1e-12 * GM / A ** 2 and np.abs(ph).max() <= 1e-12 * GM / A ** 2
    np.testing.assert_allclose(tot, np.abs(rad), rtol=1e-13)


def test_zonal_and_sectoral_closed_forms():
    lmax, r0, a = 12, 6.0e6, 6.3e6
    c = np.zeros((2, lmax + 1, lmax + 1))
    c20, c22, s22 = 1.3e-3, 4.0e-4, -2.0e-4
    c[0, 2, 0], c[0, 2, 2], c[1, 2, 2] = c20, c22, s22
    rad, th, ph, tot, pot = oracle.MakeGravGridDH(c, GM, r0, a=a, f=0.0, sampling=2, normal_gravity=0)
    n, theta = _rows(lmax, 0)
    phi = 2 * np.pi * np.arange(2 * n) / (2 * n)
    T, P = np.meshgrid(theta, phi, indexing="ij")
    z, s = np.cos(T), np.sin(T)
    p20 = np.sqrt(5.0) * (3 * z * z - 1) / 2
    p22 = np.sqrt(15.0) / 2 * s * s
    ang = c22 * np.cos(2 * P) + s22 * np.sin(2 * P)
    k = GM * r0 ** 2 / a ** 3
    pot_ref = k * (c20 * p20 + p22 * ang)
    rad_ref = -3 * k / a * (c20 * p20 + p22 * ang)
    th_ref = k / a * (c20 * (-3 * np.sqrt(5.0) * z * s) + np.sqrt(15.0) * s * z * ang)
    with np.errstate(divide="ignore", invalid="ignore"):
        ph_ref = k / a * p22 / s * 2 * (-c22 * np.sin(2 * P) + s22 * np.cos(2 * P))
    ph_ref[0] = 0.0
    th_ref[0] = 0.0
    scale = GM / a ** 2 * 1e-3
    assert np.abs(pot - pot_ref).max() <= 1e-13 * GM / a * 1e-3
    assert np.abs(rad - rad_ref).max() <= 1e-13 * scale
    assert np.abs(th - th_ref).max() <= 1e-12 * scale
    assert np.abs(ph - ph_ref).max() <= 1e-12 * scale


def test_centrifugal_point_mass():
    lmax, f = 10, 0.0
    c = np.zeros((2, lmax + 1, lmax + 1))
    c[0, 0, 0] = 1.0
    rad, th, ph, tot, pot = oracle.MakeGravGridDH(c, GM, A, a=A, f=f, omega=OMEGA, normal_gravity=0,
                                                  extend=1)
    n, theta = _rows(lmax, 1)
    s, co = np.sin(theta), np.cos(theta)
    np.testing.assert_allclose(rad[:, 0], -GM / A ** 2 + A * (s * OMEGA) ** 2, rtol=1e-14)
    np.testing.assert_allclose(th[:, 0], s * co * A * OMEGA ** 2, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(pot[:, 0], GM / A + 0.5 * (A * s * OMEGA) ** 2, rtol=1e-14)
    np.testing.assert_allclose(tot[:, 0], np.sqrt(rad[:, 0] ** 2 + th[:, 0] ** 2), rtol=1e-15)


@pytest.mark.parametrize("sampling,lmax_calc", [(2, None), (1, 9)])
def test_sphere_matches_gradient_and_synthesis(sampling, lmax_calc):
    """f = 0, a != r0: the continued coefficients c' = (r0/a)^l c give
    pot = GM/a S(c'), rad = -GM/a^2 S((l+1) c'), and theta, phi = GM/a times
    MakeGradientDH(c', radius=a)."""
    lmax, r0, a = 14, 6.1e6, 6.4e6
    rng = np.random.default_rng(11)
    c = rand_cilm(rng, lmax) * 1e-3
    c[0, 0, 0] = 1.0
    rad, th, ph, tot, pot = oracle.MakeGravGridDH(c, GM, r0, a=a, f=0.0, sampling=sampling,
                                                  lmax_calc=lmax_calc, normal_gravity=1)
    lc = lmax if lmax_calc is None else lmax_calc
    l = np.arange(lmax + 1, dtype=float)
    cc = c * ((r0 / a) ** l)[None, :, None]
    cc[:, lc + 1:, :] = 0.0
    g_u = oracle.MakeGridDH(cc, sampling=sampling, norm=1, csphase=1)
    g_r = oracle.MakeGridDH(cc * (l + 1)[None, :, None], sampling=sampling, norm=1, csphase=1)
    g_t, g_p = oracle.MakeGradientDH(cc, sampling=sampling, radius=a)
    sc = GM / a ** 2
    assert np.abs(pot - GM / a * g_u).max() <= 1e-13 * GM / a
    assert np.abs(rad + sc * g_r).max() <= 1e-13 * sc
    assert np.abs(th - GM / a * g_t).max() <= 1e-11 * sc
    assert np.abs(ph - GM / a * g_p).max() <= 1e-11 * sc
    ng = GM / a ** 2  # sphere, omega = 0: NormalGravity = GM / a^2
    np.testing.assert_allclose(tot, np.sqrt(rad ** 2 + th ** 2 + ph ** 2) - ng, rtol=0, atol=1e-13 * sc)


def test_errors():
    c = np.zeros((2, 5, 5))
    with pytest.raises(Exception):
        oracle.MakeGravGridDH(c, GM, A, sampling=3)
    with pytest.raises(Exception):
        oracle.MakeGravGridDH(c, GM, A, lmax_calc=9)
    with pytest.raises(Exception):
        oracle.MakeGravGridDH(c, GM, A, normal_gravity=2)


# ------------------------------------------------------------ SHMultiply

def test_shmultiply_zonal_square_closed_forms():
    """(P_10)^2 in each normalisation: 4pi 3 z^2 = 1 + (2/sqrt 5) Pbar_20;
    Schmidt and unnormalised z^2 = 1/3 + (2/3) P_20; orthonormal
    (3/4pi) z^2 = (1/sqrt(4pi)) Y_00 + (1/(2pi)) / sqrt(5/(4pi)) Y_20."""
    c = np.zeros((2, 2, 2))
    c[0, 1, 0] = 1.0
    want = {1: (1.0, 2 / np.sqrt(5.0)), 2: (1 / 3, 2 / 3), 3: (1 / 3, 2 / 3),
            4: (1 / np.sqrt(4 * np.pi), (1 / (2 * np.pi)) / np.sqrt(5 / (4 * np.pi)))}
    for norm, (c00, c20) in want.items():
        o = oracle.SHMultiply(c, c, norm=norm)
        assert o.shape == (2, 3, 3)
        ref = np.zeros((2, 3, 3))
        ref[0, 0, 0], ref[0, 2, 0] = c00, c20
        assert np.abs(o - ref).max() <= 1e-14, norm


def test_shmultiply_identity_and_commutation():
    rng = np.random.default_rng(21)
    f = rand_cilm(rng, 9)
    one = np.zeros((2, 4, 4))
    one[0, 0, 0] = 1.0
    o = oracle.SHMultiply(one, f)
    assert o.shape == (2, 13, 13)
    assert np.abs(o[:, :10, :10] - f).max() <= 1e-13 * np.abs(f).max()
    assert np.abs(o[:, 10:, :]).max() <= 1e-13 * np.abs(f).max()
    g = rand_cilm(rng, 5)
    np.testing.assert_array_equal(oracle.SHMultiply(f, g), oracle.SHMultiply(g, f))


@pytest.mark.parametrize("norm,csphase", [(1, 1), (2, -1), (4, 1)])
def test_shmultiply_is_the_pointwise_product(norm, csphase):
    """On a DH grid of degree >= lmax1 + lmax2 the product's synthesis equals
    the product of the syntheses (the oracle's MakeGridDH is pinned on its own)."""
    rng = np.random.default_rng(22 + norm)
    a, b = rand_cilm(rng, 7), rand_cilm(rng, 6)
    o = oracle.SHMultiply(a, b, norm=norm, csphase=csphase)
    L = 13
    pad = lambda c: np.pad(c, ((0, 0), (0, L + 1 - c.shape[1]), (0, L + 1 - c.shape[2])))
    ga = oracle.MakeGridDH(pad(a), sampling=2, norm=norm, csphase=csphase)
    gb = oracle.MakeGridDH(pad(b), sampling=2, norm=norm, csphase=csphase)
    go = oracle.MakeGridDH(pad(o), sampling=2, norm=norm, csphase=csphase)
    assert np.abs(go - ga * gb).max() <= 1e-12 * np.abs(ga * gb).max()


@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("csphase", [1, -1])
def test_shmultiply_table_branch_equals_recursion_branch(norm, csphase):
    """precomp = 1 (SHGLQ's table of the functions at the nodes, read by the
    transforms' plx branches; src/SHMultiply.f95:281-325) is the same product
    as precomp = 0: a table of the wrong degree, node order or normalisation,
    or a factor read with the wrong lmax, fails."""
    rng = np.random.default_rng(40 + norm)
    # the unnormalised functions overflow the product's degrees beyond ~15
    # (src/MakeGridDH.f95:17-18)
    cases = [(0, 0), (0, 3), (1, 0), (5, 3), (8, 6)] + ([] if norm == 3 else [(12, 9), (24, 17)])
    for l1, l2 in cases:
        a, b = rand_cilm(rng, l1), rand_cilm(rng, l2)
        o0 = oracle.SHMultiply(a, b, norm=norm, csphase=csphase)
        o1 = oracle.SHMultiply(a, b, norm=norm, csphase=csphase, precomp=1)
        assert o1.shape == o0.shape
        tol = 1e-12 if norm == 3 else 1e-13
        assert np.abs(o1 - o0).max() <= tol * np.abs(o0).max(), (l1, l2)


def test_shmultiply_errors():
    c = np.zeros((2, 3, 3))
    with pytest.raises(Exception):
        oracle.SHMultiply(c, c, lmax1=5)
    with pytest.raises(Exception):
        oracle.SHMultiply(c, c, norm=5)
    with pytest.raises(Exception):
        oracle.SHMultiply(c, c, csphase=0)
    with pytest.raises(Exception):
        oracle.SHMultiply(c, c, precomp=2)
