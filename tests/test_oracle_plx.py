"""The oracle's GLQ transforms with a precomputed table of Legendre functions
(the plx argument of SHGLQ, MakeGridGLQ and SHExpandGLQ; src/SHGLQ.f95:202-257,
src/MakeGridGLQ.f95:441-465, src/SHExpandGLQ.f95:419-441), pinned on the CPU:

* the oracle's table against the table of the reference's own SHGLQ, compiled
  from its sources (oracle/Makefile, target `ref`), for every normalisation,
  phase and cnorm;
* the transforms that read the table against the transforms that recur the
  functions on the fly (themselves pinned by closed forms and brute-force
  sums in test_oracle.py): the same linear maps, so a misplaced index, a
  transposed table or a missing factor of the table's branch fails;
* the exit codes of the branch's argument checks.
No GPU."""
import numpy as np
import pytest

import oracle
from helpers import coef_err, for_norm, grid_err, rand_cilm, rand_cilm_c

NEEDS_REF = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference helpers not built (oracle/_ref)")
# the oracle's functions recur as the transforms do; PlmSchmidt recurs
# differently (relative to each degree's largest function)
TABLE_TOL = {1: 1e-13, 2: 4e-12, 3: 2e-14, 4: 5e-13}


def _by_degree_err(a, b, L):
    """max over (l, m, node) of |a - b| / max_m,node |b| of degree l."""
    worst = 0.0
    for l in range(L + 1):
        k = l * (l + 1) // 2
        sc = max(np.abs(b[k : k + l + 1]).max(), 1e-300)
        worst = max(worst, float(np.abs(a[k : k + l + 1] - b[k : k + l + 1]).max() / sc))
    return worst


@NEEDS_REF
@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("cnorm", [0, 1])
@pytest.mark.parametrize("cs", [1, -1])
def test_table_matches_reference_shglq(norm, cnorm, cs):
    for L in ((0, 1, 5, 12) if norm == 3 else (0, 1, 2, 12, 61, 255)):
        zero, w, ref, st = oracle.ref_SHGLQ_plx(L, norm, cs, cnorm)
        assert st == 0
        z0, w0 = oracle.SHGLQ(L)
        np.testing.assert_array_equal(zero, z0)
        np.testing.assert_array_equal(w, w0)
        mine = oracle.plx_table(L, zero, norm, cs, cnorm)
        assert _by_degree_err(mine, ref, L) <= TABLE_TOL[norm], (L, norm, cs, cnorm)
        # the southern nodes are the northern ones times (-1)^(l-m), exactly
        n = L + 1
        for l in range(L + 1):
            for m in range(l + 1):
                k = l * (l + 1) // 2 + m
                np.testing.assert_array_equal(mine[k, ::-1][: n // 2], mine[k, : n // 2] * (-1.0) ** (l - m))


@NEEDS_REF
def test_complex_normalisation_divides_by_sqrt2():
    """cnorm = 1: the functions of orders m >= 1 without the real
    normalisation's sqrt(2) (the reference's table, and the oracle's)."""
    L = 40
    zero, _, re, _ = oracle.ref_SHGLQ_plx(L, 1, 1, 0)
    _, _, cx, _ = oracle.ref_SHGLQ_plx(L, 1, 1, 1)
    ms = np.concatenate([np.arange(l + 1) for l in range(L + 1)])
    ratio = np.where(ms[:, None] == 0, 1.0, np.sqrt(2.0))
    assert np.abs(cx * ratio - re).max() <= 1e-13 * np.abs(re).max()
    mine = oracle.plx_table(L, zero, 1, 1, 1)
    assert np.abs(mine - cx).max() <= 1e-13 * np.abs(cx).max()


@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("cs", [1, -1])
def test_transforms_with_table_equal_on_the_fly(norm, cs):
    rng = np.random.default_rng(10 * norm + (cs > 0))
    for L in ((0, 1, 2, 7, 12) if norm == 3 else (0, 1, 2, 7, 30, 101)):
        c = for_norm(rand_cilm(rng, L), norm)
        zero, w = oracle.SHGLQ(L)
        plx = oracle.plx_table(L, zero, norm, cs)
        for ext in (0, 1):
            g1 = oracle.MakeGridGLQ(c, zero, norm=norm, csphase=cs, extend=ext)
            g2 = oracle.MakeGridGLQ_plx(c, plx, norm=norm, csphase=cs, extend=ext)
            assert grid_err(g2, g1) <= 1e-13, (L, ext)
        g = g1[:, : 2 * L + 1]
        c1 = oracle.SHExpandGLQ(g, w, zero, norm=norm, csphase=cs)
        c2 = oracle.SHExpandGLQ_plx(g, w, plx, norm=norm, csphase=cs)
        assert coef_err(c2, c1) <= (1e-12 if norm == 3 else 1e-13), L
        # lmax_calc and a grid that comes from no synthesis
        if L >= 2:
            gr = rng.standard_normal(g.shape)
            lc = L // 2
            c1 = oracle.SHExpandGLQ(gr, w, zero, norm=norm, csphase=cs, lmax_calc=lc)
            c2 = oracle.SHExpandGLQ_plx(gr, w, plx, norm=norm, csphase=cs, lmax_calc=lc)
            assert coef_err(c2, c1) <= (1e-12 if norm == 3 else 1e-13), L
            g1 = oracle.MakeGridGLQ(c, zero, norm=norm, csphase=cs, lmax_calc=lc)
            g2 = oracle.MakeGridGLQ_plx(c, plx, norm=norm, csphase=cs, lmax_calc=lc)
            assert grid_err(g2, g1) <= 1e-13, L


@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("cs", [1, -1])
def test_complex_transforms_with_table_equal_on_the_fly(norm, cs):
    """MakeGridGLQC / SHExpandGLQC with a table of the complex normalisation
    (cnorm = 1) against their on-the-fly branches."""
    rng = np.random.default_rng(20 * norm + (cs > 0))
    for L in ((0, 1, 2, 7, 12) if norm == 3 else (0, 1, 2, 7, 40)):
        c = rand_cilm_c(rng, L, norm)
        zero, w = oracle.SHGLQ(L)
        plx = oracle.plx_table(L, zero, norm, cs, 1)
        for ext in (0, 1):
            g1 = oracle.MakeGridGLQC(c, zero, norm=norm, csphase=cs, extend=ext)
            g2 = oracle.MakeGridGLQC_plx(c, plx, norm=norm, csphase=cs, extend=ext)
            assert grid_err(g2, g1) <= 1e-13, (L, ext)
        g = g1[:, : 2 * L + 1]
        c1 = oracle.SHExpandGLQC(g, w, zero, norm=norm, csphase=cs)
        c2 = oracle.SHExpandGLQC_plx(g, w, plx, norm=norm, csphase=cs)
        assert coef_err(c2, c1) <= (1e-12 if norm == 3 else 1e-13), L
        if L >= 2:
            gr = rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)
            lc = L // 2
            c1 = oracle.SHExpandGLQC(gr, w, zero, norm=norm, csphase=cs, lmax_calc=lc)
            c2 = oracle.SHExpandGLQC_plx(gr, w, plx, norm=norm, csphase=cs, lmax_calc=lc)
            assert coef_err(c2, c1) <= (1e-12 if norm == 3 else 1e-13), L


def test_plx_branch_exit_codes():
    L = 4
    c = np.zeros((2, L + 1, L + 1))
    zero, w = oracle.SHGLQ(L)
    plx = oracle.plx_table(L, zero)
    lib = oracle.lib()
    g = np.zeros((L + 1, 2 * L + 2))
    P = oracle._p
    # extend, grid dimensions, no table, norm, csphase, lmax_calc
    assert lib.or_makegridglq_plx(P(c), L + 1, L + 1, L, P(plx), 1, 1, L, 2, P(g), L + 1, 2 * L + 2) == 2
    assert lib.or_makegridglq_plx(P(c), L + 1, L + 1, L, P(plx), 1, 1, L, 1, P(g), L + 1, 2 * L + 1) == 1
    assert lib.or_makegridglq_plx(P(c), L + 1, L + 1, L, None, 1, 1, L, 0, P(g), L + 1, 2 * L + 2) == 5
    assert lib.or_makegridglq_plx(P(c), L + 1, L + 1, L, P(plx), 5, 1, L, 0, P(g), L + 1, 2 * L + 2) == 2
    assert lib.or_makegridglq_plx(P(c), L + 1, L + 1, L, P(plx), 1, 0, L, 0, P(g), L + 1, 2 * L + 2) == 2
    assert lib.or_makegridglq_plx(P(c), L + 1, L + 1, L, P(plx), 1, 1, L + 1, 0, P(g), L + 1, 2 * L + 2) == 2
    g = np.zeros((L + 1, 2 * L + 1))
    # lmax_calc, cilm and grid dimensions, weights, no table, norm, csphase
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L + 1, L, P(w), L + 1, P(plx), 1, 1, L + 1, P(c), L + 1,
                                  L + 1) == 2
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L + 1, L, P(w), L + 1, P(plx), 1, 1, L, P(c), L, L + 1) == 1
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L, L, P(w), L + 1, P(plx), 1, 1, L, P(c), L + 1, L + 1) == 1
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L + 1, L, P(w), L, P(plx), 1, 1, L, P(c), L + 1, L + 1) == 1
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L + 1, L, P(w), L + 1, None, 1, 1, L, P(c), L + 1, L + 1) == 5
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L + 1, L, P(w), L + 1, P(plx), 0, 1, L, P(c), L + 1,
                                  L + 1) == 2
    assert lib.or_shexpandglq_plx(P(g), L + 1, 2 * L + 1, L, P(w), L + 1, P(plx), 1, 2, L, P(c), L + 1,
                                  L + 1) == 2
