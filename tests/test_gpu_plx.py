"""The plx branch of the GLQ routines on the GPU: SHGLQ's table of Legendre
functions at the nodes (shtg_shglq_plx, src/SHGLQ.f95:202-257) and the
transforms that read it (shtg_makegrid_glq_plx, shtg_shexpand_glq_plx;
src/MakeGridGLQ.f95:441-465, src/SHExpandGLQ.f95:419-441).

* The device table against the reference's own SHGLQ table (compiled from
  its sources, where that build exists) and the oracle's, per degree
  relative to the degree's largest function.
* The transforms against the oracle's plx branch on the same table (the
  device table copied to the host, and a table the caller made: the
  oracle's), grids max|d|/max|ref| <= 1e-12, coefficients ||d||/||ref|| <=
  1e-12, for every normalisation, phase, extend, lmax_calc and a grid that
  comes from no synthesis; the degree-0 field; at high degree against the
  on-the-fly transforms of the library.
* The exit statuses of the branch through the C ABI."""
import numpy as np
import pytest

import oracle
from helpers import coef_err, for_norm, grid_err, rand_cilm, rand_cilm_c

pytestmark = pytest.mark.gpu

TOL = 1e-12
# the device table recurs with the library's rescaled recursion, the
# oracle's and the reference's with the reference's: per degree relative to
# its largest function (PlmSchmidt recurs differently again)
TABLE_TOL = {1: 1e-12, 2: 4e-12, 3: 1e-13, 4: 1e-12}


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd import backend
    return backend


def _by_degree_err(a, b, L):
    worst = 0.0
    for l in range(L + 1):
        k = l * (l + 1) // 2
        sc = max(np.abs(b[k : k + l + 1]).max(), 1e-300)
        worst = max(worst, float(np.abs(a[k : k + l + 1] - b[k : k + l + 1]).max() / sc))
    return worst


@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("cnorm", [0, 1])
@pytest.mark.parametrize("cs", [1, -1])
def test_table_against_reference_and_oracle(B, norm, cnorm, cs):
    for L in ((0, 1, 6, 12) if norm == 3 else (0, 1, 2, 6, 64, 255)):
        print(f"case plx table L={L} norm={norm} cnorm={cnorm} cs={cs}", flush=True)
        zero, w, plx = B.SHGLQ_plx(L, norm=norm, csphase=cs, cnorm=cnorm)
        z0, w0 = oracle.SHGLQ(L)
        np.testing.assert_array_equal(zero, z0)
        np.testing.assert_array_equal(w, w0)
        t = plx.cpu().numpy()
        assert _by_degree_err(t, oracle.plx_table(L, zero, norm, cs, cnorm), L) <= TABLE_TOL[norm]
        if oracle.ref_lib() is not None:
            _, _, ref, st = oracle.ref_SHGLQ_plx(L, norm, cs, cnorm)
            assert st == 0
            assert _by_degree_err(t, ref, L) <= TABLE_TOL[norm]


def test_table_high_degree(B):
    """L = 1023: near the poles both tables carry their recursion's
    rounding (about 1e-12 of each degree's largest function)."""
    L = 1023
    for norm in (1, 4):
        zero, _, plx = B.SHGLQ_plx(L, norm=norm)
        t = plx.cpu().numpy()
        ref = (oracle.ref_SHGLQ_plx(L, norm, 1, 0)[2] if oracle.ref_lib() is not None
               else oracle.plx_table(L, zero, norm))
        assert _by_degree_err(t, ref, L) <= 1e-11


PLX_CASES = [(L, norm, cs, e) for L in (0, 1, 2, 7, 12, 31, 64, 127) for norm in (1, 2, 3, 4) for cs in (1, -1)
             for e in (0, 1) if not (norm == 3 and L > 12)]


@pytest.mark.parametrize("L,norm,cs,extend", PLX_CASES)
def test_transforms_against_oracle(B, L, norm, cs, extend):
    print(f"case plx L={L} norm={norm} cs={cs} e={extend}", flush=True)
    rng = np.random.default_rng(3000 + 10 * L + norm)
    c = for_norm(rand_cilm(rng, L), norm)
    zero, w, plx = B.SHGLQ_plx(L, norm=norm, csphase=cs)
    t = plx.cpu().numpy()
    g_ref = oracle.MakeGridGLQ_plx(c, t, norm=norm, csphase=cs, extend=extend)
    g = B.MakeGridGLQ_plx(c, plx, norm=norm, csphase=cs, extend=extend)
    assert g.shape == g_ref.shape
    assert grid_err(g, g_ref) <= TOL
    gin = g_ref[:, : 2 * L + 1]
    c_ref = oracle.SHExpandGLQ_plx(gin, w, t, norm=norm, csphase=cs)
    c2 = B.SHExpandGLQ_plx(gin, w, plx, norm=norm, csphase=cs)
    assert c2.shape == c_ref.shape
    assert coef_err(c2, c_ref) <= (TOL if norm != 3 else 1e-10)
    if L >= 2:
        # lmax_calc, and a grid that comes from no synthesis
        lc = L // 2
        gr = rng.standard_normal(gin.shape)
        assert coef_err(B.SHExpandGLQ_plx(gr, w, plx, norm=norm, csphase=cs, lmax_calc=lc),
                        oracle.SHExpandGLQ_plx(gr, w, t, norm=norm, csphase=cs, lmax_calc=lc)) <= TOL
        assert grid_err(B.MakeGridGLQ_plx(c, plx, norm=norm, csphase=cs, lmax_calc=lc),
                        oracle.MakeGridGLQ_plx(c, t, norm=norm, csphase=cs, lmax_calc=lc)) <= TOL


def test_callers_own_table_and_padded_arrays(B):
    """A table the caller made (the oracle's, uploaded), coefficients in a
    larger array than the grid's degree, and a truncated analysis."""
    import torch
    rng = np.random.default_rng(7)
    L = 40
    zero, w = oracle.SHGLQ(L)
    t = oracle.plx_table(L, zero, 4, -1)
    c = rand_cilm(rng, L + 9)
    g = B.MakeGridGLQ_plx(c, torch.from_numpy(t).cuda(), lmax=L, norm=4, csphase=-1)
    g_ref = oracle.MakeGridGLQ_plx(c, t, lmax=L, norm=4, csphase=-1)
    assert grid_err(g, g_ref) <= TOL
    c2 = B.SHExpandGLQ_plx(g_ref, w, t, norm=4, csphase=-1, lmax_calc=L - 5)
    assert coef_err(c2, oracle.SHExpandGLQ_plx(g_ref, w, t, norm=4, csphase=-1, lmax_calc=L - 5)) <= TOL


def test_degree_zero_field(B):
    for norm in (1, 4):
        c = np.zeros((2, 5, 5))
        c[0, 0, 0] = 2.5
        _, _, plx = B.SHGLQ_plx(4, norm=norm)
        g = B.MakeGridGLQ_plx(c, plx, lmax=4, norm=norm, lmax_calc=0, extend=1)
        g_ref = oracle.MakeGridGLQ_plx(c, plx.cpu().numpy(), lmax=4, norm=norm, lmax_calc=0, extend=1)
        np.testing.assert_array_equal(g, g_ref)


@pytest.mark.parametrize("L", [511, 1023])
def test_high_degree_against_on_the_fly(B, L):
    """The table's transforms against the library's recursion on the fly
    (the same maps; both pinned against the oracle), and round trips."""
    rng = np.random.default_rng(L)
    c = rand_cilm(rng, L)
    zero, w, plx = B.SHGLQ_plx(L, norm=1)
    g = B.MakeGridGLQ_plx(c, plx)
    g2 = B.MakeGridGLQ(c, zero)
    assert grid_err(g, g2) <= TOL
    c1 = B.SHExpandGLQ_plx(g2, w, plx)
    c2 = B.SHExpandGLQ(g2, w, zero)
    assert coef_err(c1, c2) <= TOL
    assert coef_err(c1, c) <= 1e-11


CPLX_CASES = [(L, norm, cs, e) for L in (0, 1, 2, 7, 12, 33, 100) for norm in (1, 2, 3, 4) for cs in (1, -1)
              for e in (0, 1) if not (norm == 3 and L > 12)]


@pytest.mark.parametrize("L,norm,cs,extend", CPLX_CASES)
def test_complex_transforms_against_oracle(B, L, norm, cs, extend):
    """MakeGridGLQC / SHExpandGLQC with a table of the complex normalisation
    (cnorm = 1) against the oracle's plx branch on the same table."""
    print(f"case plx complex L={L} norm={norm} cs={cs} e={extend}", flush=True)
    rng = np.random.default_rng(4000 + 10 * L + norm)
    c = rand_cilm_c(rng, L, norm)
    zero, w, plx = B.SHGLQ_plx(L, norm=norm, csphase=cs, cnorm=1)
    t = plx.cpu().numpy()
    g_ref = oracle.MakeGridGLQC_plx(c, t, norm=norm, csphase=cs, extend=extend)
    g = B.MakeGridGLQC_plx(c, plx, norm=norm, csphase=cs, extend=extend)
    assert g.shape == g_ref.shape
    assert grid_err(g, g_ref) <= TOL
    gin = g_ref[:, : 2 * L + 1]
    c_ref = oracle.SHExpandGLQC_plx(gin, w, t, norm=norm, csphase=cs)
    c2 = B.SHExpandGLQC_plx(gin, w, plx, norm=norm, csphase=cs)
    assert c2.shape == c_ref.shape
    assert coef_err(c2, c_ref) <= (TOL if norm != 3 else 1e-10)
    if L >= 2:
        lc = L // 2
        gr = rng.standard_normal(gin.shape) + 1j * rng.standard_normal(gin.shape)
        assert coef_err(B.SHExpandGLQC_plx(gr, w, plx, norm=norm, csphase=cs, lmax_calc=lc),
                        oracle.SHExpandGLQC_plx(gr, w, t, norm=norm, csphase=cs, lmax_calc=lc)) <= TOL


def test_complex_high_degree_against_on_the_fly(B):
    L = 700
    rng = np.random.default_rng(L)
    c = rand_cilm_c(rng, L)
    zero, w, plx = B.SHGLQ_plx(L, norm=1, cnorm=1)
    g = B.MakeGridGLQC_plx(c, plx)
    g2 = B.MakeGridGLQC(c, zero)
    assert grid_err(g, g2) <= TOL
    assert coef_err(B.SHExpandGLQC_plx(g2, w, plx), B.SHExpandGLQC(g2, w, zero)) <= TOL


def test_exit_statuses(B):
    import torch
    from shtools_amd import _lib
    lib = _lib.lib()
    L = 4
    c = torch.zeros((2, L + 1, L + 1), dtype=torch.float64, device="cuda")
    g = torch.zeros((L + 1, 2 * L + 2), dtype=torch.float64, device="cuda")
    _, w, plx = B.SHGLQ_plx(L)
    st = torch.cuda.current_stream().cuda_stream
    P = plx.data_ptr()
    # no table: neither table nor nodes (5); norm (2); extend (2); lmax_calc (2)
    assert lib.shtg_makegrid_glq_plx(c.data_ptr(), L + 1, L + 1, L, None, 1, 1, L, 0, g.data_ptr(), L + 1,
                                     2 * L + 2, st) == 5
    assert lib.shtg_makegrid_glq_plx(c.data_ptr(), L + 1, L + 1, L, P, 0, 1, L, 0, g.data_ptr(), L + 1, 2 * L + 2,
                                     st) == 2
    assert lib.shtg_makegrid_glq_plx(c.data_ptr(), L + 1, L + 1, L, P, 1, 1, L, 3, g.data_ptr(), L + 1, 2 * L + 2,
                                     st) == 2
    assert lib.shtg_makegrid_glq_plx(c.data_ptr(), L + 1, L + 1, L, P, 1, 1, L + 1, 0, g.data_ptr(), L + 1,
                                     2 * L + 2, st) == 2
    wq = np.ascontiguousarray(w)
    assert lib.shtg_shexpand_glq_plx(g.data_ptr(), L + 1, 2 * L + 2, B._dptr(wq), L + 1, None, 1, 1, L,
                                     c.data_ptr(), L + 1, L + 1, st) == 5
    assert lib.shtg_shexpand_glq_plx(g.data_ptr(), L + 1, 2 * L + 2, B._dptr(wq), L, P, 1, 1, L, c.data_ptr(),
                                     L + 1, L + 1, st) == 1
    z0, w0 = np.zeros(L + 1), np.zeros(L + 1)
    assert lib.shtg_shglq_plx(L, B._dptr(z0), B._dptr(w0), P, 5, 1, 0, st) == 2
    assert lib.shtg_shglq_plx(L, B._dptr(z0), B._dptr(w0), P, 1, 2, 0, st) == 2
    assert lib.shtg_shglq_plx(L, B._dptr(z0), B._dptr(w0), None, 1, 1, 0, st) == 1
    with pytest.raises(B.SHToolsError):
        B.MakeGridGLQ_plx(np.zeros((2, L + 1, L + 1)), np.zeros((3, 3)))
