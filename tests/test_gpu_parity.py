"""Parity of the HIP transforms with the oracle (the CPU restatement of the
SHTOOLS Fortran routines) on the same seeded inputs.

Bars (floating point, FP64 end to end): grids max|d|/max|ref| <= 1e-12,
coefficients ||d||/||ref|| <= 1e-12, round trips <= 1e-11.  The
unnormalised functions (norm 3) are only stable to degree ~15 in the
reference (src/MakeGridDH.f95:17-18) and are tested there."""
import numpy as np
import pytest

import oracle
from helpers import coef_err, for_norm, grid_err, rand_cilm

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd import backend
    return backend


DH_CASES = [(L, norm, cs, s, e)
            for L in (0, 1, 2, 3, 7, 12, 31)
            for norm in (1, 2, 3, 4)
            for cs in (1, -1)
            for s in (1, 2)
            for e in (0, 1)
            if not (norm == 3 and L > 15)]


@pytest.mark.parametrize("L,norm,cs,sampling,extend", DH_CASES)
def test_dh_small_all_options(B, L, norm, cs, sampling, extend):
    print(f"case DH L={L} norm={norm} cs={cs} s={sampling} e={extend}", flush=True)
    rng = np.random.default_rng(1000 + L)
    c = for_norm(rand_cilm(rng, L), norm)
    g_ref = oracle.MakeGridDH(c, norm=norm, sampling=sampling, csphase=cs, extend=extend)
    g = B.MakeGridDH(c, norm=norm, sampling=sampling, csphase=cs, extend=extend)
    assert g.shape == g_ref.shape
    assert grid_err(g, g_ref) <= TOL
    gin = g_ref[: g_ref.shape[0] - extend, : g_ref.shape[1] - extend]
    c_ref = oracle.SHExpandDH(gin, norm=norm, sampling=sampling, csphase=cs)
    c2 = B.SHExpandDH(gin, norm=norm, sampling=sampling, csphase=cs)
    assert c2.shape == c_ref.shape
    assert coef_err(c2, c_ref) <= (TOL if norm != 3 else 1e-10)


GLQ_CASES = [(L, norm, cs, e) for L in (0, 1, 2, 5, 6, 31, 64) for norm in (1, 2, 3, 4) for cs in (1, -1)
             for e in (0, 1) if not (norm == 3 and L > 15)]


@pytest.mark.parametrize("L,norm,cs,extend", GLQ_CASES)
def test_glq_small_all_options(B, L, norm, cs, extend):
    print(f"case GLQ L={L} norm={norm} cs={cs} e={extend}", flush=True)
    rng = np.random.default_rng(2000 + L)
    c = for_norm(rand_cilm(rng, L), norm)
    z, w = oracle.SHGLQ(L)
    g_ref = oracle.MakeGridGLQ(c, z, norm=norm, csphase=cs, extend=extend)
    g = B.MakeGridGLQ(c, z, norm=norm, csphase=cs, extend=extend)
    assert g.shape == g_ref.shape
    assert grid_err(g, g_ref) <= TOL
    gin = g_ref[:, : g_ref.shape[1] - extend]
    c_ref = oracle.SHExpandGLQ(gin, w, z, norm=norm, csphase=cs)
    c2 = B.SHExpandGLQ(gin, w, z, norm=norm, csphase=cs)
    assert coef_err(c2, c_ref) <= (TOL if norm != 3 else 1e-10)


def test_lmax_calc_padding_and_truncation(B):
    print("case lmax_calc/padding", flush=True)
    rng = np.random.default_rng(7)
    c = rand_cilm(rng, 20)
    for lmax, lc in ((30, 8), (20, 20), (12, 5), (40, 20)):
        g_ref = oracle.MakeGridDH(c, lmax=lmax, lmax_calc=lc, sampling=2)
        g = B.MakeGridDH(c, lmax=lmax, lmax_calc=lc, sampling=2)
        assert g.shape == g_ref.shape and grid_err(g, g_ref) <= TOL
        z, w = oracle.SHGLQ(lmax)
        g_ref = oracle.MakeGridGLQ(c, z, lmax=lmax, lmax_calc=lc)
        g = B.MakeGridGLQ(c, z, lmax=lmax, lmax_calc=lc)
        assert g.shape == g_ref.shape and grid_err(g, g_ref) <= TOL
    g = oracle.MakeGridDH(c, sampling=1)
    for lc in (0, 3, 20):
        assert coef_err(B.SHExpandDH(g, lmax_calc=lc), oracle.SHExpandDH(g, lmax_calc=lc)) <= TOL
    z, w = oracle.SHGLQ(20)
    g = oracle.MakeGridGLQ(c, z)
    for lc in (0, 4, 20):
        assert coef_err(B.SHExpandGLQ(g, w, z, lmax_calc=lc), oracle.SHExpandGLQ(g, w, z, lmax_calc=lc)) <= TOL


def test_torch_tensors_in_and_out(B):
    import torch
    print("case torch tensors", flush=True)
    rng = np.random.default_rng(11)
    c = rand_cilm(rng, 40)
    ct = torch.from_numpy(c).cuda()
    g = B.MakeGridDH(ct, sampling=2)
    assert isinstance(g, torch.Tensor) and g.is_cuda and g.dtype == torch.float64
    c2 = B.SHExpandDH(g, sampling=2)
    assert isinstance(c2, torch.Tensor)
    torch.cuda.synchronize()
    assert coef_err(c2.cpu().numpy(), c) <= 1e-12
    assert grid_err(g.cpu().numpy(), oracle.MakeGridDH(c, sampling=2)) <= TOL
    # host tensors come back as host tensors (computed on the device)
    gh = B.MakeGridDH(torch.from_numpy(c), sampling=2)
    assert isinstance(gh, torch.Tensor) and not gh.is_cuda
    assert torch.equal(gh, g.cpu())
    ch = B.SHExpandDH(gh, sampling=2)
    assert isinstance(ch, torch.Tensor) and not ch.is_cuda and torch.equal(ch, c2.cpu())
    z, w = B.SHGLQ(40)
    gq = B.MakeGridGLQ(torch.from_numpy(c), z)
    assert isinstance(gq, torch.Tensor) and not gq.is_cuda
    assert torch.equal(gq, B.MakeGridGLQ(ct, torch.from_numpy(z).cuda()).cpu())


def test_errors_match_reference(B):
    print("case errors", flush=True)
    c = np.zeros((2, 5, 5))
    with pytest.raises(B.SHToolsError, match="Improper bounds for input variable."):
        B.MakeGridDH(c, sampling=3)
    with pytest.raises(B.SHToolsError, match="Improper bounds"):
        B.MakeGridDH(c, norm=0)
    with pytest.raises(B.SHToolsError, match="Improper bounds"):
        B.SHExpandDH(np.zeros((7, 14)))
    with pytest.raises(B.SHToolsError, match="Improper dimensions of input array."):
        B.SHExpandDH(np.zeros((8, 10)), sampling=2)
    with pytest.raises(NotImplementedError):
        B.SHRotateRealCoef(c, None, None)


@pytest.mark.parametrize("L,sampling", [(127, 2), (255, 1), (300, 2)])
def test_dh_medium_vs_oracle(B, L, sampling):
    print(f"case DH medium L={L} s={sampling}", flush=True)
    rng = np.random.default_rng(L)
    c = rand_cilm(rng, L)
    g_ref = oracle.MakeGridDH(c, sampling=sampling)
    g = B.MakeGridDH(c, sampling=sampling)
    assert grid_err(g, g_ref) <= TOL
    c_ref = oracle.SHExpandDH(g_ref, sampling=sampling)
    assert coef_err(B.SHExpandDH(g_ref, sampling=sampling), c_ref) <= TOL


@pytest.mark.parametrize("L", [63, 64, 65, 128, 129, 511])
@pytest.mark.parametrize("kind", ["dh2", "glq"])
def test_block_edge_degrees_vs_oracle(B, L, kind):
    """Degrees at the edges of the 64-degree blocks in which the plain
    Legendre kernels read their recursion and coefficient rows (an order's
    degree count L - m + 1 runs through every block edge), with one and two
    ring pairs per lane (lmax 511 DH2 synthesis: 2)."""
    print(f"case block edges {kind} L={L}", flush=True)
    rng = np.random.default_rng(600 + L)
    c = rand_cilm(rng, L)
    for norm in (1, 4):
        if kind == "dh2":
            g_ref = oracle.MakeGridDH(c, sampling=2, norm=norm)
            assert grid_err(B.MakeGridDH(c, sampling=2, norm=norm), g_ref) <= TOL
            c_ref = oracle.SHExpandDH(g_ref, sampling=2, norm=norm)
            assert coef_err(B.SHExpandDH(g_ref, sampling=2, norm=norm), c_ref) <= TOL
        else:
            z, w = B.SHGLQ(L)
            g_ref = oracle.MakeGridGLQ(c, z, norm=norm)
            assert grid_err(B.MakeGridGLQ(c, z, norm=norm), g_ref) <= TOL
            c_ref = oracle.SHExpandGLQ(g_ref, w, z, norm=norm)
            assert coef_err(B.SHExpandGLQ(g_ref, w, z, norm=norm), c_ref) <= TOL


def test_dh_l1023_vs_oracle(B):
    """DH2 at the top degree of the batched configuration (lmax = 1023)."""
    print("case DH L=1023", flush=True)
    rng = np.random.default_rng(1023)
    c = rand_cilm(rng, 1023)
    g_ref = oracle.MakeGridDH(c, sampling=2)
    g = B.MakeGridDH(c, sampling=2)
    assert grid_err(g, g_ref) <= TOL
    c_ref = oracle.SHExpandDH(g_ref, sampling=2)
    assert coef_err(B.SHExpandDH(g_ref, sampling=2), c_ref) <= TOL


def test_glq_l1023_norm4_vs_oracle(B):
    """GLQ lmax = 1023, orthonormalised: 1024 x 2047 grid, Bluestein ring FFT."""
    print("case GLQ L=1023 norm 4", flush=True)
    rng = np.random.default_rng(4)
    c = rand_cilm(rng, 1023)
    z, w = B.SHGLQ(1023)
    g_ref = oracle.MakeGridGLQ(c, z, norm=4)
    g = B.MakeGridGLQ(c, z, norm=4)
    assert g.shape == (1024, 2047)
    assert grid_err(g, g_ref) <= TOL
    c_ref = oracle.SHExpandGLQ(g_ref, w, z, norm=4)
    assert coef_err(B.SHExpandGLQ(g_ref, w, z, norm=4), c_ref) <= TOL


def test_dh_l2047_roundtrip(B):
    """DH sampling 1 at lmax = 2047 (4096 x 4096): extended-range scaling near
    the poles; size-independent round-trip property on the GPU alone."""
    import torch
    print("case DH L=2047 round trip", flush=True)
    rng = np.random.default_rng(2047)
    c = rand_cilm(rng, 2047)
    ct = torch.from_numpy(c).cuda()
    g = B.MakeGridDH(ct, norm=4, sampling=1)
    c2 = B.SHExpandDH(g, norm=4, sampling=1)
    assert coef_err(c2.cpu().numpy(), c) <= 1e-11


@pytest.mark.parametrize("L", [0, 1, 2, 5, 12])
@pytest.mark.parametrize("kind", ["dh1", "dh2", "glq"])
def test_analysis_of_arbitrary_grids_vs_brute_force_quadrature(B, L, kind):
    """The HIP analysis of grids with independent random values at every
    node (not band-limited) against the mpmath quadrature fixtures
    (tests/golden/make_brute_analysis.py), every normalisation and phase:
    max error in the orthonormal scale <= 1e-13 of the largest coefficient."""
    import os
    from test_oracle import _orth_scale
    print(f"case brute-force analysis {kind} L={L}", flush=True)
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "brute_analysis.npz"),
                allow_pickle=False)
    g = d[f"{kind}_L{L}_grid"]
    for norm in (1, 2, 3, 4):
        for cs in (1, -1):
            ref = d[f"{kind}_L{L}_n{norm}_cs{cs}"]
            if kind == "glq":
                zero, w = B.SHGLQ(L)
                c = B.SHExpandGLQ(g, w, zero, norm=norm, csphase=cs)
            else:
                c = B.SHExpandDH(g, norm=norm, sampling=1 if kind == "dh1" else 2, csphase=cs)
            s = _orth_scale(L, norm)
            err = np.abs((c - ref) * s).max() / np.abs(ref * s).max()
            assert err <= 1e-13, (norm, cs, err)
