"""Batched Gauss-Legendre transforms (GLQBatch, and plans mixing GLQ with DH
fields): every field against the plain MakeGridGLQ / SHExpandGLQ path on the
nodes and weights of SHGLQ, which the plan computes itself, and against the
oracle (src/MakeGridGLQ.f95:441-553, src/SHExpandGLQ.f95:445-649)."""
import numpy as np
import pytest

import oracle
from helpers import coef_err, for_norm, grid_err, rand_cilm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _close_or_equal(torch, a, b, tol=1e-14):
    """Bitwise, or (where the plain launch splits an order's ring pairs over
    more passes than the batched one, so the same sums are associated
    differently) equal to tol in relative norm."""
    if torch.equal(a, b):
        return True
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) <= tol


@pytest.mark.parametrize("norm,cs", [(1, 1), (4, -1), (2, 1), (3, -1)])
def test_glq_batch_matches_plain_and_oracle(torch, norm, cs):
    from shtools_amd import backend as B
    from shtools_amd.batch import GLQBatch
    print(f"case glq batch norm={norm} cs={cs}", flush=True)
    # odd and even nlat, repeats, smooth and Bluestein ring lengths 2L+1
    Ls = [0, 1, 2, 5, 16, 17, 63, 40, 100, 63, 127, 96, 250, 12, 300]
    tol = 1e-12
    if norm == 3:
        # the unnormalised recursion is stable only to degree ~15
        # (src/MakeGridDH.f95:17-18); coefficients re-expressed so that the
        # grids keep an O(1) range
        Ls = [L for L in Ls if L <= 15]
        tol = 1e-10
    rng = np.random.default_rng(404 + norm)
    cs_host = [for_norm(rand_cilm(rng, L), norm) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = GLQBatch(Ls, norm=norm, csphase=cs)
    grids = b.makegrid(cilm)
    back = b.expand(grids)
    torch.cuda.synchronize()
    for i, L in enumerate(Ls):
        assert tuple(grids[i].shape) == (L + 1, 2 * L + 1)
        zero, w = B.SHGLQ(L)
        g_plain = B.MakeGridGLQ(cilm[i], zero, norm=norm, csphase=cs)
        c_plain = B.SHExpandGLQ(grids[i], w, zero, norm=norm, csphase=cs)
        torch.cuda.synchronize()
        assert torch.equal(g_plain, grids[i]), (i, L)
        assert _close_or_equal(torch, back[i], c_plain), (i, L)
        if L <= 127:
            zh, wh = oracle.SHGLQ(L)
            g_ref = oracle.MakeGridGLQ(cs_host[i], zh, norm=norm, csphase=cs)
            assert grid_err(grids[i].cpu().numpy(), g_ref) <= tol, (i, L)
            c_ref = oracle.SHExpandGLQ(g_ref, wh, zh, norm=norm, csphase=cs)
            assert coef_err(back[i].cpu().numpy(), c_ref) <= tol, (i, L)
        assert coef_err(back[i].cpu().numpy(), cs_host[i]) <= 10 * tol, (i, L)


def test_mixed_grids_plan(torch):
    """One plan holding DH1, DH2 and GLQ fields (the same degree on several
    grids too): every field equals its single-grid plan and the plain path,
    forward and adjoint."""
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch, GLQBatch, SHBatch
    print("case mixed grid plan", flush=True)
    Ls = [40, 40, 40, 17, 96, 5, 96, 63, 0, 130, 211]
    gs = ["GLQ", "DH2", "DH1", "GLQ", "GLQ", "DH1", "DH2", "GLQ", "GLQ", "DH2", "GLQ"]
    rng = np.random.default_rng(78)
    cs_host = [rand_cilm(rng, L) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = SHBatch(Ls, grid=gs, norm=1, csphase=-1)
    grids = b.makegrid(cilm)
    back = b.expand(grids)
    gbar = [torch.from_numpy(rng.standard_normal(tuple(g.shape))).cuda() for g in grids]
    adj_s = b.makegrid_adjoint(gbar)
    adj_a = b.expand_adjoint(cilm)
    torch.cuda.synchronize()
    for i, (L, g) in enumerate(zip(Ls, gs)):
        if g == "GLQ":
            single = GLQBatch([L], norm=1, csphase=-1)
            zero, w = B.SHGLQ(L)
            g_plain = B.MakeGridGLQ(cilm[i], zero, norm=1, csphase=-1)
            c_plain = B.SHExpandGLQ(grids[i], w, zero, norm=1, csphase=-1)
            s_plain = B.MakeGridGLQ_adjoint(gbar[i], tuple(cilm[i].shape), zero, norm=1, csphase=-1)
            a_plain = B.SHExpandGLQ_adjoint(cilm[i], tuple(grids[i].shape), w, zero, norm=1, csphase=-1)
            assert tuple(grids[i].shape) == (L + 1, 2 * L + 1)
        else:
            s = int(g[-1])
            single = DHBatch([L], sampling=s, norm=1, csphase=-1)
            g_plain = B.MakeGridDH(cilm[i], sampling=s, norm=1, csphase=-1)
            c_plain = B.SHExpandDH(grids[i], sampling=s, norm=1, csphase=-1)
            s_plain = B.MakeGridDH_adjoint(gbar[i], tuple(cilm[i].shape), sampling=s, norm=1, csphase=-1)
            a_plain = B.SHExpandDH_adjoint(cilm[i], tuple(grids[i].shape), sampling=s, norm=1, csphase=-1)
            assert tuple(grids[i].shape) == (2 * L + 2, s * (2 * L + 2))
        g1 = single.makegrid([cilm[i]])[0]
        c1 = single.expand([grids[i]])[0]
        torch.cuda.synchronize()
        assert torch.equal(g1, grids[i]), (i, L, g)
        assert torch.equal(c1, back[i]), (i, L, g)
        assert torch.equal(g_plain, grids[i]), (i, L, g)
        assert _close_or_equal(torch, back[i], c_plain), (i, L, g)
        # adjoints: the synthesis adjoint sums every slot, in passes sized by
        # the plan (see test_gpu_batch.py), so it agrees to rounding
        assert coef_err(adj_s[i].cpu().numpy(), s_plain.cpu().numpy()) <= 1e-14, (i, L, g)
        assert torch.equal(a_plain, adj_a[i]), (i, L, g)
        assert coef_err(back[i].cpu().numpy(), cs_host[i]) <= 1e-11, (i, L, g)
    with pytest.raises(B.SHToolsError):
        SHBatch(Ls, grid=gs[:-1])
    with pytest.raises(B.SHToolsError):
        SHBatch([4], grid=["GLQ2"])


def test_glq_batch_dot_products(torch):
    """<A c, g> = <c, A^T g> for the batched GLQ synthesis and analysis."""
    from shtools_amd.batch import GLQBatch
    print("case glq batch dot products", flush=True)
    Ls = [3, 30, 31, 64, 129]
    rng = np.random.default_rng(9)
    b = GLQBatch(Ls, norm=4, csphase=1)
    c = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    g = [torch.from_numpy(rng.standard_normal(b.grid_shape(i))).cuda() for i in range(len(Ls))]
    Ac = b.makegrid(c)
    Atg = b.makegrid_adjoint(g)
    Bg = b.expand(g)
    Btc = b.expand_adjoint(c)
    torch.cuda.synchronize()
    for i in range(len(Ls)):
        lhs = float(torch.sum(Ac[i] * g[i]))
        rhs = float(torch.sum(c[i] * Atg[i]))
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs)), (i, lhs, rhs)
        lhs = float(torch.sum(Bg[i] * c[i]))
        rhs = float(torch.sum(g[i] * Btc[i]))
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs)), (i, lhs, rhs)


def test_glq_batch_graph_tune_and_large_rings(torch):
    """A GLQ plan with rings beyond the LDS (lmax 2600: the workspace ring
    FFT), tuned and captured: every replay equals the eager result, and the
    grids equal the plain path's."""
    from shtools_amd import backend as B
    from shtools_amd.batch import GLQBatch
    print("case glq batch graph + large", flush=True)
    Ls = [2600, 1023, 700, 64, 1023, 301]
    rng = np.random.default_rng(10)
    cs_host = [rand_cilm(rng, L) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = GLQBatch(Ls)
    ref_g = [x.clone() for x in b.makegrid(cilm)]
    ref_c = [x.clone() for x in b.expand(ref_g)]
    grids = [torch.empty_like(x) for x in ref_g]
    outs = [torch.empty_like(x) for x in ref_c]
    b.tune(cilm, grids, outs, reps=1)
    graph = b.capture(cilm, grids, outs)
    for rep in range(3):
        for x in grids + outs:
            x.fill_(np.nan)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i in range(len(Ls)):
            assert torch.equal(grids[i], ref_g[i]), (rep, i)
            assert torch.equal(outs[i], ref_c[i]), (rep, i)
    for i in (0, 1):
        zero, _ = B.SHGLQ(Ls[i])
        g_plain = B.MakeGridGLQ(cilm[i], zero)
        torch.cuda.synchronize()
        assert torch.equal(g_plain, ref_g[i]), i
        assert coef_err(ref_c[i].cpu().numpy(), cs_host[i]) <= 1e-11, i
    assert max(b.field_groups()) >= 1
