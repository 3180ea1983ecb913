"""Adjoint (transpose) kernels and the differentiable transforms.

Checks, all in FP64 on the HIP device:
  * dense transpose: the synthesis adjoint equals S^T with S assembled
    column by column from the oracle's MakeGridDH at lmax 31 (DH sampling 2),
    and the analysis adjoint equals A^T with A assembled from the oracle's
    SHExpandDH (lmax 31, sampling 1), to 1e-12 norm-wise;
  * dot-product identities <S c, g> = <c, S^T g> and <A g, c> = <g, A^T c>
    for every option (norms, csphase, sampling, extend, lmax_calc, padded or
    truncated arrays), DH and GLQ, to 1e-12;
  * torch.autograd.gradcheck / gradgradcheck of the wrapped transforms;
  * the least-squares fit workload (16 fits at lmax 511, DH sampling 2,
    DH-weighted misfit + 1e-6 sum l(l+1) c^2): autograd gradients against
    central finite differences (h = 1e-6) on 64 random coefficients, to
    1e-7 relative with a floor of 1e-4 max|g|;
  * batched adjoints against the per-field adjoints.
"""
import numpy as np
import pytest

import oracle
from helpers import coef_err, for_norm, rand_cilm

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def B(torch):
    from shtools_amd import backend
    return backend


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def _coef_entries(L):
    """(i, l, m) of the coefficients a synthesis reads: C_lm (m <= l) and S_lm (1 <= m <= l)."""
    return [(0, l, m) for l in range(L + 1) for m in range(l + 1)] + \
           [(1, l, m) for l in range(L + 1) for m in range(1, l + 1)]


def test_synthesis_adjoint_dense_transpose(torch, B):
    L, sampling = 31, 2
    print(f"case dense S^T L={L} DH{sampling}", flush=True)
    ents = _coef_entries(L)
    cols = []
    for (i, l, m) in ents:
        e = np.zeros((2, L + 1, L + 1))
        e[i, l, m] = 1.0
        cols.append(oracle.MakeGridDH(e, sampling=sampling).ravel())
    S = np.stack(cols, axis=1)                         # (nlat*nlon, ncoef)
    rng = np.random.default_rng(31)
    G = rng.standard_normal((3, 2 * L + 2, sampling * (2 * L + 2)))
    for g in G:
        ref = S.T @ g.ravel()
        got = B.MakeGridDH_adjoint(torch.from_numpy(g).cuda(), (2, L + 1, L + 1), sampling=sampling).cpu().numpy()
        vec = np.array([got[i, l, m] for (i, l, m) in ents])
        assert coef_err(vec, ref) <= TOL
        # entries the synthesis does not read get no gradient
        mask = np.ones_like(got, dtype=bool)
        for (i, l, m) in ents:
            mask[i, l, m] = False
        assert np.all(got[mask] == 0.0)


def test_analysis_adjoint_dense_transpose(torch, B):
    L, sampling = 31, 1
    n = 2 * L + 2
    print(f"case dense A^T L={L} DH{sampling}", flush=True)
    cols = []
    for k in range(n * n):
        e = np.zeros(n * n)
        e[k] = 1.0
        cols.append(oracle.SHExpandDH(e.reshape(n, n), sampling=sampling).ravel())
    A = np.stack(cols, axis=1)                         # (2 (L+1)^2, n*n)
    rng = np.random.default_rng(32)
    for _ in range(3):
        cb = rng.standard_normal((2, L + 1, L + 1))
        ref = A.T @ cb.ravel()
        got = B.SHExpandDH_adjoint(torch.from_numpy(cb).cuda(), (n, n), sampling=sampling).cpu().numpy()
        assert coef_err(got.ravel(), ref) <= TOL


DH_ADJ = [(L, norm, cs, s, e, lc, dpad)
          for (L, lc, dpad) in ((0, 0, 0), (1, 1, 0), (4, 4, 0), (12, 7, 0), (12, 12, 3), (12, 12, -4), (33, 33, 0))
          for norm in (1, 2, 3, 4) for cs in (1, -1) for s in (1, 2) for e in (0, 1)
          if not (norm == 3 and L > 15)]


@pytest.mark.parametrize("L,norm,cs,sampling,extend,lmax_calc,dpad", DH_ADJ)
def test_dh_dot_product(torch, B, L, norm, cs, sampling, extend, lmax_calc, dpad):
    print(f"case DH adjoint L={L} norm={norm} cs={cs} s={sampling} e={extend} lc={lmax_calc} pad={dpad}", flush=True)
    rng = np.random.default_rng(7 + L + 100 * norm)
    d = L + 1 + dpad
    if d < 1:
        pytest.skip("empty array")
    c = torch.from_numpy(for_norm(rand_cilm(rng, d - 1), norm)).cuda()
    c[:, :, :] += 0.1 * torch.from_numpy(rng.standard_normal(c.shape)).cuda()   # also off-triangle entries
    g = B.MakeGridDH(c, lmax=L, norm=norm, sampling=sampling, csphase=cs, lmax_calc=lmax_calc, extend=extend)
    gb = torch.from_numpy(rng.standard_normal(tuple(g.shape))).cuda()
    cb = B.MakeGridDH_adjoint(gb, tuple(c.shape), lmax=L, norm=norm, sampling=sampling, csphase=cs,
                              lmax_calc=lmax_calc, extend=extend)
    lhs, rhs = _dot(g, gb), _dot(c, cb)
    assert abs(lhs - rhs) <= TOL * max(abs(lhs), float(g.norm() * gb.norm()) * 1e-3)
    # analysis: grid with (possibly) extra columns
    n = 2 * L + 2
    gin = torch.from_numpy(rng.standard_normal((n, sampling * n + max(dpad, 0)))).cuda()
    a = B.SHExpandDH(gin, norm=norm, sampling=sampling, csphase=cs, lmax_calc=lmax_calc)
    ab = torch.from_numpy(rng.standard_normal(tuple(a.shape))).cuda()
    gb2 = B.SHExpandDH_adjoint(ab, tuple(gin.shape), norm=norm, sampling=sampling, csphase=cs, lmax_calc=lmax_calc)
    lhs, rhs = _dot(a, ab), _dot(gin, gb2)
    assert abs(lhs - rhs) <= TOL * max(abs(lhs), float(a.norm() * ab.norm()) * 1e-3)


GLQ_ADJ = [(L, norm, cs, e, lc) for (L, lc) in ((0, 0), (1, 1), (5, 5), (6, 3), (20, 20))
           for norm in (1, 2, 3, 4) for cs in (1, -1) for e in (0, 1) if not (norm == 3 and L > 15)]


@pytest.mark.parametrize("L,norm,cs,extend,lmax_calc", GLQ_ADJ)
def test_glq_dot_product(torch, B, L, norm, cs, extend, lmax_calc):
    print(f"case GLQ adjoint L={L} norm={norm} cs={cs} e={extend} lc={lmax_calc}", flush=True)
    rng = np.random.default_rng(11 + L + 100 * norm)
    zero, w = B.SHGLQ(L)
    c = torch.from_numpy(for_norm(rand_cilm(rng, L), norm)).cuda()
    g = B.MakeGridGLQ(c, zero, norm=norm, csphase=cs, lmax_calc=lmax_calc, extend=extend)
    gb = torch.from_numpy(rng.standard_normal(tuple(g.shape))).cuda()
    cb = B.MakeGridGLQ_adjoint(gb, tuple(c.shape), zero, norm=norm, csphase=cs, lmax_calc=lmax_calc, extend=extend)
    lhs, rhs = _dot(g, gb), _dot(c, cb)
    assert abs(lhs - rhs) <= TOL * max(abs(lhs), float(g.norm() * gb.norm()) * 1e-3)
    gin = torch.from_numpy(rng.standard_normal((L + 1, 2 * L + 1))).cuda()
    a = B.SHExpandGLQ(gin, w, zero, norm=norm, csphase=cs, lmax_calc=lmax_calc)
    ab = torch.from_numpy(rng.standard_normal(tuple(a.shape))).cuda()
    gb2 = B.SHExpandGLQ_adjoint(ab, tuple(gin.shape), w, zero, norm=norm, csphase=cs, lmax_calc=lmax_calc)
    lhs, rhs = _dot(a, ab), _dot(gin, gb2)
    assert abs(lhs - rhs) <= TOL * max(abs(lhs), float(a.norm() * ab.norm()) * 1e-3)


def test_adjoint_errors(torch, B):
    g = torch.zeros((10, 10), dtype=torch.float64, device="cuda")
    with pytest.raises(B.SHToolsError):
        B.MakeGridDH_adjoint(g, (2, 5, 5), lmax=4, sampling=2)      # grid too small for sampling 2
    with pytest.raises(TypeError):
        B.MakeGridDH_adjoint(g.cpu(), (2, 5, 5), lmax=4)


def test_gradcheck(torch):
    from shtools_amd import autograd as AG
    from shtools_amd import backend as B
    rng = np.random.default_rng(5)
    L = 3
    c = torch.from_numpy(rand_cilm(rng, L)).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: AG.make_grid_dh(x, sampling=2, extend=1), (c,))
    assert torch.autograd.gradgradcheck(lambda x: AG.make_grid_dh(x, sampling=2), (c,))
    n = 2 * L + 2
    g = torch.from_numpy(rng.standard_normal((n, 2 * n))).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: AG.sh_expand_dh(x, sampling=2, norm=4), (g,))
    zero, w = B.SHGLQ(L)
    assert torch.autograd.gradcheck(lambda x: AG.make_grid_glq(x, zero, norm=2, csphase=-1), (c,))
    gq = torch.from_numpy(rng.standard_normal((L + 1, 2 * L + 1))).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: AG.sh_expand_glq(x, w, zero), (gq,))


def test_batched_adjoints_match_plain(torch, B):
    from shtools_amd.batch import DHBatch
    rng = np.random.default_rng(9)
    Ls = [5, 31, 17, 64, 31]
    batch = DHBatch(Ls, sampling=2, norm=4, csphase=-1)
    gbs = [torch.from_numpy(rng.standard_normal(batch.grid_shape(i))).cuda() for i in range(len(Ls))]
    cbs = [torch.from_numpy(rng.standard_normal(batch.cilm_shape(i))).cuda() for i in range(len(Ls))]
    got_c = batch.makegrid_adjoint(gbs)
    got_g = batch.expand_adjoint(cbs)
    for i, L in enumerate(Ls):
        ref_c = B.MakeGridDH_adjoint(gbs[i], batch.cilm_shape(i), sampling=2, norm=4, csphase=-1)
        ref_g = B.SHExpandDH_adjoint(cbs[i], batch.grid_shape(i), sampling=2, norm=4, csphase=-1)
        assert coef_err(got_c[i].cpu(), ref_c.cpu()) <= 1e-14
        assert coef_err(got_g[i].cpu(), ref_g.cpu()) <= 1e-14


def test_lsq_fit_gradient_vs_finite_differences(torch, B):
    """16 fits at lmax 511: loss = sum_f sum_i w_i sum_k (S c_f - d_f)^2 + 1e-6 sum l(l+1) c^2.

    Central differences (h = 1e-6) of the loss, evaluated as
    [L(c+he) - L(c-he)] / 2h = sum w (g+ - g-)(g+ + g- - 2d) / 2h + reg, which
    is the same difference with the common terms cancelled point by point
    (only the perturbed fit changes), so that its rounding stays far below
    the 1e-7 bar."""
    from shtools_amd import autograd as AG
    from shtools_amd.batch import DHBatch
    from shtools_amd.synthetic import lsq_fit_workload
    L, nfits = 511, 16
    print(f"case lsq fit L={L} fits={nfits}", flush=True)
    batch = DHBatch([L] * nfits, sampling=2, norm=1, csphase=1)
    truth, data, wlat = lsq_fit_workload(batch, seed=4)
    ll1 = torch.tensor([l * (l + 1.0) for l in range(L + 1)], dtype=torch.float64, device="cuda")[None, :, None]

    c0 = [(0.5 * t).clone().requires_grad_(True) for t in truth]
    grids = AG.batch_make_grid_dh(batch, c0)
    loss = 0.0
    for g, d, c in zip(grids, data, c0):
        loss = loss + (wlat[:, None] * (g - d) ** 2).sum() + 1e-6 * (ll1 * c * c).sum()
    grads = torch.autograd.grad(loss, c0)
    gmax = max(float(g.abs().max()) for g in grads)

    rng = np.random.default_rng(12)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for _ in range(64):
            f = int(rng.integers(nfits))
            l = int(rng.integers(1, L + 1))
            m = int(rng.integers(0, l + 1))
            i = 0 if m == 0 else int(rng.integers(2))
            cp = c0[f].detach().clone()
            cm = c0[f].detach().clone()
            cp[i, l, m] += h
            cm[i, l, m] -= h
            gp = B.MakeGridDH(cp, sampling=2)
            gm = B.MakeGridDH(cm, sampling=2)
            dmis = float((wlat[:, None] * (gp - gm) * (gp + gm - 2 * data[f])).sum())
            dreg = 1e-6 * l * (l + 1.0) * (float(cp[i, l, m]) ** 2 - float(cm[i, l, m]) ** 2)
            fd = (dmis + dreg) / (2 * h)
            ad = float(grads[f][i, l, m])
            rel = abs(fd - ad) / max(abs(ad), 1e-4 * gmax)
            worst = max(worst, rel)
    print(f"lsq fit: worst FD relative error {worst:.2e} (max |g| {gmax:.3e})", flush=True)
    assert worst <= 1e-7
