"""Batched DH transforms: every field against the plain path and the oracle."""
import numpy as np
import pytest

import oracle
from helpers import coef_err, grid_err, rand_cilm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.mark.parametrize("norm,cs,sampling", [(1, 1, 2), (4, -1, 2), (2, 1, 1)])
def test_batch_matches_plain_and_oracle(torch, norm, cs, sampling):
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    print(f"case batch norm={norm} cs={cs} s={sampling}", flush=True)
    Ls = [0, 1, 2, 5, 17, 63, 40, 100, 63, 127, 96, 250, 12]   # repeats, Bluestein and smooth sizes
    rng = np.random.default_rng(99)
    cs_host = [rand_cilm(rng, L) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = DHBatch(Ls, sampling=sampling, norm=norm, csphase=cs)
    grids = b.makegrid(cilm)
    back = b.expand(grids)
    torch.cuda.synchronize()
    for i, L in enumerate(Ls):
        g_plain = B.MakeGridDH(cilm[i], sampling=sampling, norm=norm, csphase=cs)
        c_plain = B.SHExpandDH(grids[i], sampling=sampling, norm=norm, csphase=cs)
        torch.cuda.synchronize()
        assert torch.equal(g_plain, grids[i]), L
        assert torch.equal(c_plain, back[i]), L
        if L <= 127:
            g_ref = oracle.MakeGridDH(cs_host[i], sampling=sampling, norm=norm, csphase=cs)
            assert grid_err(grids[i].cpu().numpy(), g_ref) <= 1e-12
            c_ref = oracle.SHExpandDH(g_ref, sampling=sampling, norm=norm, csphase=cs)
            assert coef_err(back[i].cpu().numpy(), c_ref) <= 1e-12
        assert coef_err(back[i].cpu().numpy(), cs_host[i]) <= 1e-11


def test_batch_repeated_calls_and_rebinding(torch):
    from shtools_amd.batch import DHBatch
    print("case batch rebinding", flush=True)
    Ls = [31, 7, 64]
    rng = np.random.default_rng(5)
    b = DHBatch(Ls)
    for rep in range(3):
        cs_host = [rand_cilm(rng, L) for L in Ls]
        cilm = [torch.from_numpy(c).cuda() for c in cs_host]
        back = b.expand(b.makegrid(cilm))
        torch.cuda.synchronize()
        for c, o in zip(cs_host, back):
            assert coef_err(o.cpu().numpy(), c) <= 1e-12


@pytest.mark.parametrize("norm,cs,sampling", [(1, 1, 2), (4, -1, 1)])
def test_repeated_degrees_scales_and_zero_field_match_plain(torch, norm, cs, sampling):
    """Many fields of equal degree (which share the plan's ring geometry and
    FFT plan), fields differing in magnitude by many orders and an all-zero
    field: every field equals the plain path bitwise, forward and adjoint."""
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    print(f"case repeated degrees norm={norm} cs={cs} s={sampling}", flush=True)
    Ls = [33] * 2 + [70] * 3 + [5] * 4 + [129] * 5 + [16] * 8 + [48, 201, 201]
    rng = np.random.default_rng(1234)
    cs_host = [rand_cilm(rng, L) * 10.0 ** rng.integers(-8, 8) for L in Ls]
    cs_host[3][...] = 0.0
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = DHBatch(Ls, sampling=sampling, norm=norm, csphase=cs)
    grids = b.makegrid(cilm)
    back = b.expand(grids)
    gbar = [torch.from_numpy(rng.standard_normal(tuple(g.shape))).cuda() for g in grids]
    adj_s = b.makegrid_adjoint(gbar)
    adj_a = b.expand_adjoint(cilm)
    torch.cuda.synchronize()
    for i, L in enumerate(Ls):
        g_plain = B.MakeGridDH(cilm[i], sampling=sampling, norm=norm, csphase=cs)
        c_plain = B.SHExpandDH(grids[i], sampling=sampling, norm=norm, csphase=cs)
        s_plain = B.MakeGridDH_adjoint(gbar[i], tuple(cilm[i].shape), sampling=sampling, norm=norm, csphase=cs)
        a_plain = B.SHExpandDH_adjoint(cilm[i], tuple(grids[i].shape), sampling=sampling, norm=norm, csphase=cs)
        torch.cuda.synchronize()
        assert torch.equal(g_plain, grids[i]), (i, L)
        assert torch.equal(c_plain, back[i]), (i, L)
        assert torch.equal(s_plain, adj_s[i]), (i, L)
        assert torch.equal(a_plain, adj_a[i]), (i, L)
    assert float(back[3].abs().max()) == 0.0
    for i in (0, 5, 9, 22):
        assert coef_err(back[i].cpu().numpy(), cs_host[i]) <= 1e-11


def test_sampling_per_field_matches_plain(torch):
    """A plan mixing DH1 and DH2 fields (the same degree under both
    samplings too): every field equals the plain path of its own sampling
    (bitwise, except the synthesis adjoint: see below) and the oracle to
    1e-12."""
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    print("case batch sampling per field", flush=True)
    Ls = [40, 40, 17, 96, 5, 96, 63, 0, 130]
    ss = [1, 2, 2, 1, 1, 2, 1, 2, 1]
    rng = np.random.default_rng(77)
    cs_host = [rand_cilm(rng, L) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = DHBatch(Ls, sampling=ss, norm=1, csphase=-1)
    grids = b.makegrid(cilm)
    back = b.expand(grids)
    gbar = [torch.from_numpy(rng.standard_normal(tuple(g.shape))).cuda() for g in grids]
    adj_s = b.makegrid_adjoint(gbar)
    adj_a = b.expand_adjoint(cilm)
    torch.cuda.synchronize()
    for i, (L, s) in enumerate(zip(Ls, ss)):
        assert tuple(grids[i].shape) == (2 * L + 2, s * (2 * L + 2))
        g_plain = B.MakeGridDH(cilm[i], sampling=s, norm=1, csphase=-1)
        c_plain = B.SHExpandDH(grids[i], sampling=s, norm=1, csphase=-1)
        s_plain = B.MakeGridDH_adjoint(gbar[i], tuple(cilm[i].shape), sampling=s, norm=1, csphase=-1)
        a_plain = B.SHExpandDH_adjoint(cilm[i], tuple(grids[i].shape), sampling=s, norm=1, csphase=-1)
        torch.cuda.synchronize()
        assert torch.equal(g_plain, grids[i]), (i, L, s)
        assert torch.equal(c_plain, back[i]), (i, L, s)
        # The synthesis adjoint sums over every slot, the pole's included; at
        # L = 63 (65 slots) the plain launch takes them as 2 ring pairs per
        # lane in one pass and the batched one, whose work items are sized
        # for the 64 slots of the analysis, as 1 per lane in two passes: the
        # same sums in another order.
        assert coef_err(adj_s[i].cpu().numpy(), s_plain.cpu().numpy()) <= 1e-14, (i, L, s)
        assert torch.equal(a_plain, adj_a[i]), (i, L, s)
        g_ref = oracle.MakeGridDH(cs_host[i], sampling=s, norm=1, csphase=-1)
        assert grid_err(grids[i].cpu().numpy(), g_ref) <= 1e-12, (i, L, s)
        assert coef_err(back[i].cpu().numpy(), cs_host[i]) <= 1e-11, (i, L, s)
    with pytest.raises(B.SHToolsError):
        DHBatch(Ls, sampling=ss[:-1])


def test_wide_analysis_waves_against_oracle_and_two_passes(torch):
    """Orders of 513 to 768 analysis slots are summed in one pass of 9 to 12
    ring pairs per lane (DH2, lmax 520, 600, 680, 760: one field per wave
    width).  The coefficients agree with the oracle's analysis of the same
    grids (src/SHExpandDH.f95:463-702) to 1e-12, and with the same plan
    capped at 8 ring pairs per lane (two passes, the sums associated
    differently) and the plain path to 1e-14."""
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    Ls = [520, 600, 680, 760]
    print(f"case wide analysis waves lmax {Ls}", flush=True)
    rng = np.random.default_rng(71)
    cs_host = [rand_cilm(rng, L) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in cs_host]
    b = DHBatch(Ls, sampling=2)
    assert b.tuning()["anal_kmax"] == 12
    grids = [g.clone() for g in b.makegrid(cilm)]
    wide = [c.clone() for c in b.expand(grids)]
    t = b.tuning()
    b.set_tuning(dict(t, anal_kmax=8))
    narrow = b.expand(grids)
    torch.cuda.synchronize()
    for i, L in enumerate(Ls):
        g_host = grids[i].cpu().numpy()
        c_ref = oracle.SHExpandDH(g_host, sampling=2, parallel=True)
        ce = coef_err(wide[i].cpu().numpy(), c_ref)
        print(f"  lmax {L}: coef err {ce:.2e}", flush=True)
        assert ce <= 1e-12, L
        assert coef_err(wide[i].cpu().numpy(), narrow[i].cpu().numpy()) <= 1e-14, L
        c_plain = B.SHExpandDH(grids[i], sampling=2)
        assert coef_err(wide[i].cpu().numpy(), c_plain.cpu().numpy()) <= 1e-14, L
        assert coef_err(wide[i].cpu().numpy(), cs_host[i]) <= 1e-11, L
