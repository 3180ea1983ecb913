"""High degrees: ring FFTs whose work array exceeds the 160 KiB LDS of a CU
run on the global-workspace path.  Parity with the oracle (grids <= 1e-12
max-relative), round trips <= 1e-11, the batched high-degree workload
(8 fields, lmax ~ U{2047..2799}, DH sampling 2, Kaula spectrum, upward
continuation filter) and the gradient of its filtered misfit."""
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


@pytest.fixture(scope="module")
def B(torch):
    from shtools_amd import backend
    return backend


@pytest.mark.parametrize("L,sampling,norm", [(1708, 2, 1), (2799, 2, 4)])
def test_dh_large_ring_parity(torch, B, L, sampling, norm):
    """L = 1708: 4(L+1) = 4*1709 (prime) takes a 10368-point Bluestein work
    array; L = 2799: an 11200-point smooth ring; both exceed the LDS."""
    print(f"case DH L={L} sampling={sampling} norm={norm}", flush=True)
    rng = np.random.default_rng(L)
    c = rand_cilm(rng, L)
    g_ref = oracle.MakeGridDH(c, sampling=sampling, norm=norm)
    ct = torch.from_numpy(c).cuda()
    g = B.MakeGridDH(ct, sampling=sampling, norm=norm)
    assert grid_err(g.cpu().numpy(), g_ref) <= 1e-12
    c2 = B.SHExpandDH(g, sampling=sampling, norm=norm)
    assert coef_err(c2.cpu().numpy(), c) <= 1e-11


def test_glq_large_ring_parity(torch, B):
    L = 2600   # 2L+1 = 5201 = 7 * 743: a 10800-point Bluestein work array
    print(f"case GLQ L={L}", flush=True)
    rng = np.random.default_rng(L)
    c = rand_cilm(rng, L)
    zero, w = B.SHGLQ(L)
    g_ref = oracle.MakeGridGLQ(c, zero, norm=4)
    g = B.MakeGridGLQ(torch.from_numpy(c).cuda(), zero, norm=4)
    assert grid_err(g.cpu().numpy(), g_ref) <= 1e-12
    c2 = B.SHExpandGLQ(g, w, zero, norm=4)
    assert coef_err(c2.cpu().numpy(), c) <= 1e-11


def test_high_degree_batch(torch, B):
    """8 fields, lmax ~ U{2047..2799}, filtered by the upward continuation:
    batched synthesis bitwise the plain one, and per field against the
    oracle: grids <= 1e-12 (max-relative), the analysis of the batched grid
    against the oracle's analysis of its own grid <= 1e-12, round trips
    <= 1e-11."""
    from shtools_amd.batch import DHBatch
    from shtools_amd.synthetic import high_degree_workload, upward_continuation
    Ls, cs = high_degree_workload()
    print(f"case high-degree batch lmax {Ls}", flush=True)
    batch = DHBatch(Ls, sampling=2, norm=1, csphase=1)
    filt = [torch.from_numpy(upward_continuation(L)).cuda()[None, :, None] for L in Ls]
    ct = [torch.from_numpy(c).cuda() * f for c, f in zip(cs, filt)]
    grids = batch.makegrid(ct)
    back = batch.expand(grids)
    for i, L in enumerate(Ls):
        c_in = ct[i].cpu().numpy()
        assert coef_err(back[i].cpu().numpy(), c_in) <= 1e-11
        g_plain = B.MakeGridDH(ct[i], sampling=2)
        assert torch.equal(g_plain, grids[i]), (i, L)
        del g_plain
        g_ref = oracle.MakeGridDH(c_in, sampling=2)
        ge = grid_err(grids[i].cpu().numpy(), g_ref)
        c_ref = oracle.SHExpandDH(g_ref, sampling=2, parallel=True)
        ce = coef_err(back[i].cpu().numpy(), c_ref)
        print(f"  field {i} lmax {L}: grid err {ge:.2e}, coefficient err {ce:.2e} vs oracle", flush=True)
        assert ge <= 1e-12, (i, L)
        assert ce <= 1e-12, (i, L)


def test_high_degree_filtered_gradient(torch, B):
    """Gradient of ||S(f c) - d||^2 through the filter: autograd (adjoint
    kernels) against central differences (h = 1e-6) of the misfit, evaluated
    point by point as sum (g+ - g-)(g+ + g- - 2d) / 2h."""
    from shtools_amd import autograd as AG
    from shtools_amd.synthetic import high_degree_workload, upward_continuation
    Ls, cs = high_degree_workload(nfields=1, lmin=2047, lmax=2799, seed=6)
    L = Ls[0]
    print(f"case filtered gradient L={L}", flush=True)
    f = torch.from_numpy(upward_continuation(L)).cuda()[None, :, None]
    c = torch.from_numpy(cs[0]).cuda()
    rng = np.random.default_rng(3)
    d = B.MakeGridDH(c * f, sampling=2)
    d = d + 1e-3 * float(d.abs().max()) * torch.from_numpy(rng.standard_normal(tuple(d.shape))).cuda()
    c0 = (0.7 * c).requires_grad_(True)
    loss = ((AG.make_grid_dh(c0 * f, sampling=2) - d) ** 2).sum()
    (grad,) = torch.autograd.grad(loss, c0)
    gmax = float(grad.abs().max())
    h = 1e-6 * float(c.abs().max())
    worst = 0.0
    with torch.no_grad():
        for _ in range(8):
            l = int(rng.integers(2, 200))
            m = int(rng.integers(0, l + 1))
            i = 0 if m == 0 else int(rng.integers(2))
            cp = c0.detach().clone()
            cm = c0.detach().clone()
            cp[i, l, m] += h
            cm[i, l, m] -= h
            gp = B.MakeGridDH(cp * f, sampling=2)
            gm = B.MakeGridDH(cm * f, sampling=2)
            fd = float(((gp - gm) * (gp + gm - 2 * d)).sum()) / (2 * h)
            ad = float(grad[i, l, m])
            worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-4 * gmax))
    print(f"filtered gradient: worst FD relative error {worst:.2e}", flush=True)
    assert worst <= 1e-7
