"""Same-degree buckets of a batched plan on the matrix cores
(shtg_batch_set_mfma): the synthesis of every field bitwise equal to the
per-field VALU path (the FP64 MFMA accumulates its products as an in-order
chain of fused multiply-adds, the order of the per-field sums) and against
the oracle (src/MakeGridDH.f95:504-643), for buckets of every column-tile
count, partial tiles, tiny degrees, mixed samplings and fields outside any
bucket in the same plan."""
import numpy as np
import pytest

import oracle
from helpers import grid_err, rand_cilm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.mark.parametrize("norm,cs", [(1, 1), (4, -1)])
def test_mfma_buckets_match_valu_and_oracle(torch, norm, cs):
    from shtools_amd.batch import DHBatch
    print(f"case mfma buckets norm={norm} cs={cs}", flush=True)
    # buckets of 16 (2 column tiles), 5, 9 (partial tiles), 33 (a group of 32
    # and a single field), tiny degrees, and singletons
    Ls = [127] * 16 + [40] * 5 + [300] * 9 + [64] * 33 + [0] * 3 + [1] * 4 + [2] * 6 + [700, 5, 250]
    ss = [2] * 16 + [1] * 5 + [2] * 9 + [2] * 33 + [2] * 3 + [1] * 4 + [2] * 6 + [2, 1, 2]
    rng = np.random.default_rng(81)
    host = [rand_cilm(rng, L) for L in Ls]
    cilm = [torch.from_numpy(c).cuda() for c in host]
    b = DHBatch(Ls, sampling=ss, norm=norm, csphase=cs)
    ref = [g.clone() for g in b.makegrid(cilm)]
    b.set_mfma(2)
    out = b.makegrid(cilm)
    torch.cuda.synchronize()
    for i, L in enumerate(Ls):
        assert torch.equal(out[i], ref[i]), (i, L)
    for i in (0, 16, 21, 30, 63, 70, 75, 78):
        g_ref = oracle.MakeGridDH(host[i], sampling=ss[i], norm=norm, csphase=cs)
        assert grid_err(out[i].cpu().numpy(), g_ref) <= 1e-12, (i, Ls[i])
    # the analysis is unchanged by the setting, and switching off restores
    # the VALU grids bitwise
    back = b.expand(out)
    for i in (0, 30, 63):
        assert float(torch.linalg.norm(back[i] - cilm[i]) / torch.linalg.norm(cilm[i])) <= 1e-11
    b.set_mfma(0)
    again = b.makegrid(cilm)
    torch.cuda.synchronize()
    for i in range(len(Ls)):
        assert torch.equal(again[i], ref[i]), i


def test_mfma_sixteen_fits_workload(torch):
    """The 16-fit least-squares workload's plan (16 fields of degree 511,
    DH2): MFMA synthesis against the VALU path and the oracle."""
    from shtools_amd.batch import DHBatch
    print("case mfma 16 x lmax 511", flush=True)
    L = 511
    rng = np.random.default_rng(82)
    host = [rand_cilm(rng, L, -3.0) for _ in range(16)]
    cilm = [torch.from_numpy(c).cuda() for c in host]
    b = DHBatch([L] * 16, sampling=2)
    ref = [g.clone() for g in b.makegrid(cilm)]
    b.set_mfma(8)
    out = b.makegrid(cilm)
    torch.cuda.synchronize()
    for i in range(16):
        assert torch.equal(out[i], ref[i]), i
    for i in (0, 15):
        assert grid_err(out[i].cpu().numpy(), oracle.MakeGridDH(host[i], sampling=2)) <= 1e-12
