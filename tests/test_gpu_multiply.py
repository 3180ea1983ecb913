"""Parity of the HIP SHMultiply with the oracle's restatement of
src/SHMultiply.f95 (two GLQ syntheses on the product's grid, pointwise
product, GLQ expansion): coefficients within 1e-12 of the largest output
coefficient, for every normalisation and both phase conventions, unequal
degrees, inputs larger than lmax+1, and device-tensor inputs; the same for
the table branch (precomp = 1, src/SHMultiply.f95:281-325) against the
oracle's table branch."""
import numpy as np
import pytest

import oracle
from helpers import rand_cilm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd import backend
    return backend


CASES = [(0, 0, 1, 1), (1, 1, 1, 1), (3, 2, 2, -1), (10, 7, 3, 1), (10, 7, 4, -1), (31, 16, 1, 1),
         (64, 64, 1, -1), (200, 55, 4, 1), (300, 300, 1, 1)]


@pytest.mark.parametrize("l1,l2,norm,csphase", CASES)
def test_shmultiply_vs_oracle(B, l1, l2, norm, csphase):
    print(f"case SHMultiply lmax1={l1} lmax2={l2} norm={norm} csphase={csphase}", flush=True)
    rng = np.random.default_rng(500 + l1 + 7 * l2)
    a, b = rand_cilm(rng, l1), rand_cilm(rng, l2)
    ref = oracle.SHMultiply(a, b, norm=norm, csphase=csphase)
    out = B.SHMultiply(a, b, norm=norm, csphase=csphase)
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def test_shmultiply_leading_blocks_and_tensors(B):
    """lmax below the array sizes reads the leading blocks; tensors in,
    tensor out."""
    import torch
    rng = np.random.default_rng(77)
    a, b = rand_cilm(rng, 20), rand_cilm(rng, 12)
    ref = oracle.SHMultiply(a, b, lmax1=15, lmax2=9)
    out = B.SHMultiply(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), lmax1=15, lmax2=9)
    assert out.is_cuda and tuple(out.shape) == ref.shape
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


# the unnormalised functions overflow beyond degree ~15 (src/MakeGridDH.f95:17-18)
PRECOMP_CASES = [(0, 0, 1, 1), (0, 5, 4, 1), (1, 1, 2, -1), (6, 4, 3, 1), (10, 7, 1, -1), (31, 16, 2, 1),
                 (64, 64, 4, -1), (120, 87, 1, 1)]


@pytest.mark.parametrize("l1,l2,norm,csphase", PRECOMP_CASES)
def test_shmultiply_table_branch_vs_oracle(B, l1, l2, norm, csphase):
    print(f"case SHMultiply precomp=1 lmax1={l1} lmax2={l2} norm={norm} csphase={csphase}", flush=True)
    rng = np.random.default_rng(900 + l1 + 7 * l2)
    a, b = rand_cilm(rng, l1), rand_cilm(rng, l2)
    ref = oracle.SHMultiply(a, b, norm=norm, csphase=csphase, precomp=1)
    out = B.SHMultiply(a, b, norm=norm, csphase=csphase, precomp=1)
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    # and the recursion branch of the library, to the tables' agreement
    rec = B.SHMultiply(a, b, norm=norm, csphase=csphase)
    assert np.abs(out - rec).max() <= 1e-12 * np.abs(rec).max()


def test_shmultiply_table_branch_repeated_and_leading_blocks(B):
    """Two table-branch calls in a row (the second of lower degree, its table
    and grids in memory the first one freed), reading leading blocks."""
    import torch
    rng = np.random.default_rng(78)
    for (n1, l1), (n2, l2) in [((40, 33), (30, 21)), ((12, 9), (9, 8))]:
        a, b = rand_cilm(rng, n1 - 1), rand_cilm(rng, n2 - 1)
        ref = oracle.SHMultiply(a, b, lmax1=l1, lmax2=l2, precomp=1)
        out = B.SHMultiply(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), lmax1=l1, lmax2=l2,
                           precomp=1)
        assert out.is_cuda and tuple(out.shape) == ref.shape
        assert np.abs(out.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_shmultiply_errors(B):
    c = np.zeros((2, 3, 3))
    with pytest.raises(B.SHToolsError):
        B.SHMultiply(c, c, lmax1=5)
    with pytest.raises(B.SHToolsError):
        B.SHMultiply(c, c, norm=5)
    with pytest.raises(B.SHToolsError):
        B.SHMultiply(c, c, csphase=0)
    with pytest.raises(B.SHToolsError):
        B.SHMultiply(c, c, precomp=2)
