"""shtg_makegrid_dh_noalloc through the C ABI: the same grids, bitwise, as
shtg_makegrid_dh on two calls in a row (the second smaller, so it runs on
scratch the first one left behind), its errors for a request beyond its
fixed scratch and for a degree whose tables no ordinary call has built,
and no device memory taken from the allocator."""
import ctypes

import numpy as np
import pytest

from helpers import rand_cilm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _call(torch, fn, c, lmax, sampling, norm, csphase, lmax_calc, extend, out=None):
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    if out is None:
        out = torch.full((g0, g1), float("nan"), dtype=torch.float64, device="cuda")
    st = fn(ctypes.c_void_p(c.data_ptr()), c.shape[1], c.shape[2], lmax, norm, sampling, csphase, lmax_calc,
            extend, ctypes.c_void_p(out.data_ptr()), g0, g1,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, out


def test_noalloc_matches_plain_on_two_calls(torch):
    from shtools_amd import _lib
    lib = _lib.lib()
    rng = np.random.default_rng(71)
    # (lmax, sampling, norm, csphase, lmax_calc, extend): the second smaller
    cases = [(900, 2, 1, 1, 900, 0), (255, 1, 4, -1, 200, 1)]
    for lmax, sampling, norm, cs, lc, ext in cases:
        print(f"case noalloc lmax={lmax} s={sampling} norm={norm} cs={cs} lc={lc} e={ext}", flush=True)
        c = torch.from_numpy(rand_cilm(rng, lmax)).cuda()
        st, ref = _call(torch, lib.shtg_makegrid_dh, c, lmax, sampling, norm, cs, lc, ext)
        assert st == 0
        out = torch.full_like(ref, float("nan"))
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        st, out = _call(torch, lib.shtg_makegrid_dh_noalloc, c, lmax, sampling, norm, cs, lc, ext, out=out)
        assert st == 0
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert torch.equal(out, ref)
        # no device memory taken (the synthesis' intermediates alone are
        # tens of MiB at lmax 900)
        assert free0 - free1 < (2 << 20), (free0, free1)
    # back to back on one stream without synchronising: stream order keeps
    # each call's scratch use apart
    c1 = torch.from_numpy(rand_cilm(rng, 600)).cuda()
    c2 = torch.from_numpy(rand_cilm(rng, 130)).cuda()
    _, r1 = _call(torch, lib.shtg_makegrid_dh, c1, 600, 2, 1, 1, 600, 0)
    _, r2 = _call(torch, lib.shtg_makegrid_dh, c2, 130, 2, 1, 1, 130, 0)
    torch.cuda.synchronize()
    s1, o1 = _call(torch, lib.shtg_makegrid_dh_noalloc, c1, 600, 2, 1, 1, 600, 0)
    s2, o2 = _call(torch, lib.shtg_makegrid_dh_noalloc, c2, 130, 2, 1, 1, 130, 0)
    torch.cuda.synchronize()
    assert s1 == 0 and s2 == 0
    assert torch.equal(o1, r1) and torch.equal(o2, r2)


def test_noalloc_errors(torch):
    from shtools_amd import _lib
    from shtools_amd.backend import _status_message
    lib = _lib.lib()
    rng = np.random.default_rng(72)
    # tables never built: degree 421 with sampling 1 is used nowhere else
    c = torch.from_numpy(rand_cilm(rng, 421)).cuda()
    st, _ = _call(torch, lib.shtg_makegrid_dh_noalloc, c, 421, 1, 1, 1, 421, 0)
    assert st == 8 and "ordinary" in _status_message(st)
    # beyond the 64 MiB scratch: lmax 1400 needs about 80 MiB
    L = 1400
    c = torch.from_numpy(rand_cilm(rng, L)).cuda()
    st, ref = _call(torch, lib.shtg_makegrid_dh, c, L, 2, 1, 1, L, 0)
    assert st == 0
    torch.cuda.synchronize()
    st, _ = _call(torch, lib.shtg_makegrid_dh_noalloc, c, L, 2, 1, 1, L, 0)
    assert st == 7 and "scratch" in _status_message(st)
    # argument errors as the ordinary call
    st, _ = _call(torch, lib.shtg_makegrid_dh_noalloc, c, L, 3, 1, 1, L, 0)
    assert st == 2
