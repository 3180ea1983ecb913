"""The HIP path on the coefficient data the reference ships for its examples
(examples/ExampleDataFiles, as arrays in golden/example_fields.npz): the
Mars gravity model GMM-3 through MakeGravGridDH exactly as
examples/python/GravMag/TestGrav.py:31-37 calls it, Earth topography
(degree 300) and Mars topography (degree 90) through the DH and GLQ
transforms and a batched plan, each against the oracle."""
import os

import numpy as np
import pytest

import oracle
from helpers import coef_err, grid_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

# pyshtools/constants/Mars.py:109-136,92 (a, b, f, angular velocity)
MARS_A, MARS_B = 3395428.0, 3377678.0
MARS_F = (MARS_A - MARS_B) / MARS_A
MARS_OMEGA = 350.891985307 * 2 * np.pi / 360 / (24 * 60 * 60)


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd import backend
    return backend


@pytest.fixture(scope="module")
def fields():
    with np.load(os.path.join(HERE, "golden", "example_fields.npz")) as d:
        return {k: d[k] for k in d.files}


def test_mars_gravity_as_in_reference_example(B, fields):
    print("case GMM-3 MakeGravGridDH lmax 719 lmax_calc 85", flush=True)
    clm = fields["gmm3_cilm"].copy()
    clm[0, 0, 0] = 1.0
    r0 = float(fields["gmm3_r0_km"]) * 1e3
    gm = float(fields["gmm3_gm_km3"]) * 1e9
    kw = dict(lmax=719, lmax_calc=85, a=MARS_A, f=MARS_F, omega=MARS_OMEGA, normal_gravity=1)
    out = B.MakeGravGridDH(clm, gm, r0, **kw)
    ref = oracle.MakeGravGridDH(clm, gm, r0, **kw)
    names = ("rad", "theta", "phi", "total", "pot")
    gs = np.abs(ref[0]).max()
    for name, x, y in zip(names, out, ref):
        assert x.shape == y.shape == (1440, 2880), name
        scale = np.abs(y).max() if name == "pot" else gs
        tol = 1e-11 if name in ("theta", "phi") else 1e-12
        assert np.abs(x - y).max() <= tol * scale, name


@pytest.mark.parametrize("name,sampling", [("srtm300", 2), ("mars90", 1)])
def test_topography_dh_and_glq(B, fields, name, sampling):
    print(f"case {name} DH sampling {sampling} and GLQ", flush=True)
    c = fields[name]
    L = c.shape[1] - 1
    g = B.MakeGridDH(c, sampling=sampling)
    g_ref = oracle.MakeGridDH(c, sampling=sampling)
    assert grid_err(g, g_ref) <= 1e-12
    back = B.SHExpandDH(g, sampling=sampling)
    assert coef_err(back, oracle.SHExpandDH(g_ref, sampling=sampling)) <= 1e-12
    assert coef_err(back, c) <= 1e-11
    zero, w = B.SHGLQ(L)
    gq = B.MakeGridGLQ(c, zero)
    assert grid_err(gq, oracle.MakeGridGLQ(c, zero)) <= 1e-12
    assert coef_err(B.SHExpandGLQ(gq, w, zero), c) <= 1e-11


def test_example_fields_in_one_batch(B, fields):
    import torch
    from shtools_amd.batch import DHBatch
    print("case example fields batched", flush=True)
    names = ("mars90", "gmm3_cilm", "srtm300")
    cs = [fields[n] for n in names]
    Ls = [c.shape[1] - 1 for c in cs]
    ss = [1, 2, 2]
    cilm = [torch.from_numpy(c).cuda() for c in cs]
    b = DHBatch(Ls, sampling=ss)
    grids = b.makegrid(cilm)
    back = b.expand(grids)
    torch.cuda.synchronize()
    for i, (c, s) in enumerate(zip(cs, ss)):
        g_plain = B.MakeGridDH(cilm[i], sampling=s)
        torch.cuda.synchronize()
        assert torch.equal(g_plain, grids[i]), names[i]
        assert coef_err(back[i].cpu().numpy(), c) <= 1e-11, names[i]
