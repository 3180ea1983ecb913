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


def test_known_answers_normalisation_notebook(B):
    """examples/notebooks/spherical-harmonic-normalizations.ipynb cells 4, 9,
    11: a single 4pi-normalised Y_{5,2} at lmax 100 has N = 4 pi under both
    quadratures (printed 12.566370614359176 for GLQ and 12.566370614359178
    for DH); here on the HIP grids, with the library's nodes and weights."""
    print("case notebook known answers", flush=True)
    lmax = 100
    c = np.zeros((2, lmax + 1, lmax + 1))
    c[0, 5, 2] = 1.0
    z, w = B.SHGLQ(lmax)
    g = B.MakeGridGLQ(c, z, extend=1)
    nlon = g.shape[1] - 1
    N = ((g[:, :nlon] ** 2) * w[:, None]).sum() * (2.0 * np.pi / nlon)
    assert abs(N - 12.566370614359176) < 4e-15 * 12.6
    g = B.MakeGridDH(c, extend=1)
    n = g.shape[0] - 1
    aj = B.DHaj(n)
    N = ((g[:n, :n] ** 2) * aj[:, None]).sum() * 2.0 * np.sqrt(2.0) * np.pi / n
    assert abs(N - 12.566370614359178) < 4e-15 * 12.6
    # and the analysis returns the single coefficient
    back = B.SHExpandDH(g[:n, :n])
    assert abs(back[0, 5, 2] - 1.0) < 1e-14 and np.abs(back).sum() - abs(back[0, 5, 2]) < 1e-12


def _accuracy_program_coefficients(maxdeg, complex_):
    """The coefficients of the reference's accuracy programs
    (examples/python/TimingAccuracy/*.py): normal random values (seed 0; real
    and imaginary parts drawn in turn for the complex programs) scaled to the
    power 1/(1+l)^1.5 per degree, zero outside m <= l and for S_l0 (the
    complex programs' negative m = 0)."""
    ls = np.arange(maxdeg + 1)
    mask = np.zeros((2, maxdeg + 1, maxdeg + 1), dtype=bool)
    for l in ls:
        mask[:, l, : l + 1] = True
    mask[1, :, 0] = False
    rs = np.random.RandomState(0)
    shape = (2, maxdeg + 1, maxdeg + 1)
    if complex_:
        cilm = np.zeros(shape, dtype=np.complex128)
        cilm.imag = rs.normal(loc=0.0, scale=1.0, size=shape)
        cilm.real = rs.normal(loc=0.0, scale=1.0, size=shape)
    else:
        cilm = rs.normal(loc=0.0, scale=1.0, size=shape)
    # power per degree (4pi normalisation): sum over m <= l of |C|^2 + |S|^2
    old_power = np.array([(np.abs(cilm[:, l, : l + 1]) ** 2).sum() for l in ls])
    cilm *= np.sqrt((1.0 / (1.0 + ls) ** 1.5) / old_power)[None, :, None]
    cilm[~mask] = 0.0
    return cilm, mask


@pytest.mark.parametrize("program,maxdeg", [("DH1", 2048), ("DH2", 2048), ("GLQ", 2048), ("DHC2", 1024),
                                            ("GLQC", 1024)])
def test_reference_accuracy_programs(B, program, maxdeg):
    """examples/python/TimingAccuracy/TimingAccuracy{DH,GLQ,DHC,GLQC}.py on
    the HIP path: round trips at lmax 2, 4, ..., maxdeg and the relative
    error of every coefficient.  The programs fail above a maximum relative
    error of 100; here the median grows like rounding over the degrees
    (measured 6e-16 at lmax 2 to 1.1e-13 at 2048 for DH) and the
    coefficient vector's relative error stays within 1e-11."""
    cilm, mask = _accuracy_program_coefficients(maxdeg, program.endswith("C") or program == "DHC2")
    lmax = 2
    while lmax <= maxdeg:
        print(f"case accuracy program {program} lmax {lmax}", flush=True)
        c = np.ascontiguousarray(cilm[:, : lmax + 1, : lmax + 1])
        m = mask[:, : lmax + 1, : lmax + 1]
        if program in ("DH1", "DH2"):
            s = int(program[2])
            c2 = B.SHExpandDH(B.MakeGridDH(c, sampling=s), sampling=s)
        elif program == "DHC2":
            c2 = B.SHExpandDHC(B.MakeGridDHC(c, sampling=2), sampling=2)
        else:
            zero, w = B.SHGLQ(lmax)
            if program == "GLQ":
                c2 = B.SHExpandGLQ(B.MakeGridGLQ(c, zero), w, zero)
            else:
                c2 = B.SHExpandGLQC(B.MakeGridGLQC(c, zero), w, zero)
        err = np.abs(c[m] - c2[m]) / np.abs(c[m])
        print(f"  maxerror {err.max():.2e} mean square {np.mean(err ** 2):.2e} median {np.median(err):.2e}",
              flush=True)
        assert err.max() <= 100.0
        # (GLQ at lmax 2: 2.7e-15 here, 2.6e-15 for the oracle's round trip)
        assert np.median(err) <= max(3e-16 * lmax, 5e-15)
        assert coef_err(c2, c) <= 1e-11
        lmax *= 2


@pytest.mark.parametrize("name,radius", [("mars90", 3389500.0), ("srtm300", 6371000.0)])
def test_topography_gradient(B, fields, name, radius):
    """MakeGradientDH of the example topographies (DH2 with the extra row
    and column) against the oracle, on the gradient's bar (max|d|/max|ref|
    <= 1e-11: both evaluate a difference that cancels near the poles)."""
    print(f"case gradient {name}", flush=True)
    c = fields[name]
    kw = dict(sampling=2, extend=1, radius=radius)
    th, ph = B.MakeGradientDH(c, **kw)
    th_ref, ph_ref = oracle.MakeGradientDH(c, **kw)
    assert th.shape == th_ref.shape and ph.shape == ph_ref.shape
    assert grid_err(th, th_ref) <= 1e-11 and grid_err(ph, ph_ref) <= 1e-11


def test_product_of_example_fields(B, fields):
    """SHMultiply of Mars topography (degree 90) with the GMM-3 gravity
    coefficients (degree 120) against the oracle (1e-12 of the largest
    product coefficient), and the product's DH synthesis against the
    pointwise product of the factors' syntheses."""
    print("case SHMultiply example fields", flush=True)
    a, b = fields["mars90"], fields["gmm3_cilm"]
    out = B.SHMultiply(a, b)
    ref = oracle.SHMultiply(a, b)
    assert out.shape == ref.shape == (2, 211, 211)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    L = 210
    pad = lambda c: np.pad(c, ((0, 0), (0, L + 1 - c.shape[1]), (0, L + 1 - c.shape[2])))
    ga, gb = B.MakeGridDH(pad(a), sampling=2), B.MakeGridDH(pad(b), sampling=2)
    assert grid_err(B.MakeGridDH(out, sampling=2), ga * gb) <= 1e-12
