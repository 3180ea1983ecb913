"""Parity of the HIP MakeGravGridDH with the oracle's restatement of
src/MakeGravGridDH.f95 on seeded Earth-like potential coefficients.

The GPU path evaluates four radially continued syntheses (rho^l weights in
the recursion) and combines them per grid point; the reference sums the
same terms per ring.  Bars: rad, pot and the magnitude within 1e-12 of the
gravity (potential) scale; theta and phi, which carry the gradient's
(a P_l-1 - l z P_l) / u^2 cancellation near the poles, within 1e-11."""
import numpy as np
import pytest

import oracle
from helpers import rand_cilm

pytestmark = pytest.mark.gpu

GM, A, F, OMEGA = 3.986004418e14, 6378137.0, 1.0 / 298.257223563, 7.292115e-5


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd import backend
    return backend


def earth_like(rng, L, pad=0):
    """C00 = 1, C20 = -4.84e-4 and Kaula-like 1e-5/l^2 RMS above."""
    c = rand_cilm(rng, L + pad)
    l = np.arange(L + pad + 1, dtype=float)
    rms = np.zeros_like(l)
    rms[2:] = 1e-5 / l[2:] ** 2
    c *= (rms * np.sqrt(2 * l + 1))[None, :, None]
    c[0, 0, 0] = 1.0
    if L + pad >= 2:
        c[0, 2, 0] = -4.84e-4
    return c


def check(out, ref, L):
    rad, th, ph, tot, pot = (np.asarray(x.cpu()) if hasattr(x, "cpu") else x for x in out)
    r_rad, r_th, r_ph, r_tot, r_pot = ref
    for x, y in zip((rad, th, ph, tot, pot), ref):
        assert x.shape == y.shape
    gs = np.abs(r_rad).max()
    assert np.abs(rad - r_rad).max() <= 1e-12 * gs
    assert np.abs(pot - r_pot).max() <= 1e-12 * np.abs(r_pot).max()
    assert np.abs(tot - r_tot).max() <= 1e-12 * gs
    tol_h = 1e-11 if L > 0 else 1e-12
    assert np.abs(th - r_th).max() <= tol_h * gs
    assert np.abs(ph - r_ph).max() <= tol_h * gs


CASES = [
    # (L, sampling, extend, lmax_calc, r0/a, f, omega, normal_gravity, pad)
    (0, 2, 0, 0, 1.0, F, OMEGA, 1, 0),
    (1, 1, 1, 1, 1.0, F, 0.0, 1, 0),
    (2, 2, 1, 2, 0.99, F, OMEGA, 0, 0),
    (15, 1, 0, 15, 1.0, 0.0, 0.0, 1, 0),
    (15, 2, 1, 15, 1.02, F, OMEGA, 1, 0),
    (16, 2, 0, 9, 1.0, F, OMEGA, 1, 3),
    (16, 1, 1, 16, 0.98, 0.01, OMEGA, 1, -4),
    (64, 2, 1, 64, 1.0, F, OMEGA, 1, 0),
    (127, 1, 0, 127, 1.001, F, 0.0, 0, 0),
    (255, 2, 0, 255, 1.0, F, OMEGA, 1, 0),
    (1023, 2, 0, 1023, 1.0, F, OMEGA, 1, 0),
]


@pytest.mark.parametrize("L,sampling,extend,lmax_calc,rr,f,omega,ng,pad", CASES)
def test_gravity_vs_oracle(B, L, sampling, extend, lmax_calc, rr, f, omega, ng, pad):
    print(f"case grav L={L} s={sampling} e={extend} lc={lmax_calc} r0/a={rr} f={f} w={omega} ng={ng}",
          flush=True)
    rng = np.random.default_rng(9100 + L)
    c = earth_like(rng, L, pad)
    kw = dict(lmax=L, a=A, f=f, sampling=sampling, lmax_calc=lmax_calc, omega=omega, normal_gravity=ng,
              extend=extend)
    ref = oracle.MakeGravGridDH(c, GM, A * rr, **kw)
    out = B.MakeGravGridDH(c, GM, A * rr, **kw)
    check(out, ref, L)


def test_gravity_defaults_and_tensor_io(B):
    """pyshtools defaults (a = r0, f = 0, sampling 2, omega 0, normal gravity
    removed with gm) on a device tensor; outputs stay on the device."""
    import torch
    rng = np.random.default_rng(9200)
    c = earth_like(rng, 31)
    ref = oracle.MakeGravGridDH(c, GM, A)
    out = B.MakeGravGridDH(torch.from_numpy(c).cuda(), GM, A)
    assert all(o.is_cuda for o in out)
    check(out, ref, 31)
    # f = 0, a = r0, omega = 0: the magnitude minus GM/a^2 is what remains
    assert abs(float(out[3].abs().max())) < 1e-2 * GM / A ** 2


def test_gravity_normal_gravity_gm(B):
    rng = np.random.default_rng(9300)
    c = earth_like(rng, 20)
    kw = dict(a=A, f=F, omega=OMEGA, normal_gravity=1, normal_gravity_gm=GM * 1.001)
    check(B.MakeGravGridDH(c, GM, A, **kw), oracle.MakeGravGridDH(c, GM, A, **kw), 20)


def test_gravity_errors(B):
    c = np.zeros((2, 6, 6))
    c[0, 0, 0] = 1.0
    with pytest.raises(B.SHToolsError):
        B.MakeGravGridDH(c, GM, A, sampling=3)
    with pytest.raises(B.SHToolsError):
        B.MakeGravGridDH(c, GM, A, lmax_calc=9)
    with pytest.raises(B.SHToolsError):
        B.MakeGravGridDH(c, GM, A, normal_gravity=2)
