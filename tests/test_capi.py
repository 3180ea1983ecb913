"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, validates arguments with the reference's exit codes before
touching the GPU, and its host helpers match the oracle.  No GPU compute."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from shtools_amd import _lib
from shtools_amd import backend as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "shtools_amd.h")).read()
    return sorted(set(re.findall(r"\b(shtg_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert "shtg_makegrid_dh" in names and "shtg_shexpand_glq" in names
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, name
    assert L.shtg_version().startswith(b"shtools_amd")


def test_shglq_matches_oracle():
    for L in (0, 1, 2, 5, 63, 200):
        z, w = B.SHGLQ(L)
        z2, w2 = oracle.SHGLQ(L)
        np.testing.assert_array_equal(z, z2)
        np.testing.assert_array_equal(w, w2)


def test_dhaj_and_coords_match_oracle():
    for n in (2, 8, 64, 130):
        for ext in (0, 1):
            np.testing.assert_allclose(B.DHaj(n, ext), oracle.DHaj(n, ext), rtol=0, atol=1e-15)
    for ext in (0, 1):
        a, b = B.GLQGridCoord(17, ext)
        a2, b2 = oracle.GLQGridCoord(17, ext)
        np.testing.assert_allclose(a, a2, atol=1e-13)
        np.testing.assert_allclose(b, b2, atol=1e-13)
    with pytest.raises(B.SHToolsError, match="Improper bounds"):
        B.DHaj(5)


@pytest.mark.parametrize("norm", [1, 2, 3, 4])
@pytest.mark.parametrize("cs", [1, -1])
def test_plm_tables_match_closed_form(golden, norm, cs):
    # the GPU recursion tables (rescaled three-term recursion, seeds and
    # extended-range powers) evaluated on the host against mpmath
    L = int(golden["plm_lmax"])
    ref = golden[f"plm_n{norm}_cs{cs}"]
    lib = _lib.lib()
    for iz, z in enumerate(golden["plm_z"]):
        p = np.zeros((L + 1, L + 1))
        assert lib.shtg_plm(L, norm, cs, float(z), B._dptr(p)) == 0
        for l in range(L + 1):
            scale = max(np.abs(ref[iz, l, : l + 1]).max(), 1e-300)
            assert np.abs(p[l, : l + 1] - ref[iz, l, : l + 1]).max() < 5e-14 * scale, (norm, cs, z, l)


def test_plm_tables_high_degree(golden):
    lib = _lib.lib()
    for l, m, z, norm, val in golden["plmhi"]:
        l, m, norm = int(l), int(m), int(norm)
        p = np.zeros((l + 1, l + 1))
        assert lib.shtg_plm(l, norm, 1, float(z), B._dptr(p)) == 0
        assert abs(p[l, m] - val) <= 2e-12 * abs(val) + 1e-300, (l, m, z, norm)


def test_packed_plm_wrappers():
    p = B.PlmBar(6, 0.3)
    full = oracle.plm(6, 0.3, norm=1)
    il, im = np.tril_indices(7)
    np.testing.assert_allclose(p, full[il, im], rtol=1e-14, atol=1e-15)
    assert B.PlmON(4, 0.1).size == 15


def test_argument_checks_before_any_device_work():
    """Invalid arguments return the reference's codes without a GPU."""
    lib = _lib.lib()
    nul = ctypes.c_void_p(0)
    dummy = ctypes.c_void_p(16)
    # MakeGridDH: sampling, extend, dims, norm, csphase, lmax_calc
    assert lib.shtg_makegrid_dh(dummy, 5, 5, 4, 1, 3, 1, 4, 0, dummy, 10, 20, nul) == 2
    assert lib.shtg_makegrid_dh(dummy, 5, 5, 4, 1, 2, 1, 4, 2, dummy, 10, 20, nul) == 2
    assert lib.shtg_makegrid_dh(dummy, 5, 5, 4, 1, 2, 1, 4, 0, dummy, 9, 20, nul) == 1
    assert lib.shtg_makegrid_dh(dummy, 5, 5, 4, 5, 2, 1, 4, 0, dummy, 10, 20, nul) == 2
    assert lib.shtg_makegrid_dh(dummy, 5, 5, 4, 1, 2, 0, 4, 0, dummy, 10, 20, nul) == 2
    assert lib.shtg_makegrid_dh(dummy, 5, 5, 4, 1, 2, 1, 5, 0, dummy, 10, 20, nul) == 2
    # SHExpandDH: lmax_calc, sampling, odd n, cilm dims, grid dims
    assert lib.shtg_shexpand_dh(dummy, 10, 20, 1, 2, 1, 5, dummy, 6, 6, nul) == 2
    assert lib.shtg_shexpand_dh(dummy, 10, 20, 1, 3, 1, 4, dummy, 5, 5, nul) == 2
    assert lib.shtg_shexpand_dh(dummy, 9, 18, 1, 2, 1, 3, dummy, 5, 5, nul) == 2
    assert lib.shtg_shexpand_dh(dummy, 10, 20, 1, 2, 1, 4, dummy, 4, 5, nul) == 1
    assert lib.shtg_shexpand_dh(dummy, 10, 19, 1, 2, 1, 4, dummy, 5, 5, nul) == 1
    # GLQ: missing zero -> 5, short zero -> 1, short w -> 1
    dp = ctypes.POINTER(ctypes.c_double)
    z = np.zeros(5)
    assert lib.shtg_makegrid_glq(dummy, 5, 5, 4, dp(), 5, 1, 1, 4, 0, dummy, 5, 9, nul) == 5
    assert lib.shtg_makegrid_glq(dummy, 5, 5, 4, B._dptr(z), 4, 1, 1, 4, 0, dummy, 5, 9, nul) == 1
    assert lib.shtg_shexpand_glq(dummy, 5, 9, B._dptr(z), 4, B._dptr(z), 5, 1, 1, 4, dummy, 5, 5, nul) == 1
    # batched plans: sampling (uniform or per field), norm, csphase, degree, field count
    ci = ctypes.c_int
    Ls = (ci * 3)(4, 7, 4)
    plan = ctypes.c_void_p(16)
    assert lib.shtg_batch_create_dh(3, Ls, 3, 1, 1, ctypes.byref(plan)) == 2 and not plan.value
    assert lib.shtg_batch_create_dh(0, Ls, 2, 1, 1, ctypes.byref(plan)) == 1 and not plan.value
    assert lib.shtg_batch_create_dh_fields(3, Ls, (ci * 3)(1, 2, 0), 1, 1, ctypes.byref(plan)) == 2
    assert lib.shtg_batch_create_dh_fields(3, Ls, (ci * 3)(1, 2, 1), 5, 1, ctypes.byref(plan)) == 2
    assert lib.shtg_batch_create_dh_fields(3, Ls, (ci * 3)(1, 2, 1), 1, 0, ctypes.byref(plan)) == 2
    assert lib.shtg_batch_create_dh_fields(3, (ci * 3)(4, -1, 4), (ci * 3)(1, 2, 1), 1, 1, ctypes.byref(plan)) == 2
    assert lib.shtg_batch_create_dh_fields(3, Ls, None, 1, 1, ctypes.byref(plan)) == 1
    assert not plan.value


def test_normal_gravity_matches_oracle():
    """Host helper (no GPU): the library's NormalGravity equals the oracle's
    restatement (pinned by WGS84 values in test_oracle_grav.py)."""
    GM, a, b, w = 3.986004418e14, 6378137.0, 6356752.314245, 7.292115e-5
    for lat in (-90.0, -45.5, 0.0, 12.25, 60.0, 90.0):
        assert abs(B.NormalGravity(lat, GM, w, a, b) - oracle.NormalGravity(lat, GM, w, a, b)) <= 1e-14 * 9.8
    assert B.NormalGravity(10.0, GM, 0.0, a, a) == GM / a ** 2


@pytest.mark.parametrize("L", [16, 200, 700, 2799])
def test_polar_skip_is_exact_and_tight(L):
    """The DH polar skip (host table, no GPU): at every ring slot, every
    4pi-normalised P_lm with m above kmax stays below 1e-32 for all l <= L
    (the oracle's P_lm, pinned to 60 digits near the poles at high degree in
    test_oracle.py), and the order kmax itself reaches it (nothing more is
    dropped than the bound allows)."""
    P = L + 2
    km = (ctypes.c_int * P)()
    assert _lib.lib().shtg_dh_polar_orders(L, km) == 0
    km = np.array(km[:])
    assert km[0] == 0 and km[-1] == L and np.all(np.diff(km) >= 0)
    n = 2 * L + 2
    step = max(1, P // (60 if L < 1000 else 15))
    for s in list(range(0, P, step)) + [P - 1]:
        z = 0.0 if s == n // 2 else np.cos(np.pi * s / n)
        p = oracle.plm(L, z)  # [l, m]
        mx = np.abs(p).max(axis=0)      # max over l per order m
        if km[s] < L:
            assert mx[km[s] + 1:].max() < 1e-32, (s, km[s])
        assert mx[km[s]] >= 1e-32, (s, km[s])
    # the skip is substantial at high degree: the recursion work it removes
    if L >= 700:
        work = sum((L - m + 1) for s in range(P) for m in range(km[s] + 1, L + 1))
        total = P * (L + 1) * (L + 2) / 2
        assert work / total > 0.1
