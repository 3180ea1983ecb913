"""ORACLE -- test infrastructure only.

ctypes front end of ``sht_oracle.c``, a CPU restatement of the SHTOOLS
Fortran transforms.  It mirrors the pyshtools f2py signatures and defaults
(src/pyshtools.pyf:235-276,321-377,1895-1905) so parity tests read like the
reference's own calls.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may use it, and only as the checker.

Every function is pinned by checks that do not retype its formulas
(tests/test_oracle.py): mpmath closed forms of the associated Legendre
functions in all four normalisations with and without the Condon-Shortley
phase, mpmath DH weights and Gauss-Legendre nodes, brute-force direct
summation of small expansions at every grid node, quadrature exactness,
round trips, numpy.fft for the longitude DFT and the known answers of the
reference's normalisation notebook.  The complex transforms (MakeGridDHC,
SHExpandDHC, MakeGridGLQC, SHExpandGLQC) are pinned through the real ones:
for coefficients of a real field, c_lm = (C_lm - i S_lm)/2 and
c_l,-m = (-1)^m (C_lm + i S_lm)/2 (c_l0 = C_l0), and the complex maps are
complex-linear, which fixes them completely.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libsht_oracle.so")
_lib = None


class OracleError(Exception):
    pass


_MESSAGES = {
    1: "Improper dimensions of input array.",
    2: "Improper bounds for input variable.",
    3: "Error allocating memory.",
    4: "File IO error.",
}


_REFSRC = "/root/reference/src"
_REFLIB = os.path.join(_HERE, "_ref", "libshtools_ref.so")
_FLANG = "/opt/rocm/lib/llvm/bin/flang"
_ref = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    # the reference's own helper routines, where its sources and flang exist
    if os.path.isdir(_REFSRC) and os.path.exists(_FLANG):
        r = subprocess.run(["make", "-s", "-C", _HERE, "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            print("oracle: building the reference helpers failed:\n" + r.stdout + r.stderr)


def ref_lib():
    """The reference's helper routines compiled from its sources (DHaj,
    PreGLQ, SHGLQ, GLQGridCoord, PlmBar, PlmSchmidt, PlmON, PLegendreA,
    NormalGravity) behind the C entry points of ref_glue.f95, or None where
    they have not been built."""
    global _ref
    if _ref is None and os.path.exists(_REFLIB):
        L = ctypes.CDLL(_REFLIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        i, d = ctypes.c_int, ctypes.c_double
        L.ref_dhaj.argtypes = [i, i, dp, ip]
        L.ref_preglq.argtypes = [d, d, i, dp, dp, ip]
        L.ref_shglq.argtypes = [i, dp, dp, ip]
        L.ref_shglq_plx.argtypes = [i, dp, dp, dp, i, i, i, ip]
        L.ref_glqgridcoord.argtypes = [i, i, dp, dp, ip, ip, ip]
        L.ref_plm.argtypes = [i, i, d, i, i, dp, ip]
        L.ref_normalgravity.argtypes = [d, d, d, d, d]
        L.ref_normalgravity.restype = d
        for name, nargs in (("ref_nglq", 1), ("ref_nglqsh", 1), ("ref_nglqshn", 2)):
            getattr(L, name).argtypes = [i] * nargs
            getattr(L, name).restype = i
        _ref = L
    return _ref


def _ref_call(fn, *args):
    st = ctypes.c_int(-1)
    fn(*args, ctypes.byref(st))
    return st.value


def ref_DHaj(n, extend=0):
    """(aj, exit status) of the reference's DHaj."""
    aj = np.zeros(n + extend)
    return aj, _ref_call(ref_lib().ref_dhaj, n, extend, _p(aj))


def ref_SHGLQ(lmax):
    """(zero, w, exit status) of the reference's SHGLQ (no plx table)."""
    zero, w = np.zeros(lmax + 1), np.zeros(lmax + 1)
    return zero, w, _ref_call(ref_lib().ref_shglq, lmax, _p(zero), _p(w))


def ref_SHGLQ_plx(lmax, norm=1, csphase=1, cnorm=0):
    """(zero, w, plx, exit status) of the reference's SHGLQ with its plx
    table, plx as an array [k, i] (the Fortran plx(i, k) in C order)."""
    zero, w = np.zeros(lmax + 1), np.zeros(lmax + 1)
    plx = np.zeros(((lmax + 1) * (lmax + 2) // 2, lmax + 1))
    st = _ref_call(ref_lib().ref_shglq_plx, lmax, _p(zero), _p(w), _p(plx), norm, csphase, cnorm)
    return zero, w, plx, st


def ref_PreGLQ(x1, x2, n):
    zero, w = np.zeros(n), np.zeros(n)
    return zero, w, _ref_call(ref_lib().ref_preglq, float(x1), float(x2), n, _p(zero), _p(w))


def ref_GLQGridCoord(lmax, extend=0):
    """(lat, lon, nlat, nlong, exit status) of the reference's GLQGridCoord."""
    lat, lon = np.zeros(lmax + 1), np.zeros(2 * lmax + 1 + extend)
    nlat, nlong, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(-1)
    ref_lib().ref_glqgridcoord(lmax, extend, _p(lat), _p(lon), ctypes.byref(nlat), ctypes.byref(nlong),
                               ctypes.byref(st))
    return lat, lon, nlat.value, nlong.value, st.value


# transform normalisation -> the reference's standalone routine (kind of ref_plm)
_PLM_KIND = {1: 1, 2: 2, 3: 4, 4: 3}


def ref_plm(lmax, z, norm=1, csphase=1, cnorm=0):
    """(p, exit status): the reference's PlmBar / PlmSchmidt / PLegendreA /
    PlmON (norm 1 / 2 / 3 / 4) at z, packed by l(l+1)/2 + m."""
    p = np.zeros((lmax + 1) * (lmax + 2) // 2)
    return p, _ref_call(ref_lib().ref_plm, _PLM_KIND[norm], lmax, float(z), csphase, cnorm, _p(p))


def ref_NormalGravity(geocentric_lat, gm, omega, a, b):
    return ref_lib().ref_normalgravity(float(geocentric_lat), float(gm), float(omega), float(a), float(b))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i = ctypes.c_int
        L.or_dhaj.argtypes = [i, i, dp]
        L.or_preglq.argtypes = [i, dp, dp]
        L.or_glqgridcoord.argtypes = [i, i, dp, dp]
        L.or_plm.argtypes = [i, i, i, ctypes.c_double, dp]
        L.or_makegriddh.argtypes = [dp, i, i, i, i, i, i, i, i, dp, i, i]
        L.or_shexpanddh.argtypes = [dp, i, i, i, i, i, i, dp, i, i]
        L.or_shexpanddh_par.argtypes = [dp, i, i, i, i, i, i, dp, i, i]
        L.or_shexpandglq_par.argtypes = [dp, i, i, i, dp, i, dp, i, i, i, i, dp, i, i]
        L.or_makegridglq.argtypes = [dp, i, i, i, dp, i, i, i, i, i, dp, i, i]
        L.or_shexpandglq.argtypes = [dp, i, i, i, dp, i, dp, i, i, i, i, dp, i, i]
        L.or_plx_table.argtypes = [i, dp, i, i, i, i, dp]
        L.or_makegridglq_plx.argtypes = [dp, i, i, i, dp, i, i, i, i, dp, i, i]
        L.or_shexpandglq_plx.argtypes = [dp, i, i, i, dp, i, dp, i, i, i, dp, i, i]
        L.or_makegridglqc_plx.argtypes = [dp, i, i, i, dp, i, i, i, i, dp, i, i]
        L.or_shexpandglqc_plx.argtypes = [dp, i, i, i, dp, i, dp, i, i, i, dp, i, i]
        L.or_makegriddhc.argtypes = [dp, i, i, i, i, i, i, i, i, dp, i, i]
        L.or_shexpanddhc.argtypes = [dp, i, i, i, i, i, i, dp, i, i]
        L.or_makegridglqc.argtypes = [dp, i, i, i, dp, i, i, i, i, i, dp, i, i]
        L.or_shexpandglqc.argtypes = [dp, i, i, i, dp, i, dp, i, i, i, i, dp, i, i]
        L.or_makegradientdh.argtypes = [dp, i, i, i, i, i, i, ctypes.c_double, dp, dp, i, i]
        d = ctypes.c_double
        L.or_makegravgriddh.argtypes = [dp, i, i, i, d, d, d, d, i, i, d, i, d, i, dp, dp, dp, dp, dp, i, i]
        L.or_shmultiply.argtypes = [dp, i, i, dp, i, i, i, dp, i, i, i, i, i, i]
        L.or_normalgravity.argtypes = [d, d, d, d, d]
        L.or_normalgravity.restype = d
        L.or_r2c.argtypes = [i, dp, dp]
        L.or_c2r.argtypes = [i, dp, dp]
        lp = ctypes.POINTER(ctypes.c_longlong)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_roundtrip_dh_batch.argtypes = [i, ip, lp, lp, i, i, i, dp, dp, dp, i]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(st):
    if st != 0:
        raise OracleError(_MESSAGES.get(st, "Unhandled Fortran 95 error."))


def DHaj(n, extend=0):
    aj = np.zeros(n + extend)
    _check(lib().or_dhaj(n, extend, _p(aj)))
    return aj


def SHGLQ(lmax):
    zero = np.zeros(lmax + 1)
    w = np.zeros(lmax + 1)
    _check(lib().or_preglq(lmax + 1, _p(zero), _p(w)))
    return zero, w


def GLQGridCoord(lmax, extend=0):
    lat = np.zeros(lmax + 1)
    lon = np.zeros(2 * lmax + 1 + extend)
    _check(lib().or_glqgridcoord(lmax, extend, _p(lat), _p(lon)))
    return lat, lon


def plm(lmax, z, norm=1, csphase=1):
    """P_lm(z) of the transforms' recursion, as an [l, m] array."""
    p = np.zeros((lmax + 1, lmax + 1))
    _check(lib().or_plm(lmax, norm, csphase, float(z), _p(p)))
    return p


def MakeGridDH(cilm, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0):
    cilm = np.ascontiguousarray(cilm, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    grid = np.zeros((g0, g1))
    _check(lib().or_makegriddh(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, norm, sampling,
                               csphase, lmax_calc, extend, _p(grid), g0, g1))
    return grid


def SHExpandDH(griddh, norm=1, sampling=1, csphase=1, lmax_calc=None, parallel=False):
    """parallel: sum the orders on all OpenMP threads (or_shexpanddh_par,
    bitwise the same coefficients; for high degrees)."""
    grid = np.ascontiguousarray(griddh, dtype=np.float64)
    n = grid.shape[0]
    if lmax_calc is None:
        lmax_calc = n // 2 - 1
    cilm = np.zeros((2, lmax_calc + 1, lmax_calc + 1))
    fn = lib().or_shexpanddh_par if parallel else lib().or_shexpanddh
    _check(fn(_p(grid), n, grid.shape[1], norm, sampling, csphase, lmax_calc,
              _p(cilm), lmax_calc + 1, lmax_calc + 1))
    return cilm


def MakeGridGLQ(cilm, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    cilm = np.ascontiguousarray(cilm, dtype=np.float64)
    zero = np.ascontiguousarray(zero, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = np.zeros((g0, g1))
    _check(lib().or_makegridglq(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, _p(zero), zero.size,
                                norm, csphase, lmax_calc, extend, _p(grid), g0, g1))
    return grid


def SHExpandGLQ(gridglq, w, zero, norm=1, csphase=1, lmax_calc=None, parallel=False):
    """parallel: as for SHExpandDH (or_shexpandglq_par)."""
    grid = np.ascontiguousarray(gridglq, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    zero = np.ascontiguousarray(zero, dtype=np.float64)
    lmax = grid.shape[0] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    cilm = np.zeros((2, lmax_calc + 1, lmax_calc + 1))
    fn = lib().or_shexpandglq_par if parallel else lib().or_shexpandglq
    _check(fn(_p(grid), grid.shape[0], grid.shape[1], lmax, _p(w), w.size,
              _p(zero), zero.size, norm, csphase, lmax_calc, _p(cilm),
              lmax_calc + 1, lmax_calc + 1))
    return cilm


def plx_table(lmax, zero, norm=1, csphase=1, cnorm=0):
    """SHGLQ's plx table at the nodes zero (src/SHGLQ.f95:202-257), as the
    array plx[k, i] = P_lm(zero[i]), k = l(l+1)/2 + m (the Fortran array
    plx(i, k) in C order)."""
    zero = np.ascontiguousarray(zero, dtype=np.float64)
    plx = np.zeros(((lmax + 1) * (lmax + 2) // 2, lmax + 1))
    _check(lib().or_plx_table(lmax, _p(zero), zero.size, norm, csphase, cnorm, _p(plx)))
    return plx


def MakeGridGLQ_plx(cilm, plx, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """MakeGridGLQ with the table plx in place of the nodes
    (src/MakeGridGLQ.f95:441-465)."""
    cilm = np.ascontiguousarray(cilm, dtype=np.float64)
    plx = np.ascontiguousarray(plx, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    assert plx.shape == ((lmax + 1) * (lmax + 2) // 2, lmax + 1)
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = np.zeros((g0, g1))
    _check(lib().or_makegridglq_plx(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, _p(plx), norm, csphase,
                                    lmax_calc, extend, _p(grid), g0, g1))
    return grid


def SHExpandGLQ_plx(gridglq, w, plx, norm=1, csphase=1, lmax_calc=None):
    """SHExpandGLQ with the table plx in place of the nodes
    (src/SHExpandGLQ.f95:419-441)."""
    grid = np.ascontiguousarray(gridglq, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    plx = np.ascontiguousarray(plx, dtype=np.float64)
    lmax = grid.shape[0] - 1
    assert plx.shape == ((lmax + 1) * (lmax + 2) // 2, lmax + 1)
    if lmax_calc is None:
        lmax_calc = lmax
    cilm = np.zeros((2, lmax_calc + 1, lmax_calc + 1))
    _check(lib().or_shexpandglq_plx(_p(grid), grid.shape[0], grid.shape[1], lmax, _p(w), w.size, _p(plx), norm,
                                    csphase, lmax_calc, _p(cilm), lmax_calc + 1, lmax_calc + 1))
    return cilm


def MakeGridDHC(cilm, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0):
    cilm = np.ascontiguousarray(cilm, dtype=np.complex128)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    grid = np.zeros((g0, g1), dtype=np.complex128)
    _check(lib().or_makegriddhc(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, norm, sampling,
                                csphase, lmax_calc, extend, _p(grid), g0, g1))
    return grid


def SHExpandDHC(griddh, norm=1, sampling=1, csphase=1, lmax_calc=None):
    grid = np.ascontiguousarray(griddh, dtype=np.complex128)
    n = grid.shape[0]
    if lmax_calc is None:
        lmax_calc = n // 2 - 1
    cilm = np.zeros((2, lmax_calc + 1, lmax_calc + 1), dtype=np.complex128)
    _check(lib().or_shexpanddhc(_p(grid), n, grid.shape[1], norm, sampling, csphase, lmax_calc,
                                _p(cilm), lmax_calc + 1, lmax_calc + 1))
    return cilm


def MakeGridGLQC(cilm, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    cilm = np.ascontiguousarray(cilm, dtype=np.complex128)
    zero = np.ascontiguousarray(zero, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = np.zeros((g0, g1), dtype=np.complex128)
    _check(lib().or_makegridglqc(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, _p(zero), zero.size,
                                 norm, csphase, lmax_calc, extend, _p(grid), g0, g1))
    return grid


def SHExpandGLQC(gridglq, w, zero, norm=1, csphase=1, lmax_calc=None):
    grid = np.ascontiguousarray(gridglq, dtype=np.complex128)
    w = np.ascontiguousarray(w, dtype=np.float64)
    zero = np.ascontiguousarray(zero, dtype=np.float64)
    lmax = grid.shape[0] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    cilm = np.zeros((2, lmax_calc + 1, lmax_calc + 1), dtype=np.complex128)
    _check(lib().or_shexpandglqc(_p(grid), grid.shape[0], grid.shape[1], lmax, _p(w), w.size,
                                 _p(zero), zero.size, norm, csphase, lmax_calc, _p(cilm),
                                 lmax_calc + 1, lmax_calc + 1))
    return cilm


def MakeGridGLQC_plx(cilm, plx, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """MakeGridGLQC with the table plx (of complex normalisation) in place of
    the nodes."""
    cilm = np.ascontiguousarray(cilm, dtype=np.complex128)
    plx = np.ascontiguousarray(plx, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    assert plx.shape == ((lmax + 1) * (lmax + 2) // 2, lmax + 1)
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = np.zeros((g0, g1), dtype=np.complex128)
    _check(lib().or_makegridglqc_plx(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, _p(plx), norm, csphase,
                                     lmax_calc, extend, _p(grid), g0, g1))
    return grid


def SHExpandGLQC_plx(gridglq, w, plx, norm=1, csphase=1, lmax_calc=None):
    """SHExpandGLQC with the table plx (of complex normalisation) in place of
    the nodes."""
    grid = np.ascontiguousarray(gridglq, dtype=np.complex128)
    w = np.ascontiguousarray(w, dtype=np.float64)
    plx = np.ascontiguousarray(plx, dtype=np.float64)
    lmax = grid.shape[0] - 1
    assert plx.shape == ((lmax + 1) * (lmax + 2) // 2, lmax + 1)
    if lmax_calc is None:
        lmax_calc = lmax
    cilm = np.zeros((2, lmax_calc + 1, lmax_calc + 1), dtype=np.complex128)
    _check(lib().or_shexpandglqc_plx(_p(grid), grid.shape[0], grid.shape[1], lmax, _p(w), w.size, _p(plx), norm,
                                     csphase, lmax_calc, _p(cilm), lmax_calc + 1, lmax_calc + 1))
    return cilm


def MakeGradientDH(cilm, lmax=None, sampling=2, lmax_calc=None, extend=0, radius=1.0):
    cilm = np.ascontiguousarray(cilm, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    theta = np.zeros((g0, g1))
    phi = np.zeros((g0, g1))
    _check(lib().or_makegradientdh(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, sampling, lmax_calc, extend,
                                   float(radius), _p(theta), _p(phi), g0, g1))
    return theta, phi


def MakeGravGridDH(cilm, gm, r0, lmax=None, a=None, f=0.0, sampling=2, lmax_calc=None, omega=0.0,
                   normal_gravity=1, normal_gravity_gm=None, extend=0):
    """(rad, theta, phi, total, pot) grids; argument order and defaults of
    pyshtools' MakeGravGridDH (src/pyshtools.pyf:1440-1477)."""
    cilm = np.ascontiguousarray(cilm, dtype=np.float64)
    if lmax is None:
        lmax = cilm.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if a is None:
        a = r0
    if normal_gravity_gm is None:
        normal_gravity_gm = gm
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    out = [np.zeros((g0, g1)) for _ in range(5)]
    _check(lib().or_makegravgriddh(_p(cilm), cilm.shape[1], cilm.shape[2], lmax, float(gm), float(r0), float(a),
                                   float(f), sampling, lmax_calc, float(omega), normal_gravity,
                                   float(normal_gravity_gm), extend, *[_p(x) for x in out], g0, g1))
    return tuple(out)


def SHMultiply(sh1, sh2, lmax1=None, lmax2=None, norm=1, csphase=1, precomp=0):
    """Coefficients of the product of two fields (src/SHMultiply.f95;
    f2py src/pyshtools.pyf:596-620, which fixes precomp = 0; precomp = 1 is
    the table branch, src/SHMultiply.f95:281-325)."""
    sh1 = np.ascontiguousarray(sh1, dtype=np.float64)
    sh2 = np.ascontiguousarray(sh2, dtype=np.float64)
    if lmax1 is None:
        lmax1 = sh1.shape[1] - 1
    if lmax2 is None:
        lmax2 = sh2.shape[1] - 1
    n = lmax1 + lmax2 + 1
    out = np.zeros((2, n, n))
    _check(lib().or_shmultiply(_p(out), n, n, _p(sh1), sh1.shape[1], sh1.shape[2], lmax1, _p(sh2), sh2.shape[1],
                               sh2.shape[2], lmax2, precomp, norm, csphase))
    return out


def NormalGravity(geocentric_lat, gm, omega, a, b):
    return lib().or_normalgravity(float(geocentric_lat), float(gm), float(omega), float(a), float(b))


def rfft(x):
    """FFTW-style unnormalised r2c of a real vector."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    out = np.zeros(2 * (n // 2 + 1))
    lib().or_r2c(n, _p(x), _p(out))
    return out[0::2] + 1j * out[1::2]


def irfft(c, n):
    """FFTW-style unnormalised c2r to length n."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    ri = np.zeros(2 * (n // 2 + 1))
    ri[0::2] = c.real[: n // 2 + 1]
    ri[1::2] = c.imag[: n // 2 + 1]
    out = np.zeros(n)
    lib().or_c2r(n, _p(ri), _p(out))
    return out


def roundtrip_dh_batch(cilms, norm=1, sampling=2, csphase=1, nthreads=1, return_grids=False):
    """MakeGridDH -> SHExpandDH for a list of cilm arrays, OpenMP over fields.
    Returns the coefficients (and, with return_grids, the grids first)."""
    lmaxs = np.array([c.shape[1] - 1 for c in cilms], dtype=np.int32)
    csz = np.array([2 * (L + 1) ** 2 for L in lmaxs], dtype=np.int64)
    gsz = np.array([(2 * L + 2) * sampling * (2 * L + 2) for L in lmaxs], dtype=np.int64)
    coff = np.concatenate([[0], np.cumsum(csz)[:-1]]).astype(np.int64)
    goff = np.concatenate([[0], np.cumsum(gsz)[:-1]]).astype(np.int64)
    cin = np.concatenate([np.ascontiguousarray(c, dtype=np.float64).ravel() for c in cilms])
    grid = np.zeros(int(gsz.sum()))
    out = np.zeros_like(cin)
    st = lib().or_roundtrip_dh_batch(
        len(cilms), lmaxs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
        coff.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
        goff.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
        norm, sampling, csphase, _p(cin), _p(grid), _p(out), nthreads)
    _check(st)
    coefs = [out[o:o + s].reshape(2, L + 1, L + 1) for o, s, L in zip(coff, csz, lmaxs)]
    if not return_grids:
        return coefs
    grids = [grid[o:o + s].reshape(2 * L + 2, sampling * (2 * L + 2)) for o, s, L in zip(goff, gsz, lmaxs)]
    return grids, coefs
