"""pyshtools backend module running the transforms on an MI355X.

Drop-in for the names that ``pyshtools.backends`` injects into
``pyshtools.expand`` (pyshtools/expand/__init__.py:91-110), with the f2py
signatures, defaults and errors of the Fortran backend
(src/pyshtools.pyf:235-276,321-377,424-438,1895-1905;
pyshtools/backends/shtools.py:189-229):

    MakeGridDH(cilm, lmax=cilm.shape[1]-1, norm=1, sampling=1, csphase=1,
               lmax_calc=lmax, extend=0) -> griddh
    SHExpandDH(griddh, norm=1, sampling=1, csphase=1, lmax_calc=n//2-1) -> cilm
    MakeGridGLQ(cilm, zero, lmax=cilm.shape[1]-1, norm=1, csphase=1,
                lmax_calc=lmax, extend=0) -> gridglq
    SHExpandGLQ(gridglq, w, zero, norm=1, csphase=1,
                lmax_calc=gridglq.shape[0]-1) -> cilm
    SHGLQ(lmax) -> (zero, w);  GLQGridCoord(lmax, extend=0) -> (lat, lon);
    DHaj(n, extend=0) -> aj
    MakeGridDHC, SHExpandDHC, MakeGridGLQC, SHExpandGLQC: the complex
    transforms with the same arguments (src/pyshtools.pyf:278-319,379-422)
    MakeGradientDH(cilm, lmax, sampling=2, lmax_calc, extend=0, radius=1.0)
        -> (theta, phi)

Array arguments may be numpy arrays (copied to the GPU and back, like f2py
copies between C and Fortran order) or float64 (complex128 for the complex
transforms) torch tensors on a HIP device
(used in place; the result is a tensor on the same device, produced
asynchronously on the current stream).  Nonzero exit codes raise
``SHToolsError`` with the reference's messages.
"""
import ctypes

import numpy as np

from . import _lib

__all__ = [
    "SHToolsError", "MakeGridDH", "SHExpandDH", "MakeGridGLQ", "SHExpandGLQ", "SHGLQ", "GLQGridCoord", "DHaj",
    "NGLQ", "NGLQSH", "NGLQSHN",
    "PlmBar", "PlmSchmidt", "PlmON", "PLegendreA",
    "SHExpandDHC", "MakeGridDHC", "SHExpandGLQC", "MakeGridGLQC", "MakeGradientDH", "SHRotateRealCoef",
    "MakeGridDH_adjoint", "SHExpandDH_adjoint", "MakeGridGLQ_adjoint", "SHExpandGLQ_adjoint",
    "SHGLQ_plx", "MakeGridGLQ_plx", "SHExpandGLQ_plx", "MakeGridGLQC_plx", "SHExpandGLQC_plx",
]


class SHToolsError(Exception):
    pass


def _status_message(status):
    # pyshtools/backends/shtools.py:197-208
    if status == 1:
        return "Improper dimensions of input array."
    if status == 2:
        return "Improper bounds for input variable."
    if status == 3:
        return "Error allocating memory."
    if status == 4:
        return "File IO error."
    if status == 6:
        # this library's own status: no counterpart in the reference
        return ("Call not allowed under HIP graph capture: it needs device tables or array bindings that "
                "do not exist yet; make one eager call of the same shape on the same arrays first.")
    if status == 7:
        return "The no-allocation synthesis needs more scratch than its fixed device array (64 MiB) holds."
    if status == 8:
        return ("The no-allocation synthesis found no device tables for this degree, sampling and "
                "normalisation: make one ordinary MakeGridDH call with them on this device first.")
    return "Unhandled Fortran 95 error."


def _check(status):
    if status != 0:
        raise SHToolsError(_status_message(status))


def _torch():
    import torch
    return torch


def _is_tensor(x):
    try:
        torch = _torch()
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _device_of(*arrays):
    torch = _torch()
    for a in arrays:
        if _is_tensor(a) and a.is_cuda:
            return a.device
    if not torch.cuda.is_available():
        raise RuntimeError("shtools_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _to_dev(x, device):
    torch = _torch()
    if _is_tensor(x):
        t = x.to(device=device, dtype=torch.float64)
    else:
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)
    return t.contiguous()


def _host_f64(x):
    if _is_tensor(x):
        x = x.detach().to("cpu").numpy()
    return np.ascontiguousarray(x, dtype=np.float64).ravel()


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stream(device):
    return ctypes.c_void_p(_torch().cuda.current_stream(device).cuda_stream)


def _out_kind(*inputs):
    """How results go back: "cuda" (device tensors) if any input is a device
    tensor, else "cpu" for host tensors, else None (numpy arrays)."""
    kinds = [("cuda" if x.is_cuda else "cpu") for x in inputs if _is_tensor(x)]
    if "cuda" in kinds:
        return "cuda"
    return "cpu" if kinds else None


def _finish(out, kind):
    if kind == "cuda":
        return out
    if kind == "cpu":
        return out.cpu()
    return out.cpu().numpy()


# ------------------------------------------------------------ DH

def MakeGridDH(cilm, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0):
    """Evaluate real coefficients on a Driscoll-Healy grid (src/MakeGridDH.f95)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if sampling not in (1, 2) or extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    grid = torch.empty((g0, g1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegrid_dh(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), int(norm), int(sampling),
                                         int(csphase), int(lmax_calc), int(extend), grid.data_ptr(), g0, g1,
                                         _stream(dev))
    _check(st)
    return _finish(grid, as_tensor)


def SHExpandDH(griddh, norm=1, sampling=1, csphase=1, lmax_calc=None):
    """Expand a Driscoll-Healy grid into real coefficients (src/SHExpandDH.f95)."""
    torch = _torch()
    as_tensor = _out_kind(griddh)
    dev = _device_of(griddh)
    g = _to_dev(griddh, dev)
    if g.dim() != 2:
        raise SHToolsError(_status_message(1))
    n = g.shape[0]
    if lmax_calc is None:
        lmax_calc = n // 2 - 1
    if lmax_calc < 0:
        raise SHToolsError(_status_message(2))
    cilm = torch.empty((2, lmax_calc + 1, lmax_calc + 1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shexpand_dh(g.data_ptr(), g.shape[0], g.shape[1], int(norm), int(sampling),
                                         int(csphase), int(lmax_calc), cilm.data_ptr(), lmax_calc + 1,
                                         lmax_calc + 1, _stream(dev))
    _check(st)
    return _finish(cilm, as_tensor)


# ------------------------------------------------------------ GLQ

def _plx_dev(plx, lmax, dev):
    """The table plx of degree lmax as a contiguous float64 tensor on dev,
    shape ((lmax+1)(lmax+2)/2, lmax+1) (the reference's Fortran array
    plx(lmax+1, k) in C order)."""
    torch = _torch()
    t = plx.to(dev) if _is_tensor(plx) else torch.from_numpy(np.asarray(plx, dtype=np.float64)).to(dev)
    t = t.to(torch.float64).contiguous()
    if tuple(t.shape) != ((lmax + 1) * (lmax + 2) // 2, lmax + 1):
        raise SHToolsError(_status_message(1))
    return t


def SHGLQ_plx(lmax, norm=1, csphase=1, cnorm=0, device=None):
    """SHGLQ with its table of Legendre functions (src/SHGLQ.f95:202-257):
    (zero, w, plx), zero and w numpy arrays as SHGLQ returns them, plx a
    device tensor of shape ((lmax+1)(lmax+2)/2, lmax+1) holding P_lm at node
    i in plx[l(l+1)/2 + m, i], for MakeGridGLQ(..., plx=plx) and
    SHExpandGLQ(..., plx=plx)."""
    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if lmax < 0:
        raise SHToolsError(_status_message(2))
    zero = np.zeros(lmax + 1)
    w = np.zeros(lmax + 1)
    plx = torch.empty(((lmax + 1) * (lmax + 2) // 2, lmax + 1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        _check(_lib.lib().shtg_shglq_plx(int(lmax), _dptr(zero), _dptr(w), plx.data_ptr(), int(norm), int(csphase),
                                         int(cnorm), _stream(dev)))
    return zero, w, plx


def MakeGridGLQ(cilm, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """Evaluate real coefficients on a Gauss-Legendre grid (src/MakeGridGLQ.f95)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    z = _host_f64(zero)
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = torch.empty((g0, g1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegrid_glq(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), _dptr(z), z.size,
                                          int(norm), int(csphase), int(lmax_calc), int(extend), grid.data_ptr(),
                                          g0, g1, _stream(dev))
    _check(st)
    return _finish(grid, as_tensor)


def MakeGridGLQ_plx(cilm, plx, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """MakeGridGLQ with the table plx of SHGLQ_plx in place of the nodes (the
    reference's plx branch, src/MakeGridGLQ.f95:441-465): the functions are
    read from the table, which fixes normalisation and phase (norm sets the
    value of a degree-0 field)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    p = _plx_dev(plx, lmax, dev)
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = torch.empty((g0, g1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegrid_glq_plx(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), p.data_ptr(),
                                              int(norm), int(csphase), int(lmax_calc), int(extend), grid.data_ptr(),
                                              g0, g1, _stream(dev))
    _check(st)
    return _finish(grid, as_tensor)


def SHExpandGLQ_plx(gridglq, w, plx, norm=1, csphase=1, lmax_calc=None):
    """SHExpandGLQ with the table plx of SHGLQ_plx in place of the nodes (the
    reference's plx branch, src/SHExpandGLQ.f95:419-441); norm selects the
    normalisation epilogue."""
    torch = _torch()
    as_tensor = _out_kind(gridglq)
    dev = _device_of(gridglq)
    g = _to_dev(gridglq, dev)
    if g.dim() != 2:
        raise SHToolsError(_status_message(1))
    wq = _host_f64(w)
    lmax = g.shape[0] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if lmax_calc < 0:
        raise SHToolsError(_status_message(2))
    p = _plx_dev(plx, lmax, dev)
    cilm = torch.empty((2, lmax_calc + 1, lmax_calc + 1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shexpand_glq_plx(g.data_ptr(), g.shape[0], g.shape[1], _dptr(wq), wq.size, p.data_ptr(),
                                              int(norm), int(csphase), int(lmax_calc), cilm.data_ptr(),
                                              lmax_calc + 1, lmax_calc + 1, _stream(dev))
    _check(st)
    return _finish(cilm, as_tensor)


def SHExpandGLQ(gridglq, w, zero, norm=1, csphase=1, lmax_calc=None):
    """Expand a Gauss-Legendre grid into real coefficients (src/SHExpandGLQ.f95)."""
    torch = _torch()
    as_tensor = _out_kind(gridglq)
    dev = _device_of(gridglq)
    g = _to_dev(gridglq, dev)
    if g.dim() != 2:
        raise SHToolsError(_status_message(1))
    wq = _host_f64(w)
    z = _host_f64(zero)
    lmax = g.shape[0] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if lmax_calc < 0:
        raise SHToolsError(_status_message(2))
    cilm = torch.empty((2, lmax_calc + 1, lmax_calc + 1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shexpand_glq(g.data_ptr(), g.shape[0], g.shape[1], _dptr(wq), wq.size, _dptr(z),
                                          z.size, int(norm), int(csphase), int(lmax_calc), cilm.data_ptr(),
                                          lmax_calc + 1, lmax_calc + 1, _stream(dev))
    _check(st)
    return _finish(cilm, as_tensor)


# ------------------------------------------------------------ complex transforms

def _to_dev_c(x, device):
    torch = _torch()
    if _is_tensor(x):
        t = x.to(device=device, dtype=torch.complex128)
    else:
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.complex128), device=device)
    return t.contiguous()


def MakeGridDHC(cilm, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0):
    """Evaluate complex coefficients on a Driscoll-Healy grid (src/MakeGridDHC.f95).
    cilm[0] holds the orders +m, cilm[1] the orders -m."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev_c(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if sampling not in (1, 2) or extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    grid = torch.empty((g0, g1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegrid_dhc(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), int(norm), int(sampling),
                                          int(csphase), int(lmax_calc), int(extend), grid.data_ptr(), g0, g1,
                                          _stream(dev))
    _check(st)
    return _finish(grid, as_tensor)


def SHExpandDHC(griddh, norm=1, sampling=1, csphase=1, lmax_calc=None):
    """Expand a complex Driscoll-Healy grid into complex coefficients (src/SHExpandDHC.f95)."""
    torch = _torch()
    as_tensor = _out_kind(griddh)
    dev = _device_of(griddh)
    g = _to_dev_c(griddh, dev)
    if g.dim() != 2:
        raise SHToolsError(_status_message(1))
    n = g.shape[0]
    if lmax_calc is None:
        lmax_calc = n // 2 - 1
    if lmax_calc < 0:
        raise SHToolsError(_status_message(2))
    cilm = torch.empty((2, lmax_calc + 1, lmax_calc + 1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shexpand_dhc(g.data_ptr(), g.shape[0], g.shape[1], int(norm), int(sampling),
                                          int(csphase), int(lmax_calc), cilm.data_ptr(), lmax_calc + 1,
                                          lmax_calc + 1, _stream(dev))
    _check(st)
    return _finish(cilm, as_tensor)


def MakeGridGLQC(cilm, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """Evaluate complex coefficients on a Gauss-Legendre grid (src/MakeGridGLQC.f95)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev_c(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    z = _host_f64(zero)
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = torch.empty((g0, g1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegrid_glqc(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), _dptr(z), z.size,
                                           int(norm), int(csphase), int(lmax_calc), int(extend), grid.data_ptr(),
                                           g0, g1, _stream(dev))
    _check(st)
    return _finish(grid, as_tensor)


def SHExpandGLQC(gridglq, w, zero, norm=1, csphase=1, lmax_calc=None):
    """Expand a complex Gauss-Legendre grid into complex coefficients (src/SHExpandGLQC.f95)."""
    torch = _torch()
    as_tensor = _out_kind(gridglq)
    dev = _device_of(gridglq)
    g = _to_dev_c(gridglq, dev)
    if g.dim() != 2:
        raise SHToolsError(_status_message(1))
    wq = _host_f64(w)
    z = _host_f64(zero)
    lmax = g.shape[0] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if lmax_calc < 0:
        raise SHToolsError(_status_message(2))
    cilm = torch.empty((2, lmax_calc + 1, lmax_calc + 1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shexpand_glqc(g.data_ptr(), g.shape[0], g.shape[1], _dptr(wq), wq.size, _dptr(z),
                                           z.size, int(norm), int(csphase), int(lmax_calc), cilm.data_ptr(),
                                           lmax_calc + 1, lmax_calc + 1, _stream(dev))
    _check(st)
    return _finish(cilm, as_tensor)


def MakeGridGLQC_plx(cilm, plx, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """MakeGridGLQC with the table plx of SHGLQ_plx (cnorm=1) in place of the
    nodes (the reference's plx branch, src/MakeGridGLQC.f95:443-468)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev_c(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    p = _plx_dev(plx, lmax, dev)
    g0, g1 = lmax + 1, 2 * lmax + 1 + extend
    grid = torch.empty((g0, g1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegrid_glqc_plx(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), p.data_ptr(),
                                               int(norm), int(csphase), int(lmax_calc), int(extend), grid.data_ptr(),
                                               g0, g1, _stream(dev))
    _check(st)
    return _finish(grid, as_tensor)


def SHExpandGLQC_plx(gridglq, w, plx, norm=1, csphase=1, lmax_calc=None):
    """SHExpandGLQC with the table plx of SHGLQ_plx (cnorm=1) in place of the
    nodes (the reference's plx branch of src/SHExpandGLQC.f95)."""
    torch = _torch()
    as_tensor = _out_kind(gridglq)
    dev = _device_of(gridglq)
    g = _to_dev_c(gridglq, dev)
    if g.dim() != 2:
        raise SHToolsError(_status_message(1))
    wq = _host_f64(w)
    lmax = g.shape[0] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if lmax_calc < 0:
        raise SHToolsError(_status_message(2))
    p = _plx_dev(plx, lmax, dev)
    cilm = torch.empty((2, lmax_calc + 1, lmax_calc + 1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shexpand_glqc_plx(g.data_ptr(), g.shape[0], g.shape[1], _dptr(wq), wq.size,
                                               p.data_ptr(), int(norm), int(csphase), int(lmax_calc), cilm.data_ptr(),
                                               lmax_calc + 1, lmax_calc + 1, _stream(dev))
    _check(st)
    return _finish(cilm, as_tensor)


# ------------------------------------------------------------ gradient

def MakeGradientDH(cilm, lmax=None, sampling=2, lmax_calc=None, extend=0, radius=1.0):
    """Horizontal gradient (theta, phi components) of a 4pi-normalised real
    field on DH grids (src/MakeGradientDH.f95; f2py src/pyshtools.pyf:2021-2043)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if sampling not in (1, 2) or extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    theta = torch.empty((g0, g1), dtype=torch.float64, device=dev)
    phi = torch.empty((g0, g1), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegradient_dh(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), int(sampling),
                                             int(lmax_calc), int(extend), float(radius), theta.data_ptr(),
                                             phi.data_ptr(), g0, g1, _stream(dev))
    _check(st)
    return _finish(theta, as_tensor), _finish(phi, as_tensor)


def MakeGravGridDH(cilm, gm, r0, lmax=None, a=None, f=0.0, sampling=2, lmax_calc=None, omega=0.0,
                   normal_gravity=1, normal_gravity_gm=None, extend=0):
    """Gravity (rad, theta, phi), its magnitude and the potential on a
    flattened ellipsoid from 4pi-normalised potential coefficients
    (src/MakeGravGridDH.f95; f2py src/pyshtools.pyf:1440-1477, whose
    defaults these are: a = r0, f = 0, sampling 2, omega 0, normal gravity
    removed).  Returns (rad, theta, phi, total, pot)."""
    torch = _torch()
    as_tensor = _out_kind(cilm)
    dev = _device_of(cilm)
    c = _to_dev(cilm, dev)
    if c.dim() != 3 or c.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax is None:
        lmax = c.shape[1] - 1
    if lmax_calc is None:
        lmax_calc = lmax
    if a is None:
        a = r0
    if normal_gravity_gm is None:
        normal_gravity_gm = gm
    if sampling not in (1, 2) or extend not in (0, 1) or lmax < 0:
        raise SHToolsError(_status_message(2))
    g0 = 2 * (lmax + 1) + extend
    g1 = sampling * 2 * (lmax + 1) + extend
    outs = [torch.empty((g0, g1), dtype=torch.float64, device=dev) for _ in range(5)]
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_makegravgrid_dh(c.data_ptr(), c.shape[1], c.shape[2], int(lmax), float(gm), float(r0),
                                             float(a), float(f), int(sampling), int(lmax_calc), float(omega),
                                             int(normal_gravity), float(normal_gravity_gm), int(extend),
                                             *[o.data_ptr() for o in outs], g0, g1, _stream(dev))
    _check(st)
    return tuple(_finish(o, as_tensor) for o in outs)


def NormalGravity(geocentric_lat, gm, omega, a, b):
    """Normal gravity of a rotating ellipsoid (src/NormalGravity.f95)."""
    out = ctypes.c_double(0.0)
    _check(_lib.lib().shtg_normal_gravity(float(geocentric_lat), float(gm), float(omega), float(a), float(b),
                                          ctypes.byref(out)))
    return out.value


# ------------------------------------------------------------ products

def SHMultiply(sh1, sh2, lmax1=None, lmax2=None, norm=1, csphase=1, precomp=0):
    """Coefficients of degree lmax1 + lmax2 of the product of two real fields
    (src/SHMultiply.f95; f2py src/pyshtools.pyf:596-620).  precomp, which the
    f2py interface fixes at 0, is the Fortran routine's optional argument:
    1 builds SHGLQ's table of the Legendre functions at the nodes and reads
    it in the three transforms (src/SHMultiply.f95:281-325)."""
    torch = _torch()
    as_tensor = _out_kind(sh1, sh2)
    dev = _device_of(sh1, sh2)
    a = _to_dev(sh1, dev)
    b = _to_dev(sh2, dev)
    if a.dim() != 3 or b.dim() != 3 or a.shape[0] < 2 or b.shape[0] < 2:
        raise SHToolsError(_status_message(1))
    if lmax1 is None:
        lmax1 = a.shape[1] - 1
    if lmax2 is None:
        lmax2 = b.shape[1] - 1
    n = int(lmax1) + int(lmax2) + 1
    if n < 1:
        raise SHToolsError(_status_message(2))
    out = torch.empty((2, n, n), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().shtg_shmultiply(out.data_ptr(), n, n, a.data_ptr(), a.shape[1], a.shape[2], int(lmax1),
                                        b.data_ptr(), b.shape[1], b.shape[2], int(lmax2), int(precomp), int(norm),
                                        int(csphase), _stream(dev))
    _check(st)
    return _finish(out, as_tensor)


# ------------------------------------------------------------ host helpers

def NGLQ(degree):
    """Gauss-Legendre points integrating a polynomial of this degree exactly
    (src/PreGLQ.f95:153-178)."""
    return _nglq_call("shtg_nglq", degree)


def NGLQSH(degree):
    """Gauss-Legendre points integrating a field of this degree exactly
    (src/PreGLQ.f95:181-205)."""
    return _nglq_call("shtg_nglqsh", degree)


def NGLQSHN(degree, n):
    """Gauss-Legendre points integrating a field of this degree raised to the
    power n exactly (src/PreGLQ.f95:208-233)."""
    return _nglq_call("shtg_nglqshn", degree, n)


def _nglq_call(name, *args):
    r = getattr(_lib.lib(), name)(*[int(a) for a in args])
    if r < 0:
        raise SHToolsError(_status_message(2))
    return r


def SHGLQ(lmax):
    """Gauss-Legendre nodes (descending) and weights for degree lmax (src/SHGLQ.f95)."""
    zero = np.zeros(lmax + 1)
    w = np.zeros(lmax + 1)
    _check(_lib.lib().shtg_shglq(int(lmax), _dptr(zero), _dptr(w)))
    return zero, w


def GLQGridCoord(lmax, extend=0):
    """Latitudes and longitudes (degrees) of the GLQ grid (src/GLQGridCoord.f95)."""
    if extend not in (0, 1):
        raise SHToolsError(_status_message(2))
    lat = np.zeros(lmax + 1)
    lon = np.zeros(2 * lmax + 1 + extend)
    _check(_lib.lib().shtg_glqgridcoord(int(lmax), int(extend), _dptr(lat), _dptr(lon)))
    return lat, lon


def DHaj(n, extend=0):
    """Driscoll-Healy quadrature weights (src/DHaj.f95)."""
    if extend not in (0, 1):
        raise SHToolsError(_status_message(2))
    aj = np.zeros(n + extend)
    _check(_lib.lib().shtg_dhaj(int(n), int(extend), _dptr(aj)))
    return aj


def _plm_packed(lmax, z, norm, csphase):
    p = np.zeros((lmax + 1, lmax + 1))
    _check(_lib.lib().shtg_plm(int(lmax), int(norm), int(csphase), float(z), _dptr(p)))
    il, im = np.tril_indices(lmax + 1)
    return p[il, im]   # packed index l(l+1)/2 + m (src/PlmIndex.f95:41)


def PlmBar(lmax, z, csphase=1, cnorm=0):
    """4pi-normalised P_lm(z), packed (src/PlmBar.f95); cnorm=1 is not supported."""
    if cnorm != 0:
        raise NotImplementedError("PlmBar(cnorm=1) (complex normalisation) is not provided by shtools_amd")
    return _plm_packed(lmax, z, 1, csphase)


def PlmSchmidt(lmax, z, csphase=1, cnorm=0):
    """Schmidt semi-normalised P_lm(z), packed (src/PlmSchmidt.f95)."""
    if cnorm != 0:
        raise NotImplementedError("PlmSchmidt(cnorm=1) is not provided by shtools_amd")
    return _plm_packed(lmax, z, 2, csphase)


def PLegendreA(lmax, z, csphase=1):
    """Unnormalised P_lm(z), packed (src/PLegendreA.f95)."""
    return _plm_packed(lmax, z, 3, csphase)


def PlmON(lmax, z, csphase=1, cnorm=0):
    """Orthonormalised P_lm(z), packed (src/PlmON.f95)."""
    if cnorm != 0:
        raise NotImplementedError("PlmON(cnorm=1) is not provided by shtools_amd")
    return _plm_packed(lmax, z, 4, csphase)


# ------------------------------------------------------------ names injected but not on this path

def _not_provided(name):
    def f(*args, **kwargs):
        raise NotImplementedError(f"{name} is not provided by the shtools_amd backend "
                                  "(real scalar DH/GLQ transforms only); select the 'shtools' backend for it")
    f.__name__ = name
    return f


SHRotateRealCoef = _not_provided("SHRotateRealCoef")


# ------------------------------------------------------------ adjoints (device tensors only)

def _dev_tensor(x, name):
    torch = _torch()
    if not (_is_tensor(x) and x.is_cuda and x.dtype == torch.float64):
        raise TypeError(f"{name} must be a float64 tensor on a HIP device")
    return x.contiguous()


def MakeGridDH_adjoint(grid_bar, cilm_shape, lmax=None, norm=1, sampling=1, csphase=1, lmax_calc=None, extend=0):
    """Transpose of MakeGridDH: cilm_bar (shape cilm_shape) = S^T grid_bar,
    for the same arguments as the forward call that read a cilm of that
    shape."""
    torch = _torch()
    g = _dev_tensor(grid_bar, "grid_bar")
    d0, d1, d2 = (int(x) for x in cilm_shape)
    if lmax is None:
        lmax = d1 - 1
    if lmax_calc is None:
        lmax_calc = lmax
    cbar = torch.empty((d0, d1, d2), dtype=torch.float64, device=g.device)
    if d0 > 2:
        cbar[2:].zero_()
    with torch.cuda.device(g.device):
        st = _lib.lib().shtg_makegrid_dh_adjoint(g.data_ptr(), g.shape[0], g.shape[1], int(lmax), int(norm),
                                                 int(sampling), int(csphase), int(lmax_calc), int(extend),
                                                 cbar.data_ptr(), d1, d2, _stream(g.device))
    _check(st)
    return cbar


def SHExpandDH_adjoint(cilm_bar, grid_shape, norm=1, sampling=1, csphase=1, lmax_calc=None):
    """Transpose of SHExpandDH: grid_bar (shape grid_shape) = A^T cilm_bar."""
    torch = _torch()
    c = _dev_tensor(cilm_bar, "cilm_bar")
    g0, g1 = (int(x) for x in grid_shape)
    if lmax_calc is None:
        lmax_calc = g0 // 2 - 1
    gbar = torch.empty((g0, g1), dtype=torch.float64, device=c.device)
    with torch.cuda.device(c.device):
        st = _lib.lib().shtg_shexpand_dh_adjoint(c.data_ptr(), c.shape[1], c.shape[2], int(norm), int(sampling),
                                                 int(csphase), int(lmax_calc), gbar.data_ptr(), g0, g1,
                                                 _stream(c.device))
    _check(st)
    return gbar


def MakeGridGLQ_adjoint(grid_bar, cilm_shape, zero, lmax=None, norm=1, csphase=1, lmax_calc=None, extend=0):
    """Transpose of MakeGridGLQ."""
    torch = _torch()
    g = _dev_tensor(grid_bar, "grid_bar")
    z = _host_f64(zero)
    d0, d1, d2 = (int(x) for x in cilm_shape)
    if lmax is None:
        lmax = d1 - 1
    if lmax_calc is None:
        lmax_calc = lmax
    cbar = torch.empty((d0, d1, d2), dtype=torch.float64, device=g.device)
    if d0 > 2:
        cbar[2:].zero_()
    with torch.cuda.device(g.device):
        st = _lib.lib().shtg_makegrid_glq_adjoint(g.data_ptr(), g.shape[0], g.shape[1], int(lmax), _dptr(z), z.size,
                                                  int(norm), int(csphase), int(lmax_calc), int(extend),
                                                  cbar.data_ptr(), d1, d2, _stream(g.device))
    _check(st)
    return cbar


def SHExpandGLQ_adjoint(cilm_bar, grid_shape, w, zero, norm=1, csphase=1, lmax_calc=None):
    """Transpose of SHExpandGLQ."""
    torch = _torch()
    c = _dev_tensor(cilm_bar, "cilm_bar")
    wq = _host_f64(w)
    z = _host_f64(zero)
    g0, g1 = (int(x) for x in grid_shape)
    if lmax_calc is None:
        lmax_calc = g0 - 1
    gbar = torch.empty((g0, g1), dtype=torch.float64, device=c.device)
    with torch.cuda.device(c.device):
        st = _lib.lib().shtg_shexpand_glq_adjoint(c.data_ptr(), c.shape[1], c.shape[2], _dptr(wq), wq.size,
                                                  _dptr(z), z.size, int(norm), int(csphase), int(lmax_calc),
                                                  gbar.data_ptr(), g0, g1, _stream(c.device))
    _check(st)
    return gbar
