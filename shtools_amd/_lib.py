"""ctypes binding of libshtools_amd.so (the C ABI of include/shtools_amd.h).

There is no fallback: if the library is missing or no HIP device is
visible, calls raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SHTOOLS_AMD_LIB selects another build of the library (same C ABI).
LIBPATH = os.environ.get("SHTOOLS_AMD_LIB") or os.path.join(_HERE, "libshtools_amd.so")
_lib = None

_c_int = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

SIGNATURES = {
    "shtg_makegrid_dh": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _vp],
    "shtg_makegrid_dh_noalloc": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp,
                                 _c_int, _c_int, _vp],
    "shtg_shexpand_dh": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _vp],
    "shtg_makegrid_glq": [_vp, _c_int, _c_int, _c_int, _dp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int,
                          _c_int, _vp],
    "shtg_shexpand_glq": [_vp, _c_int, _c_int, _dp, _c_int, _dp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int,
                          _vp],
    "shtg_makegrid_dhc": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int,
                          _vp],
    "shtg_shexpand_dhc": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _vp],
    "shtg_makegrid_glqc": [_vp, _c_int, _c_int, _c_int, _dp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int,
                           _c_int, _vp],
    "shtg_shexpand_glqc": [_vp, _c_int, _c_int, _dp, _c_int, _dp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int,
                           _c_int, _vp],
    "shtg_makegradient_dh": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_double, _vp, _vp, _c_int,
                             _c_int, _vp],
    "shtg_makegravgrid_dh": [_vp, _c_int, _c_int, _c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, _c_int, _c_int, ctypes.c_double, _c_int, ctypes.c_double, _c_int,
                             _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp],
    "shtg_normal_gravity": [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                            _dp],
    "shtg_dh_polar_orders": [_c_int, ctypes.POINTER(ctypes.c_int)],
    "shtg_shmultiply": [_vp, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _c_int,
                        _c_int, _c_int, _vp],
    "shtg_makegrid_dh_adjoint": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int,
                                 _c_int, _vp],
    "shtg_shexpand_dh_adjoint": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _vp],
    "shtg_makegrid_glq_adjoint": [_vp, _c_int, _c_int, _c_int, _dp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp,
                                  _c_int, _c_int, _vp],
    "shtg_shexpand_glq_adjoint": [_vp, _c_int, _c_int, _dp, _c_int, _dp, _c_int, _c_int, _c_int, _c_int, _vp,
                                  _c_int, _c_int, _vp],
    "shtg_nglq": [_c_int],
    "shtg_nglqsh": [_c_int],
    "shtg_nglqshn": [_c_int, _c_int],
    "shtg_shglq": [_c_int, _dp, _dp],
    "shtg_shglq_plx": [_c_int, _dp, _dp, _vp, _c_int, _c_int, _c_int, _vp],
    "shtg_makegrid_glq_plx": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int,
                              _vp],
    "shtg_shexpand_glq_plx": [_vp, _c_int, _c_int, _dp, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _c_int, _c_int,
                              _vp],
    "shtg_makegrid_glqc_plx": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int,
                               _vp],
    "shtg_shexpand_glqc_plx": [_vp, _c_int, _c_int, _dp, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _c_int, _c_int,
                               _vp],
    "shtg_dhaj": [_c_int, _c_int, _dp],
    "shtg_glqgridcoord": [_c_int, _c_int, _dp, _dp],
    "shtg_plm": [_c_int, _c_int, _c_int, ctypes.c_double, _dp],
    "shtg_batch_create_dh": [_c_int, ctypes.POINTER(_c_int), _c_int, _c_int, _c_int, ctypes.POINTER(_vp)],
    "shtg_batch_create_dh_fields": [_c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), _c_int, _c_int,
                                    ctypes.POINTER(_vp)],
    "shtg_batch_create_glq": [_c_int, ctypes.POINTER(_c_int), _c_int, _c_int, ctypes.POINTER(_vp)],
    "shtg_batch_create_fields": [_c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), _c_int, _c_int,
                                 ctypes.POINTER(_vp)],
    "shtg_batch_makegrid_dh": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp],
    "shtg_batch_shexpand_dh": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp],
    "shtg_batch_makegrid_dh_adjoint": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp],
    "shtg_batch_shexpand_dh_adjoint": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp],
    "shtg_batch_destroy": [_vp],
    "shtg_batch_tune": [_vp, _c_int, _c_int, _vp],
    "shtg_batch_get_shapes": [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), _c_int],
    "shtg_batch_field_groups": [_vp, ctypes.POINTER(_c_int), _c_int],
    "shtg_batch_set_mfma": [_vp, _c_int, _vp],
    "shtg_batch_get_tuning": [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)],
    "shtg_batch_set_tuning": [_vp, _c_int, _c_int, _c_int],
    "shtg_batch_set_shapes": [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), _c_int],
    "shtg_profile_enable": [_c_int],
    "shtg_profile_collect": [_dp, ctypes.POINTER(ctypes.c_longlong), _c_int],
    "shtg_device_count": [],
    "shtg_version": [],
}


def lib():
    """Load the library (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise OSError(f"{LIBPATH} is missing: build it with `python -m shtools_amd.build`")
        L = ctypes.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.shtg_version.restype = ctypes.c_char_p
        _lib = L
    return _lib
