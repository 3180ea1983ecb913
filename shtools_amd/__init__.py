"""shtools_amd: SHTOOLS' real spherical-harmonic transforms on AMD MI355X.

``shtools_amd.backend`` is a drop-in pyshtools backend module (same names,
signatures, defaults and errors as the Fortran backend);
``register_with_pyshtools()`` adds it to pyshtools' backend registry when
pyshtools is importable.
"""
from . import backend
from .backend import (DHaj, GLQGridCoord, MakeGridDH, MakeGridGLQ, SHExpandDH, SHExpandGLQ, SHGLQ,
                      SHToolsError)

__all__ = ["backend", "DHaj", "GLQGridCoord", "MakeGridDH", "MakeGridGLQ", "SHExpandDH", "SHExpandGLQ", "SHGLQ",
           "SHToolsError", "register_with_pyshtools"]


def register_with_pyshtools(name="shtools_amd", select=True):
    """Insert this backend into pyshtools.backends (backends.py:26-31) and,
    optionally, make it the preferred backend (backends.py:87-118)."""
    import pyshtools.backends.backends as _b  # noqa: raises ImportError without pyshtools
    _b._available_backends[name] = backend
    if select:
        _b.select_preferred_backend(name)
    return backend
