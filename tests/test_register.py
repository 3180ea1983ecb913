"""register_with_pyshtools() against a stand-in of pyshtools' backend
registry (pyshtools is not installed here).

The stand-in package written below reproduces the parts of the registry
that a backend plugs into, with the behaviour of
pyshtools/backends/backends.py:25-118 (``_available_backends``,
``_preferred_backend``, ``select_preferred_backend`` that injects the
selected module's functions), pyshtools/expand/__init__.py:91-110 (the
nine injected transform names) and pyshtools/rotate/__init__.py:18-21
(``SHRotateRealCoef``).  No GPU is used: only which function objects the
names resolve to is checked, and that SHRotateRealCoef raises.
"""
import importlib
import sys
import textwrap

import pytest

EXPAND_NAMES = ["SHExpandGLQ", "SHExpandGLQC", "SHExpandDH", "SHExpandDHC", "MakeGridGLQ", "MakeGridGLQC",
                "MakeGridDH", "MakeGridDHC", "MakeGradientDH"]

STANDIN = {
    "pyshtools/__init__.py": """
        from . import backends, expand, rotate
    """,
    "pyshtools/backends/__init__.py": """
        from .backends import select_preferred_backend, preferred_backend, preferred_backend_module, backend_module
    """,
    "pyshtools/backends/shtools.py": """
        # the Fortran backend's names, as placeholders
        def _mk(name):
            def f(*a, **k):
                return ("fortran", name)
            f.__name__ = name
            return f
        for _n in %(names)r + ["SHRotateRealCoef"]:
            globals()[_n] = _mk(_n)
    """,
    "pyshtools/backends/backends.py": """
        from . import shtools
        _preferred_backend = "shtools"
        _available_backends = {"shtools": shtools}

        def preferred_backend():
            return _preferred_backend

        def preferred_backend_module():
            return _available_backends[_preferred_backend]

        def backend_module(backend=None, nthreads=None):
            if backend is None:
                return preferred_backend_module()
            backend = backend.lower()
            if backend not in _available_backends:
                raise RuntimeError
            return _available_backends[backend]

        def select_preferred_backend(backend=_preferred_backend, nthreads=None):
            global _preferred_backend
            backend = backend.lower()
            if backend not in _available_backends:
                raise RuntimeError
            _preferred_backend = backend
            from ..expand import inject_backend_specific_functions_for_expand
            inject_backend_specific_functions_for_expand()
            from ..rotate import inject_backend_specific_functions_for_rotate
            inject_backend_specific_functions_for_rotate()
    """,
    "pyshtools/expand/__init__.py": """
        from ..backends import backend_module, select_preferred_backend

        def inject_backend_specific_functions_for_expand():
            mod = backend_module()
            for name in %(names)r:
                globals()[name] = getattr(mod, name)

        select_preferred_backend()
    """,
    "pyshtools/rotate/__init__.py": """
        from ..backends import backend_module, select_preferred_backend

        def inject_backend_specific_functions_for_rotate():
            global SHRotateRealCoef
            SHRotateRealCoef = backend_module().SHRotateRealCoef

        select_preferred_backend()
    """,
}


@pytest.fixture
def standin(tmp_path, monkeypatch):
    for rel, src in STANDIN.items():
        p = tmp_path / rel
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(textwrap.dedent(src % {"names": EXPAND_NAMES} if "%(names)" in src else src))
    monkeypatch.syspath_prepend(str(tmp_path))
    for mod in [m for m in sys.modules if m == "pyshtools" or m.startswith("pyshtools.")]:
        monkeypatch.delitem(sys.modules, mod)
    pysh = importlib.import_module("pyshtools")
    yield pysh
    for mod in [m for m in sys.modules if m == "pyshtools" or m.startswith("pyshtools.")]:
        del sys.modules[mod]


def test_register_selects_and_injects_every_name(standin):
    import shtools_amd
    from shtools_amd import backend
    pysh = standin
    # before: the Fortran placeholders
    assert pysh.backends.preferred_backend() == "shtools"
    assert pysh.expand.MakeGridDH() == ("fortran", "MakeGridDH")
    mod = shtools_amd.register_with_pyshtools()
    assert mod is backend
    assert pysh.backends.preferred_backend() == "shtools_amd"
    assert pysh.backends.backend_module("shtools_amd") is backend
    assert pysh.backends.preferred_backend_module() is backend
    for name in EXPAND_NAMES:
        assert getattr(pysh.expand, name) is getattr(backend, name), name
    assert pysh.rotate.SHRotateRealCoef is backend.SHRotateRealCoef
    with pytest.raises(NotImplementedError, match="SHRotateRealCoef"):
        pysh.rotate.SHRotateRealCoef(None, None, None)
    # switching back restores the Fortran backend's names
    pysh.backends.select_preferred_backend("shtools")
    assert pysh.expand.SHExpandDH() == ("fortran", "SHExpandDH")
    assert pysh.rotate.SHRotateRealCoef() == ("fortran", "SHRotateRealCoef")


def test_register_without_selecting(standin):
    import shtools_amd
    pysh = standin
    shtools_amd.register_with_pyshtools(name="gpu", select=False)
    assert pysh.backends.preferred_backend() == "shtools"
    assert "gpu" in pysh.backends.backends._available_backends
    assert pysh.expand.MakeGridDH() == ("fortran", "MakeGridDH")
    pysh.backends.select_preferred_backend("GPU")   # names are case-insensitive
    from shtools_amd import backend
    assert pysh.expand.MakeGridGLQ is backend.MakeGridGLQ


def test_backend_module_signatures_match_f2py_defaults():
    """The injected names take the f2py defaults (src/pyshtools.pyf:235-276,
    278-319, 321-377, 379-422, 2021-2043)."""
    import inspect
    from shtools_amd import backend
    sig = {n: inspect.signature(getattr(backend, n)) for n in EXPAND_NAMES}
    assert list(sig["MakeGridDH"].parameters) == ["cilm", "lmax", "norm", "sampling", "csphase", "lmax_calc",
                                                  "extend"]
    assert sig["MakeGridDH"].parameters["norm"].default == 1
    assert sig["MakeGridDH"].parameters["sampling"].default == 1
    assert list(sig["SHExpandDH"].parameters) == ["griddh", "norm", "sampling", "csphase", "lmax_calc"]
    assert list(sig["MakeGridGLQ"].parameters) == ["cilm", "zero", "lmax", "norm", "csphase", "lmax_calc",
                                                   "extend"]
    assert list(sig["SHExpandGLQ"].parameters) == ["gridglq", "w", "zero", "norm", "csphase", "lmax_calc"]
    assert list(sig["MakeGradientDH"].parameters)[:5] == ["cilm", "lmax", "sampling", "lmax_calc", "extend"]
    assert sig["MakeGradientDH"].parameters["sampling"].default == 2
    for n in ("MakeGridDHC", "SHExpandDHC", "MakeGridGLQC", "SHExpandGLQC"):
        base = n[:-1]
        assert list(sig[n].parameters) == list(sig[base].parameters), n
