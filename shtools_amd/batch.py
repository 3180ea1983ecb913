"""Batched spherical-harmonic transforms of many fields of uneven degree.

One plan per set of degrees and grids; every call covers all fields with a
handful of launches (pack, Legendre, ring FFTs grouped by size, unpack)
whose work items are ordered longest first.  Per field the results are
those of ``backend.MakeGridDH(c, sampling=s, norm=n, csphase=p)`` and
``backend.SHExpandDH(g, sampling=s, norm=n, csphase=p)`` for DH fields, and
of ``backend.MakeGridGLQ(c, zero, norm=n, csphase=p)`` and
``backend.SHExpandGLQ(g, w, zero, norm=n, csphase=p)`` with ``zero, w =
backend.SHGLQ(lmax)`` for GLQ fields.
"""
import ctypes

from . import _lib
from .backend import SHToolsError, _status_message


def _ptr_array(tensors):
    arr = (ctypes.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr


def _read_cache(path):
    import json
    import os
    if not path or not os.path.exists(path):
        return {}
    with open(path) as fh:
        try:
            return json.load(fh)
        except ValueError:
            return {}


def _write_cache(path, key, value):
    import json
    import os
    data = _read_cache(path)
    data[key] = value
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


# grid codes of the C interface (shtg_batch_create_fields)
_GRID_CODES = {"DH1": 1, "DH2": 2, "GLQ": 3, 1: 1, 2: 2, 3: 3}


class SHBatch:
    """Plan for the fields of degrees ``lmax`` on grids ``grid``: one of
    "DH1", "DH2" (DH with sampling 1 or 2) or "GLQ" for every field, or a
    sequence with one per field."""

    def __init__(self, lmax, grid="DH2", norm=1, csphase=1, device=None):
        import torch
        self.torch = torch
        self.lmax = [int(L) for L in lmax]
        grids = list(grid) if isinstance(grid, (list, tuple)) else [grid] * len(self.lmax)
        if len(grids) != len(self.lmax) or any(g not in _GRID_CODES for g in grids):
            raise SHToolsError(_status_message(1 if len(grids) != len(self.lmax) else 2))
        self.grid = [_GRID_CODES[g] for g in grids]
        self.norm, self.csphase = int(norm), int(csphase)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        arr = (ctypes.c_int * len(self.lmax))(*self.lmax)
        codes = (ctypes.c_int * len(self.lmax))(*self.grid)
        plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            st = _lib.lib().shtg_batch_create_fields(len(self.lmax), arr, codes, self.norm, self.csphase,
                                                     ctypes.byref(plan))
        if st != 0:
            raise SHToolsError(_status_message(st))
        self._plan = plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        self._plan = None
        # at interpreter shutdown the module globals may already be gone
        if plan and _lib is not None and getattr(_lib, "lib", None) is not None:
            try:
                _lib.lib().shtg_batch_destroy(plan)
            except Exception:
                pass

    def grid_shape(self, i):
        L = self.lmax[i]
        if self.grid[i] == 3:
            return (L + 1, 2 * L + 1)
        n = 2 * L + 2
        return (n, self.grid[i] * n)

    def cilm_shape(self, i):
        L = self.lmax[i]
        return (2, L + 1, L + 1)

    def _check(self, tensors, shapes):
        torch = self.torch
        if len(tensors) != len(self.lmax):
            raise SHToolsError(_status_message(1))
        for t, shp in zip(tensors, shapes):
            if (t.dtype != torch.float64 or not t.is_contiguous() or tuple(t.shape) != shp
                    or t.device != self.device):
                raise SHToolsError(_status_message(1))

    def makegrid(self, cilms, out=None):
        """Synthesis of every field; returns the list of grids (``out`` if given)."""
        torch = self.torch
        self._check(cilms, [self.cilm_shape(i) for i in range(len(self.lmax))])
        if out is None:
            out = [torch.empty(self.grid_shape(i), dtype=torch.float64, device=self.device)
                   for i in range(len(self.lmax))]
        self._check(out, [self.grid_shape(i) for i in range(len(self.lmax))])
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        with torch.cuda.device(self.device):
            st = _lib.lib().shtg_batch_makegrid_dh(self._plan, _ptr_array(cilms), _ptr_array(out), stream)
        if st != 0:
            raise SHToolsError(_status_message(st))
        return out

    def expand(self, grids, out=None):
        """Analysis of every field; returns the list of coefficient arrays."""
        torch = self.torch
        self._check(grids, [self.grid_shape(i) for i in range(len(self.lmax))])
        if out is None:
            out = [torch.empty(self.cilm_shape(i), dtype=torch.float64, device=self.device)
                   for i in range(len(self.lmax))]
        self._check(out, [self.cilm_shape(i) for i in range(len(self.lmax))])
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        with torch.cuda.device(self.device):
            st = _lib.lib().shtg_batch_shexpand_dh(self._plan, _ptr_array(grids), _ptr_array(out), stream)
        if st != 0:
            raise SHToolsError(_status_message(st))
        return out

    def makegrid_adjoint(self, grid_bars, out=None):
        """Transpose of makegrid per field: cilm_bar[i] = S_i^T grid_bar[i]."""
        torch = self.torch
        self._check(grid_bars, [self.grid_shape(i) for i in range(len(self.lmax))])
        if out is None:
            out = [torch.empty(self.cilm_shape(i), dtype=torch.float64, device=self.device)
                   for i in range(len(self.lmax))]
        self._check(out, [self.cilm_shape(i) for i in range(len(self.lmax))])
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        with torch.cuda.device(self.device):
            st = _lib.lib().shtg_batch_makegrid_dh_adjoint(self._plan, _ptr_array(grid_bars), _ptr_array(out),
                                                           stream)
        if st != 0:
            raise SHToolsError(_status_message(st))
        return out

    def expand_adjoint(self, cilm_bars, out=None):
        """Transpose of expand per field: grid_bar[i] = A_i^T cilm_bar[i]."""
        torch = self.torch
        self._check(cilm_bars, [self.cilm_shape(i) for i in range(len(self.lmax))])
        if out is None:
            out = [torch.empty(self.grid_shape(i), dtype=torch.float64, device=self.device)
                   for i in range(len(self.lmax))]
        self._check(out, [self.grid_shape(i) for i in range(len(self.lmax))])
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        with torch.cuda.device(self.device):
            st = _lib.lib().shtg_batch_shexpand_dh_adjoint(self._plan, _ptr_array(cilm_bars), _ptr_array(out),
                                                           stream)
        if st != 0:
            raise SHToolsError(_status_message(st))
        return out

    def capture(self, cilms, grids, outs, warmup=1):
        """Capture one synthesis + analysis of the batch (cilms -> grids ->
        outs) in a HIP graph and return it (``graph.replay()`` reruns the whole
        step on the current stream).  The calls are made once eagerly first,
        which binds the plan to these tensors; the graph holds no memset and
        no allocation, and must be replayed with the same tensors alive."""
        torch = self.torch
        for _ in range(max(1, warmup)):
            self.makegrid(cilms, out=grids)
            self.expand(grids, out=outs)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.makegrid(cilms, out=grids)
            self.expand(grids, out=outs)
        return graph

    def tune(self, cilms, grids, outs, reps=3, cache=None, analysis=False):
        """Autotune this plan's launches on these arrays (timed on the
        device): ring-FFT workgroup sizes, ring pairs per lane of the
        synthesis Legendre waves and the MFMA bucket size (shtg_batch_tune;
        none of these changes any result), and with analysis=True ring pairs
        per lane of the analysis waves (coefficients then change by
        rounding); then recompute grids and outs.  Returns the choice (see
        tuning()).

        cache: path of a JSON file (default: $SHTOOLS_AMD_TUNE_CACHE, if set)
        in which choices are kept per GPU architecture and plan (the fields'
        degrees and samplings); a plan found there takes the stored choice
        instead of being timed."""
        import os
        torch = self.torch
        cache = cache or os.environ.get("SHTOOLS_AMD_TUNE_CACHE")
        key = self._tune_key() + (":a" if analysis else "") if cache else None
        stored = _read_cache(cache).get(key) if cache else None
        if stored is not None:
            self.set_tuning(stored)
        else:
            self.makegrid(cilms, out=grids)
            self.expand(grids, out=outs)
            stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            with torch.cuda.device(self.device):
                st = _lib.lib().shtg_batch_tune(self._plan, int(reps), 1 if analysis else 0, stream)
            if st != 0:
                raise SHToolsError(_status_message(st))
            if cache:
                _write_cache(cache, key, self.tuning())
        self.makegrid(cilms, out=grids)
        self.expand(grids, out=outs)
        return self.tuning()

    def _tune_key(self):
        import hashlib
        arch = self.torch.cuda.get_device_properties(self.device).gcnArchName.split(":")[0]
        # grid codes 1 and 2 are the DH samplings
        sig = ",".join(f"{L}/{g}" for L, g in zip(self.lmax, self.grid))
        return arch + ":" + hashlib.sha1(sig.encode()).hexdigest()[:16]

    def tuning(self):
        """The plan's launch choices: {"fft_threads": [(synthesis, analysis)
        threads per ring-FFT group], "synth_kmax", "anal_kmax": ring pairs per
        lane of the Legendre waves, "mfma_min": bucket size sent to the MFMA
        kernel (0: none)}."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        st = _lib.lib().shtg_batch_get_tuning(self._plan, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        if st != 0:
            raise SHToolsError(_status_message(st))
        return {"fft_threads": self.shapes(), "synth_kmax": a.value, "anal_kmax": b.value, "mfma_min": c.value}

    def set_tuning(self, t):
        """Apply a choice returned by tuning() (for instance one stored by an
        earlier run)."""
        lib = _lib.lib()
        st = lib.shtg_batch_set_tuning(self._plan, int(t["synth_kmax"]), int(t["anal_kmax"]), int(t["mfma_min"]))
        if st == 0 and t.get("fft_threads"):
            n = len(t["fft_threads"])
            ts = (ctypes.c_int * n)(*[int(x[0]) for x in t["fft_threads"]])
            ta = (ctypes.c_int * n)(*[int(x[1]) for x in t["fft_threads"]])
            st = lib.shtg_batch_set_shapes(self._plan, ts, ta, n)
        if st != 0:
            raise SHToolsError(_status_message(st))

    def set_mfma(self, min_bucket):
        """Compute the synthesis Legendre sums of same-degree buckets of at
        least ``min_bucket`` fields as FP64 MFMA products (0: off)."""
        stream = ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)
        st = _lib.lib().shtg_batch_set_mfma(self._plan, int(min_bucket), stream)
        if st != 0:
            raise SHToolsError(_status_message(st))

    def field_groups(self):
        """The ring-FFT group of each field (fields of one group share a
        launch and an LDS size class; the highest number is the group whose
        rings run on the global workspace, if any)."""
        arr = (ctypes.c_int * len(self.lmax))()
        ng = _lib.lib().shtg_batch_field_groups(self._plan, arr, len(self.lmax))
        if ng < 0:
            raise SHToolsError(_status_message(5))
        return [arr[i] for i in range(len(self.lmax))]

    def shapes(self):
        """(synthesis, analysis) threads per workgroup of each ring-FFT group."""
        n = 64
        ts = (ctypes.c_int * n)()
        ta = (ctypes.c_int * n)()
        ng = _lib.lib().shtg_batch_get_shapes(self._plan, ts, ta, n)
        return [(ts[i], ta[i]) for i in range(min(ng, n))]


class DHBatch(SHBatch):
    """Plan for DH fields of degrees ``lmax``; ``sampling`` is one value for
    every field or a sequence with one per field."""

    def __init__(self, lmax, sampling=2, norm=1, csphase=1, device=None):
        n = len(list(lmax))
        samp = list(sampling) if isinstance(sampling, (list, tuple)) else [sampling] * n
        if len(samp) != n:
            raise SHToolsError(_status_message(1))
        if any(int(x) not in (1, 2) for x in samp):
            raise SHToolsError(_status_message(2))
        self.sampling = [int(x) for x in samp]
        super().__init__(lmax, grid=self.sampling, norm=norm, csphase=csphase, device=device)


class GLQBatch(SHBatch):
    """Plan for Gauss-Legendre fields of degrees ``lmax`` (grids of
    lmax+1 x 2lmax+1 on the nodes of SHGLQ)."""

    def __init__(self, lmax, norm=1, csphase=1, device=None):
        super().__init__(lmax, grid="GLQ", norm=norm, csphase=csphase, device=device)
