"""Seeded synthetic spectra and the work model of the transforms.

Coefficients have Var(C_lm) = Var(S_lm) = S(l) / (2l+1) for a power law
S(l) = (l+1)^slope (the random models of pyshtools' SHCoeffs.from_random
with a power-law spectrum); S_l0 = 0 and entries with m > l are zero.
"""
import math

import numpy as np


def power_law_cilm(rng, lmax, slope=-2.0):
    c = rng.standard_normal((2, lmax + 1, lmax + 1))
    l = np.arange(lmax + 1)
    c *= np.sqrt((l + 1.0) ** slope / (2 * l + 1.0))[None, :, None]
    iu = np.triu_indices(lmax + 1, 1)
    c[:, iu[0], iu[1]] = 0.0
    c[1, :, 0] = 0.0
    return c


def mixed_lmax_workload(nfields=512, lmin=63, lmax=1023, slope_range=(-4.0, -2.0), seed=3):
    """The batched uneven-degree workload: lmax ~ U{lmin..lmax} and a
    spectral slope ~ U[slope_range] per field (seeded)."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(lmin, lmax + 1, size=nfields)
    slopes = rng.uniform(slope_range[0], slope_range[1], size=nfields)
    return [(int(L), float(s)) for L, s in zip(Ls, slopes)], rng


def transform_flops(lmax, nlat, nlon):
    """Algorithmic FP64 work of one transform direction: 8 flops per
    (degree, order, ring pair) for the recursion and the even/odd
    accumulation plus 2.5 N log2 N per real ring FFT."""
    P = (nlat + 1) // 2
    leg = 8.0 * P * (lmax + 1) * (lmax + 2) / 2.0
    fft = 2.5 * nlat * nlon * math.log2(nlon) if nlon > 1 else 0.0
    return leg, fft


def dh_flops(lmax, sampling=2):
    n = 2 * lmax + 2
    return transform_flops(lmax, n, sampling * n)


def transform_bytes(lmax, nlat, nlon):
    """Algorithmic HBM bytes of one direction: coefficients once, grid once."""
    return 8.0 * (2 * (lmax + 1) ** 2 + nlat * nlon)


def lsq_fit_workload(batch, seed=4, noise=0.01):
    """The differentiated least-squares workload on a DHBatch plan: per field a
    truth with power spectrum (l+1)^-3 (Var C_lm = (l+1)^-3/(2l+1)), data =
    MakeGridDH(truth) + Gaussian noise of sigma = noise * grid RMS, and the
    per-ring DH weights w_lat = DHaj(n).  Returns (truths, data, w_lat) as
    device tensors (w_lat for the first field's grid)."""
    import torch
    from . import backend
    rng = np.random.default_rng(seed)
    dev = batch.device
    truth = [torch.from_numpy(power_law_cilm(rng, L, -3.0)).to(dev) for L in batch.lmax]
    grids = batch.makegrid(truth)
    data = []
    for g in grids:
        rms = float(torch.sqrt((g * g).mean()))
        data.append(g + noise * rms * torch.from_numpy(rng.standard_normal(tuple(g.shape))).to(dev))
    n = 2 * batch.lmax[0] + 2
    w = torch.from_numpy(backend.DHaj(n)).to(dev)
    return truth, data, w


def kaula_cilm(rng, lmax, amp=1e-5):
    """Kaula-like spectrum: per-coefficient RMS amp / l^2 for l >= 2, zero for
    degrees 0 and 1 (the geodesy convention); S_l0 = 0, zero for m > l."""
    c = rng.standard_normal((2, lmax + 1, lmax + 1))
    l = np.arange(lmax + 1, dtype=np.float64)
    rms = np.zeros(lmax + 1)
    rms[2:] = amp / l[2:] ** 2
    c *= rms[None, :, None]
    iu = np.triu_indices(lmax + 1, 1)
    c[:, iu[0], iu[1]] = 0.0
    c[1, :, 0] = 0.0
    return c


def upward_continuation(lmax, r_over_R=1.01):
    """Per-degree filter (l+1) (R/r)^(l+2) of the high-degree workload."""
    l = np.arange(lmax + 1, dtype=np.float64)
    return (l + 1.0) * (1.0 / r_over_R) ** (l + 2.0)


def high_degree_workload(nfields=8, lmin=2047, lmax=2799, seed=5):
    """Degrees ~ U{lmin..lmax} (seeded) and Kaula-spectrum coefficients."""
    rng = np.random.default_rng(seed)
    Ls = [int(L) for L in rng.integers(lmin, lmax + 1, size=nfields)]
    return Ls, [kaula_cilm(rng, L) for L in Ls]
