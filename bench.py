#!/usr/bin/env python
"""Benchmark: batched DH synthesis + analysis, fields/s (FP64).

Workload: 512 independent real 4pi-normalised fields with lmax ~ U{63..1023}
(seeded) and spectral slopes ~ U[-4, -2], on Driscoll-Healy sampling-2 grids
of (2L+2) x (4L+4).  One step = MakeGridDH followed by SHExpandDH of every
field, inputs resident in HBM.  Prints one JSON line.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0
# profiling ids of the library's stages and the batched kernels behind them
KERNEL_NAMES = {0: "k_pack_synth", 1: "k_synth_leg", 2: "k_synth_fft", 3: "k_anal_fft", 4: "k_anal_leg",
                5: "k_unpack_anal"}
BATCH_KERNELS = {"k_pack_synth": "kb_pack", "k_synth_leg": "kb_synth_leg", "k_synth_fft": "kb_synth_fft",
                 "k_anal_fft": "kb_anal_fft", "k_anal_leg": "kb_anal_leg", "k_unpack_anal": "kb_unpack"}
CALIB_BYTES = 1 << 30     # plain device-to-device copy used to scale the HBM counters
PROF_STEPS = 3            # eager steps averaged for the per-stage HIP-event durations


def stage_model(L, sampling=2):
    """Work of each stage for one DH field of degree L, one direction:
    (algorithmic FP64 flops, algorithmic HBM bytes, intermediate bytes).

    The algorithmic figures split the transform's work model (per direction
    F = 8 P (L+1)(L+2)/2 + 2.5 nlat nlon log2 nlon flops and B = 8 (2 (L+1)^2
    + nlat nlon) bytes: coefficients once, grid once) over the stages that do
    it: the Legendre stages the recursion and accumulation flops, the ring
    FFTs the FFT flops and the grid, pack/unpack the coefficient array.  The
    intermediate bytes are what the split into stages adds on top (the packed
    coefficient triangle and the order-major ring sums written by one stage
    and read by the next); they are reported as overhead, not credited."""
    from shtools_amd.synthetic import dh_flops
    n = 2 * L + 2
    nlon = sampling * n
    P = n // 2 + 1
    leg, fft = dh_flops(L, sampling)
    tri = (L + 1) * (L + 2) // 2
    coef_b = 8.0 * 2 * (L + 1) ** 2
    grid_b = 8.0 * n * nlon
    ring_sums = 32.0 * P * (L + 1)
    packed = 16.0 * tri
    return {"k_pack_synth": (0.0, coef_b, packed), "k_synth_leg": (leg, 0.0, packed + ring_sums),
            "k_synth_fft": (fft, grid_b, ring_sums), "k_anal_fft": (fft, grid_b, ring_sums),
            "k_anal_leg": (leg, 0.0, packed + ring_sums), "k_unpack_anal": (0.0, coef_b, packed)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nfields", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--loop", action="store_true", help="time the per-field loop of the plain kernels instead")
    ap.add_argument("--no-loop", action="store_true", help="skip the per-field loop comparison")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM traffic per kernel")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of HIP-graph replays")
    ap.add_argument("--no-tune", action="store_true", help="skip the ring-FFT launch-shape autotuner")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the timings of the other workloads (GLQ 1023, DH1 2047, 16 fits, high degree)")
    ap.add_argument("--configs-only", action="store_true",
                    help="time only the other workloads (for profiling them); prints their JSON")
    ap.add_argument("--counter-pass", action="store_true",
                    help="(internal) one untimed step then a calibration copy, for a rocprofv3 --pmc run")
    return ap.parse_args()


def make_workload(nfields, dev):
    import torch
    from shtools_amd.synthetic import mixed_lmax_workload, power_law_cilm
    spec, rng = mixed_lmax_workload(nfields=nfields)
    host_cilm = [power_law_cilm(rng, L, s) for L, s in spec]
    cilm = [torch.from_numpy(c).to(dev) for c in host_cilm]
    return [L for L, _ in spec], host_cilm, cilm


def counter_pass(args):
    """Run under rocprofv3 --pmc: one batched step, then a plain copy of
    CALIB_BYTES whose counter readings scale the others to bytes."""
    import torch
    from shtools_amd.batch import DHBatch
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    Ls, _, cilm = make_workload(args.nfields, dev)
    batch = DHBatch(Ls, sampling=2, norm=1, csphase=1, device=dev)
    grids = batch.makegrid(cilm)
    batch.expand(grids)
    src = torch.ones(CALIB_BYTES // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def run_counter(counter, nfields, timeout=240):
    """One rocprofv3 --pmc pass of the counter pass; returns {kernel name:
    summed KiB} and the reading of the last dispatch (the calibration copy)."""
    res = run_counters([counter], nfields, timeout)
    return res[counter]


def run_counters(counters, nfields, timeout=240):
    """One rocprofv3 --pmc pass over the counter pass collecting several
    counters (of blocks that fit one pass); returns {counter: ({kernel name:
    summed value}, value of the last dispatch)}."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="shtg_pmc_")
    names = " ".join(counters)
    cmd = [exe, "--pmc"] + list(counters) + ["--output-format", "csv", "-d", out, "-o", "run", "--",
                                             sys.executable, os.path.abspath(__file__), "--counter-pass",
                                             "--nfields", str(nfields)]
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {names} exited with {r.returncode}: {r.stderr[-400:]}")
    path = None
    for dp, _, fs in os.walk(out):
        for f in fs:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(dp, f)
    if path is None:
        raise RuntimeError(f"rocprofv3 --pmc {names} wrote no counter_collection.csv")
    res = {c: ({}, (-1, 0.0)) for c in counters}
    with open(path) as fh:
        for row in csv.DictReader(fh):
            c = row["Counter_Name"]
            if c not in res:
                continue
            per_kernel, last = res[c]
            v = float(row["Counter_Value"])
            name = row["Kernel_Name"]
            per_kernel[name] = per_kernel.get(name, 0.0) + v
            d = int(row["Dispatch_Id"])
            if d > last[0]:
                res[c] = (per_kernel, (d, v))
    shutil.rmtree(out, ignore_errors=True)
    return {c: (pk, last[1]) for c, (pk, last) in res.items()}


# FP64 VALU instruction counters (per wave instruction; 64 lanes each, an FMA
# counting 2 flops): the counter-measured FP64 work of each kernel
FP64_COUNTERS = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64")


def measure_fp64(nfields):
    """FP64 flops per step of each batched stage from the VALU F64
    instruction counters (64 lanes per wave instruction, 2 flops per FMA;
    inactive lanes count as well, so this is an upper bound); None if the
    profiler is unavailable."""
    try:
        res = run_counters(FP64_COUNTERS, nfields)
    except Exception as e:  # reported in the JSON line, never fatal
        return None, str(e)
    out = {}
    for stage, kname in BATCH_KERNELS.items():
        tot = {}
        for c in FP64_COUNTERS:
            per_kernel = res[c][0]
            tot[c] = sum(v for k, v in per_kernel.items() if k.startswith(f"shtg::{kname}") or
                         k.startswith(f"void shtg::{kname}"))
        out[stage] = 64.0 * (2.0 * tot["SQ_INSTS_VALU_FMA_F64"] + tot["SQ_INSTS_VALU_MUL_F64"] +
                             tot["SQ_INSTS_VALU_ADD_F64"])
    return out, None


def measure_traffic(nfields):
    """HBM bytes per step of each batched kernel from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by the reading of a plain copy
    of CALIB_BYTES; None if the profiler is unavailable."""
    res = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            per_kernel, calib = run_counter(counter, nfields)
            if calib <= 0:
                raise RuntimeError(f"{counter}: calibration copy read {calib}")
            scale = CALIB_BYTES / calib
            res[counter] = {"scale_bytes_per_unit": scale, "calib_kib": calib, "kernels": {}}
            for stage, kname in BATCH_KERNELS.items():
                tot = sum(v for k, v in per_kernel.items() if k.startswith(f"shtg::{kname}") or
                          k.startswith(f"void shtg::{kname}"))
                res[counter]["kernels"][stage] = tot * scale
    except Exception as e:  # reported in the JSON line, never fatal
        return None, str(e)
    return res, None


def _timed(fn, reps):
    """Median wall time (ms) of reps calls of fn, each bracketed by device
    synchronisations, after one untimed call."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    ts.sort()
    return ts[len(ts) // 2]


def other_configs(dev, reps=5):
    """The other workloads of the transform path, each timed on the GPU with
    inputs resident (median of reps), priced by the same work model, with an
    accuracy check of its own:
      glq_L1023: MakeGridGLQ + SHExpandGLQ, lmax 1023, orthonormal, 1024 x 2047;
      dh1_L2047: MakeGridDH + SHExpandDH, lmax 2047, sampling 1, orthonormal,
                 4096 x 4096;
      lsq_16fits_L511: one gradient of the 16-fit least-squares loss (DH2
                 1024 x 2048, batched synthesis + batched adjoint);
      high_degree_8fields: 8 fields lmax ~ U{2047..2799}, DH2, upward-
                 continuation filter, batched synthesis + analysis."""
    import numpy as np
    import torch
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    from shtools_amd.synthetic import (dh_flops, high_degree_workload, lsq_fit_workload, power_law_cilm,
                                       transform_flops, upward_continuation)
    out = {}

    def rec(name, ms, flops, extra):
        out[name] = dict(ms=ms, model_tflops=flops / (ms * 1e-3) / 1e12,
                         frac_fp64_peak=flops / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, **extra)
        print(f"[bench] {name}: {ms:.2f} ms, {out[name]['model_tflops']:.1f} model TFLOP/s", file=sys.stderr,
              flush=True)

    rng = np.random.default_rng(11)
    # Gauss-Legendre, lmax 1023, orthonormal
    L = 1023
    zero, w = B.SHGLQ(L)
    c = torch.from_numpy(power_law_cilm(rng, L, -2.0)).to(dev)
    state = {}

    def glq():
        state["g"] = B.MakeGridGLQ(c, zero, norm=4)
        state["c"] = B.SHExpandGLQ(state["g"], w, zero, norm=4)
    ms = _timed(glq, reps)
    err = float(torch.linalg.norm(state["c"] - c) / torch.linalg.norm(c))
    leg, fft = transform_flops(L, L + 1, 2 * L + 1)
    rec("glq_L1023", ms, 2 * (leg + fft), {"grid": [L + 1, 2 * L + 1], "roundtrip_err": err})
    # Driscoll-Healy sampling 1, lmax 2047, orthonormal
    L = 2047
    c = torch.from_numpy(power_law_cilm(rng, L, -2.0)).to(dev)

    def dh1():
        state["g"] = B.MakeGridDH(c, norm=4, sampling=1)
        state["c"] = B.SHExpandDH(state["g"], norm=4, sampling=1)
    ms = _timed(dh1, reps)
    err = float(torch.linalg.norm(state["c"] - c) / torch.linalg.norm(c))
    leg, fft = dh_flops(L, 1)
    rec("dh1_L2047", ms, 2 * (leg + fft), {"grid": [2 * L + 2, 2 * L + 2], "roundtrip_err": err})
    del state["g"], state["c"], c
    # 16 least-squares fits at lmax 511: gradient = S^T (2 w (S c - d)) + 2e-6 l(l+1) c
    L, nfits = 511, 16
    batch = DHBatch([L] * nfits, sampling=2, norm=1, csphase=1, device=dev)
    truth, data, wlat = lsq_fit_workload(batch, seed=4)
    ll1 = torch.tensor([l * (l + 1.0) for l in range(L + 1)], dtype=torch.float64, device=dev)[None, :, None]
    coef = [0.5 * t for t in truth]
    grids = [torch.empty(batch.grid_shape(i), dtype=torch.float64, device=dev) for i in range(nfits)]
    grads = [torch.empty_like(t) for t in truth]
    w2 = 2.0 * wlat[:, None]
    reg = [2e-6 * ll1 * torch.ones_like(t) for t in truth]

    def lsq_grad():
        batch.makegrid(coef, out=grids)
        torch._foreach_sub_(grids, data)          # residuals, in place of the grids
        torch._foreach_mul_(grids, [w2] * nfits)  # 2 w (S c - d)
        batch.makegrid_adjoint(grids, out=grads)
        torch._foreach_addcmul_(grads, reg, coef)  # + 2e-6 l(l+1) c
    ms = _timed(lsq_grad, reps)
    # the same gradient through torch.autograd (whose backward is the adjoint)
    from shtools_amd import autograd as AG
    cg = [c.clone().requires_grad_(True) for c in coef]
    gg = AG.batch_make_grid_dh(batch, cg)
    loss = sum((wlat[:, None] * (g - d) ** 2).sum() + 1e-6 * (ll1 * c * c).sum() for g, d, c in zip(gg, data, cg))
    ref = torch.autograd.grad(loss, cg)
    gerr = max(float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in zip(grads, ref))
    leg, fft = dh_flops(L, 2)
    rec("lsq_16fits_L511", ms, nfits * 2 * (leg + fft),
        {"gradients_per_s": nfits / (ms * 1e-3), "gradient_vs_autograd_err": gerr,
         "note": "one gradient of all 16 fits: batched synthesis, weighted residual, batched adjoint "
                 "(analysis kernels)"})
    del batch, truth, data, coef, grids, grads, cg, gg, ref
    # high-degree batch: 8 fields lmax ~ U{2047..2799}, filtered synthesis + analysis
    Ls, cs = high_degree_workload()
    batch = DHBatch(Ls, sampling=2, norm=1, csphase=1, device=dev)
    filt = [torch.from_numpy(upward_continuation(L)).to(dev)[None, :, None] for L in Ls]
    cin = [torch.from_numpy(x).to(dev) * f for x, f in zip(cs, filt)]
    grids = [torch.empty(batch.grid_shape(i), dtype=torch.float64, device=dev) for i in range(len(Ls))]
    back = [torch.empty_like(x) for x in cin]

    def high():
        batch.makegrid(cin, out=grids)
        batch.expand(grids, out=back)
    ms = _timed(high, max(3, reps // 2))
    err = max(float(torch.linalg.norm(b - x) / torch.linalg.norm(x)) for b, x in zip(back, cin))
    rec("high_degree_8fields", ms, sum(2 * sum(dh_flops(L, 2)) for L in Ls),
        {"lmax": Ls, "fields_per_s": len(Ls) / (ms * 1e-3), "roundtrip_err": err})
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.counter_pass:
        counter_pass(args)
        return
    if args.configs_only:
        import torch
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        print(json.dumps({"other_configs": other_configs(dev)}))
        return

    # counter passes first, in child processes, before this one touches the GPU
    traffic, traffic_err = None, "skipped (--no-traffic)"
    fp64 = None
    if not args.no_traffic and not args.loop:
        t0 = time.perf_counter()
        traffic, traffic_err = measure_traffic(args.nfields)
        fp64, fp64_err = measure_fp64(args.nfields)
        print(f"[bench] counter passes: {time.perf_counter() - t0:.1f} s"
              + (f" ({traffic_err})" if traffic_err else "") + (f" ({fp64_err})" if fp64_err else ""),
              file=sys.stderr, flush=True)

    import numpy as np
    import torch

    from shtools_amd import _lib
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    from shtools_amd.synthetic import dh_flops

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    Ls, host_cilm, cilm = make_workload(args.nfields, dev)
    print(f"[bench] {len(Ls)} fields, lmax {min(Ls)}..{max(Ls)}", file=sys.stderr, flush=True)

    batch = DHBatch(Ls, sampling=2, norm=1, csphase=1, device=dev)
    grids = [torch.empty(batch.grid_shape(i), dtype=torch.float64, device=dev) for i in range(len(Ls))]
    outs = [torch.empty(batch.cilm_shape(i), dtype=torch.float64, device=dev) for i in range(len(Ls))]

    def loop_step():
        res = []
        for c in cilm:
            g = B.MakeGridDH(c, sampling=2)
            res.append(B.SHExpandDH(g, sampling=2))
        return res

    def batch_step():
        batch.makegrid(cilm, out=grids)
        batch.expand(grids, out=outs)
        return outs

    step = loop_step if args.loop else batch_step
    shapes = None
    if not args.loop and not args.no_tune:
        t0 = time.perf_counter()
        shapes = batch.tune(cilm, grids, outs)
        print(f"[bench] tuned launches in {time.perf_counter() - t0:.1f} s: {shapes}",
              file=sys.stderr, flush=True)

    # the per-field loop of the plain kernels, for the batched-vs-loop ratio
    # (best of three passes after a warm-up pass: the loop is partly bound by
    # the host's calls, whose time varies from pass to pass)
    loop_rate = None
    if not args.loop and not args.no_loop:
        loop_step()
        torch.cuda.synchronize()
        for _ in range(3):
            l0 = time.perf_counter()
            loop_step()
            torch.cuda.synchronize()
            loop_rate = max(loop_rate or 0.0, len(cilm) / (time.perf_counter() - l0))
        print(f"[bench] per-field loop: {loop_rate:.1f} fields/s (best of 3 passes)", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timed = step
    if not args.loop and not args.no_graph:
        graph = batch.capture(cilm, grids, outs)

        def timed():
            graph.replay()
            return outs
    for _ in range(args.warmup):
        timed()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = timed()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = len(cilm) * args.steps / (t1 - t0)
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step", file=sys.stderr, flush=True)
    # parity sanity on the last step's outputs (round trip, norm-wise)
    rt = max(float(torch.linalg.norm(o - c) / torch.linalg.norm(c)) for o, c in zip(res, cilm))
    print(f"[bench] worst round-trip err {rt:.2e}", file=sys.stderr, flush=True)

    # per-stage durations with HIP events recorded on the launch stream
    # around every launch, averaged over PROF_STEPS eager steps
    lib = _lib.lib()
    nk = 16
    ms = (ctypes.c_double * nk)()
    cnt = (ctypes.c_longlong * nk)()
    lib.shtg_profile_enable(1)
    for _ in range(PROF_STEPS):
        step()
    torch.cuda.synchronize()
    lib.shtg_profile_collect(ms, cnt, nk)
    lib.shtg_profile_enable(0)
    kern = {KERNEL_NAMES.get(k, str(k)): ms[k] / PROF_STEPS for k in range(nk) if cnt[k]}
    launches = {KERNEL_NAMES.get(k, str(k)): cnt[k] / PROF_STEPS for k in range(nk) if cnt[k]}
    print("[bench] stage ms: " + ", ".join(f"{k}={v:.2f}" for k, v in kern.items()), file=sys.stderr, flush=True)

    # algorithmic work per stage over the batch, and the roofline of each
    work = {}
    for L in Ls:
        for k, fbi in stage_model(L).items():
            w = work.setdefault(k, [0.0, 0.0, 0.0])
            for j in range(3):
                w[j] += fbi[j]
    stages = {}
    for k, t in kern.items():
        f, b, inter = work.get(k, (0.0, 0.0, 0.0))
        bound = "fp64" if k in ("k_synth_leg", "k_anal_leg") else "hbm"
        if bound == "fp64":
            ach, peak, unit = f / (t * 1e-3) / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
        else:
            ach, peak, unit = b / (t * 1e-3) / 1e9, PEAK_HBM_GBS, "GB/s"
        st = {"ms": t, "launches": launches[k], "bound": bound, "achieved": ach, "peak": peak, "unit": unit,
              "frac": ach / peak, "algo_flops": f, "algo_bytes": b, "intermediate_bytes": inter}
        if traffic:
            st["hbm_bytes"] = (traffic["FETCH_SIZE"]["kernels"][k] + traffic["WRITE_SIZE"]["kernels"][k])
            st["hbm_gbs_counters"] = st["hbm_bytes"] / (t * 1e-3) / 1e9
            if b > 0:
                st["hbm_over_algo"] = st["hbm_bytes"] / b
        if fp64:
            st["fp64_flops_counters"] = fp64[k]
            st["fp64_tflops_counters"] = fp64[k] / (t * 1e-3) / 1e12
        stages[k] = st
    dom = max(stages, key=lambda k: stages[k]["ms"])
    d = stages[dom]
    roofline = {"bound": d["bound"], "achieved": d["achieved"], "peak": d["peak"], "unit": d["unit"],
                "frac": d["frac"], "traffic": d.get("hbm_bytes"), "kernel": BATCH_KERNELS.get(dom, dom),
                "ms": d["ms"], "algo_bytes": d["algo_bytes"], "algo_flops": d["algo_flops"],
                "intermediate_bytes": d["intermediate_bytes"],
                "launch": "one step's dispatches of the stage (FFT stages: one per LDS class); achieved = "
                          "algorithmic bytes (grid once) or flops of those dispatches / their summed duration"}
    if traffic_err:
        roofline["traffic_note"] = traffic_err
    if traffic:
        roofline["traffic_counters"] = {
            c: {"kib_of_1GiB_copy": traffic[c]["calib_kib"], "bytes_per_kib": traffic[c]["scale_bytes_per_unit"],
                "bytes": traffic[c]["kernels"][dom]} for c in ("FETCH_SIZE", "WRITE_SIZE")}
    leg_flops = sum(dh_flops(L)[0] for L in Ls)
    fft_flops = sum(dh_flops(L)[1] for L in Ls)
    total_flops = 2 * (leg_flops + fft_flops)

    configs = None
    if not args.loop and not args.no_configs:
        configs = other_configs(dev)

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        nproc = os.cpu_count() or 1
        try:
            naff = len(os.sched_getaffinity(0))
        except AttributeError:
            naff = nproc
        # every core of this process's affinity mask, unless the host
        # limits the process's thread count (OMP_NUM_THREADS, the CPU share
        # a shared GPU host gives one GPU's jobs)
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        nthreads = min(naff, omp) if omp > 0 else naff
        budget = args.cpu_seconds * 1.5e9 * nthreads   # ~1.5 GFLOP/s per core for the Fortran-style loops
        idx, acc = [], 0.0
        for i, L in enumerate(Ls):
            if acc >= budget:
                break
            idx.append(i)
            acc += 2 * sum(dh_flops(L))
        sample = [host_cilm[i] for i in idx]
        c0 = time.perf_counter()
        gref, outc = oracle.roundtrip_dh_batch(sample, norm=1, sampling=2, csphase=1, nthreads=nthreads,
                                               return_grids=True)
        c1 = time.perf_counter()
        cerr = max(float(np.linalg.norm(o - c) / np.linalg.norm(c)) for o, c in zip(outc, sample))
        # the oracle's grids and coefficients of these fields against the GPU
        # batched outputs of the timed step (same inputs)
        if not args.loop:
            g_errs = [float(np.abs(grids[i].cpu().numpy() - g).max() / np.abs(g).max())
                      for i, g in zip(idx, gref)]
            c_errs = [float(np.linalg.norm(outs[i].cpu().numpy() - c) / np.linalg.norm(c))
                      for i, c in zip(idx, outc)]
            parity = {"grid_err_max": max(g_errs), "coef_err_max": max(c_errs), "fields": len(idx),
                      "measure": "per field: grids max|gpu-oracle|/max|oracle|, coefficients "
                                 "||gpu-oracle||/||oracle||; the oracle is the C restatement of the Fortran "
                                 "MakeGridDH/SHExpandDH, run on the CPU-baseline sample"}
            print(f"[bench] parity vs oracle on {len(idx)} fields: grid {parity['grid_err_max']:.2e}, "
                  f"coef {parity['coef_err_max']:.2e}", file=sys.stderr, flush=True)
        cpu = {"value": len(sample) / (c1 - c0), "unit": "fields/s", "cores": nthreads, "kind": "port",
               "nproc": nproc, "affinity_cores": naff, "omp_num_threads": omp or None,
               "sample": f"first {len(sample)} of the {len(Ls)} fields in workload order, as many as ~"
                         f"{args.cpu_seconds:.0f} s of CPU work at 1.5 GFLOP/s per thread allow (lmax "
                         f"{min(Ls[i] for i in idx)}..{max(Ls[i] for i in idx)}), MakeGridDH+SHExpandDH per "
                         f"field, OpenMP over fields on {nthreads} threads, {c1 - c0:.1f} s wall, round-trip "
                         f"err {cerr:.1e}",
               "model_gflops": acc / (c1 - c0) / 1e9}
        print(f"[bench] cpu baseline {cpu['value']:.3f} fields/s on {nthreads} threads", file=sys.stderr, flush=True)
        # the same loop on one thread, over a quarter of the time budget
        idx1, acc1 = [], 0.0
        for i in idx:
            if idx1 and acc1 >= args.cpu_seconds / 4 * 1.5e9:
                break
            idx1.append(i)
            acc1 += 2 * sum(dh_flops(Ls[i]))
        c0 = time.perf_counter()
        oracle.roundtrip_dh_batch([host_cilm[i] for i in idx1], norm=1, sampling=2, csphase=1, nthreads=1)
        c1 = time.perf_counter()
        cpu["one_thread"] = {"value": len(idx1) / (c1 - c0), "unit": "fields/s", "cores": 1,
                             "model_gflops": acc1 / (c1 - c0) / 1e9,
                             "sample": f"first {len(idx1)} fields of the same sample, {c1 - c0:.1f} s"}
        print(f"[bench] cpu baseline {cpu['model_gflops']:.1f} model GFLOP/s on {nthreads} threads, "
              f"{cpu['one_thread']['model_gflops']:.2f} on one", file=sys.stderr, flush=True)

    line = {
        "metric": "batched DH synth+analysis fields/s, FP64, batched vs per-field, rel err <=1e-12",
        "value": value,
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(Ls)} DH2 fields, lmax~U{{63..1023}} seed 3, slope~U[-4,-2], norm 4pi, "
                               "synthesis+analysis per step, "
                               + ("per-field loop of the plain kernels" if args.loop else
                                  "batched launches" + ("" if args.no_graph else " replayed as one HIP graph")),
                   "fields": len(Ls), "sampling": 2, "norm": 1, "csphase": 1},
        "roofline": roofline,
        "model_tflops": total_flops * args.steps / (t1 - t0) / 1e12,
        "roundtrip_err": rt,
        "parity_err": max(parity["grid_err_max"], parity["coef_err_max"]) if parity else None,
        "parity": parity,
        "loop_fields_per_s": loop_rate,
        "batched_vs_loop": (value / loop_rate) if loop_rate else None,
        "stages": stages,
        "tuning": shapes,
        "other_configs": configs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
