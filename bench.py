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
KERNEL_NAMES = {0: "k_pack_synth", 1: "k_synth_leg", 2: "k_synth_fft", 3: "k_anal_fft", 4: "k_anal_leg",
                5: "k_unpack_anal"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nfields", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--loop", action="store_true", help="time the per-field loop of the plain kernels instead")
    ap.add_argument("--no-loop", action="store_true", help="skip the per-field loop comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from shtools_amd import _lib
    from shtools_amd import backend as B
    from shtools_amd.synthetic import dh_flops, mixed_lmax_workload, power_law_cilm

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    spec, rng = mixed_lmax_workload(nfields=args.nfields)
    host_cilm = [power_law_cilm(rng, L, s) for L, s in spec]
    cilm = [torch.from_numpy(c).to(dev) for c in host_cilm]
    Ls = [L for L, _ in spec]
    print(f"[bench] {len(Ls)} fields, lmax {min(Ls)}..{max(Ls)}", file=sys.stderr, flush=True)

    from shtools_amd.batch import DHBatch
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

    # the per-field loop of the plain kernels, for the batched-vs-loop ratio
    loop_rate = None
    if not args.loop and not args.no_loop:
        loop_step()
        torch.cuda.synchronize()
        l0 = time.perf_counter()
        loop_outs = loop_step()
        torch.cuda.synchronize()
        loop_rate = len(cilm) / (time.perf_counter() - l0)
        print(f"[bench] per-field loop: {loop_rate:.1f} fields/s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        outs = step()
    torch.cuda.synchronize()
    # parity sanity on the warm-up outputs (round trip, norm-wise)
    rt = max(float(torch.linalg.norm(o - c) / torch.linalg.norm(c)) for o, c in zip(outs, cilm))
    print(f"[bench] warmup done, worst round-trip err {rt:.2e}", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = len(cilm) * args.steps / (t1 - t0)
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step", file=sys.stderr, flush=True)

    # per-kernel durations with HIP events on the launch stream
    lib = _lib.lib()
    lib.shtg_profile_enable(1)
    step()
    torch.cuda.synchronize()
    nk = 16
    ms = (ctypes.c_double * nk)()
    cnt = (ctypes.c_longlong * nk)()
    lib.shtg_profile_collect(ms, cnt, nk)
    lib.shtg_profile_enable(0)
    kern = {KERNEL_NAMES.get(k, str(k)): (ms[k], cnt[k]) for k in range(nk) if cnt[k]}
    leg_flops = sum(dh_flops(L)[0] for L in Ls)
    fft_flops = sum(dh_flops(L)[1] for L in Ls)
    algo = {"k_synth_leg": leg_flops, "k_anal_leg": leg_flops, "k_synth_fft": fft_flops, "k_anal_fft": fft_flops}
    dom = max(kern, key=lambda k: kern[k][0])
    dom_ms, dom_n = kern[dom]
    achieved = algo.get(dom, 0.0) / (dom_ms * 1e-3) / 1e12 if dom in algo else None
    total_flops = 2 * (leg_flops + fft_flops)
    print("[bench] kernels (ms total, launches): " + ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in kern.items()),
          file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        import oracle
        try:
            nthreads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            nthreads = min(16, os.cpu_count() or 1)
        budget = args.cpu_seconds * 1.5e9 * nthreads   # ~1.5 GFLOP/s per core for the Fortran-style loops
        idx, acc = [], 0.0
        for i, L in enumerate(Ls):
            if acc >= budget:
                break
            idx.append(i)
            acc += 2 * sum(dh_flops(L))
        sample = [host_cilm[i] for i in idx]
        c0 = time.perf_counter()
        outc = oracle.roundtrip_dh_batch(sample, norm=1, sampling=2, csphase=1, nthreads=nthreads)
        c1 = time.perf_counter()
        cerr = max(float(np.linalg.norm(o - c) / np.linalg.norm(c)) for o, c in zip(outc, sample))
        cpu = {"value": len(sample) / (c1 - c0), "unit": "fields/s", "cores": nthreads, "kind": "port",
               "sample": f"first {len(sample)} of the {len(Ls)} fields (lmax {min(Ls[i] for i in idx)}.."
                         f"{max(Ls[i] for i in idx)}), MakeGridDH+SHExpandDH per field, OpenMP over fields, "
                         f"{c1 - c0:.1f} s wall, round-trip err {cerr:.1e}"}
        print(f"[bench] cpu baseline {cpu['value']:.3f} fields/s on {nthreads} threads", file=sys.stderr, flush=True)

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
                               + ("per-field loop of the plain kernels" if args.loop else "batched launches"),
                   "fields": len(Ls), "sampling": 2, "norm": 1, "csphase": 1},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": (achieved / PEAK_FP64_TFLOPS) if achieved else None,
                     "traffic": None},
        "model_tflops": total_flops * args.steps / (t1 - t0) / 1e12,
        "roundtrip_err": rt,
        "loop_fields_per_s": loop_rate,
        "batched_vs_loop": (value / loop_rate) if loop_rate else None,
        "kernels_ms": {k: v[0] for k, v in kern.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
