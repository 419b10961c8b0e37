#!/usr/bin/env python3
"""Benchmark: two-site sigma vector H_eff|psi> on one MI355X (FP64 block GEMMs on MFMA).

Workload (the headline configuration): open Hubbard chain L=32, t=1, U=4, half filling, 2Sz=0, bond dimension M=512,
two-site centre at the middle bond, compact nearest-neighbour MPO.  Sector dimensions follow a converged chain's
spectrum shape and every quantum-number-allowed block holds random N(0,1)/sqrt(m) entries (synthetic data, see
mpsxx_amd/workloads.py).  One step = one sigma-vector application with inputs resident in HBM.

Prints one JSON line.  `value` = algorithmic TFLOP/s of the sigma vector (reference contraction order, flops counted
by the planner), `roofline` = the dominant kernel against its bound, `cpu_baseline` = the oracle (FP64 CPU
restatement of the reference algorithm) timed on the host on the same inputs.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mpsxx_amd as mx  # noqa: E402
from mpsxx_amd.workloads import hubbard_twosite  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--M", type=int, default=512)
    ap.add_argument("--U", type=float, default=4.0)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X: the sigma-vector path runs on a single GPU in this build "
              "(multi-GPU sharding over operator groups is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    prec = mx.FP64 if args.precision == "fp64" else mx.FP32
    lm, rm, le, re_, psi = hubbard_twosite(args.L, args.M, args.U)
    dl = [mx.DTensor(x) for x in (lm, rm, le, re_, psi)]
    lib = mx.lib()

    flops = ctypes.c_int64()
    nbytes = ctypes.c_int64()
    nsteps = ctypes.c_int32()
    mx._check(lib.mpsxx_sigma_twosite_stats(*[d.h for d in dl], ctypes.c_int(prec), ctypes.byref(flops),
                                            ctypes.byref(nbytes), ctypes.byref(nsteps)))

    # the first application shapes the output (plan built, storage allocated); warmup then runs the accumulate plan
    out = mx.DTensor()
    mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
    for _ in range(max(args.warmup, 1)):
        mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
    mx.synchronize()
    # timed region: K sigma applications (output cleared as Davidson does before each multH)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
    mx.synchronize()
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    tflops = flops.value / sec / 1e12

    # per-step device times (HIP events on the launch stream)
    ms = (ctypes.c_double * 16)()
    sf = (ctypes.c_int64 * 16)()
    sk = (ctypes.c_int32 * 16)()
    ns = ctypes.c_int32()
    mx._check(lib.mpsxx_sigma_twosite_profile(*[d.h for d in dl], ctypes.c_int(prec), ctypes.c_int(50),
                                              ctypes.c_int32(16), ms, sf, sk, ctypes.byref(ns), None))
    steps = [dict(kind=["gemm", "axpy", "copy"][sk[i]], ms=ms[i], flops=int(sf[i])) for i in range(ns.value)]
    dom = max(range(ns.value), key=lambda i: ms[i])
    dk = steps[dom]
    peak = FP64_PEAK_TFLOPS if prec == mx.FP64 else FP32_PEAK_TFLOPS
    if dk["kind"] == "gemm":
        achieved = dk["flops"] / (dk["ms"] * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak, traffic=None,
                    kernel="gemm_grouped_kernel", step=dom)
    else:
        achieved = dk["flops"] * 4 / (dk["ms"] * 1e-3) / 1e9  # 2 flops per 8-byte read+... see DESIGN
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=None, kernel="axpy_grouped_kernel", step=dom)

    # parity of this very run against the oracle + CPU baseline (oracle, 1 thread)
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        oracle.set_threads(1)
        got = out.to_qst()
        ref = None
        reps = 0
        t0 = time.perf_counter()
        while True:
            ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
            reps += 1
            if time.perf_counter() - t0 > args.cpu_seconds or reps >= 200:
                break
        tc = (time.perf_counter() - t0) / reps
        cpu = dict(value=flops.value / tc / 1e12, unit="TFLOP/s", cores=1, kind="port",
                   sample="%d sigma applications of the same workload (%.3f s each)" % (reps, tc),
                   sigma_per_s=1.0 / tc)
        # the timed output accumulated warmup+steps sigmas onto a cleared tensor: compare one application
        one = mx.DTensor()
        mx.compute_sigma_vector_twosite(*dl, sgv0=one, precision=prec)
        g1 = one.to_qst()
        scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values())
        parity = dict(same_blocks=g1.block_set() == ref.block_set(), max_abs_diff=g1.max_abs_diff(ref) / scale)
        del got

    res = {
        "metric": "H_psi TFLOP/s (two-site sigma vector, FP64 MFMA)",
        "value": tflops,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == mx.FP64 else "f32",
        "data": "synthetic",
        "config": {"workload": "hubbard_sigma_twosite", "L": args.L, "M": args.M, "U": args.U, "t": 1.0,
                   "N": args.L, "twoSz": 0, "mpo": "compact nearest-neighbour (D_w=6)",
                   "psi_blocks": len(psi.blocks), "psi_elems": psi.nelem(), "precision": args.precision},
        "sigma_per_s": 1.0 / sec,
        "flops_per_sigma": flops.value,
        "bytes_per_sigma": nbytes.value,
        "steps_detail": steps,
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
