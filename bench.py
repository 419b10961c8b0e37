#!/usr/bin/env python3
"""Benchmark: two-site sigma vector H_eff|psi> on one MI355X (block GEMMs on MFMA).

Headline workload: open Hubbard chain L=32, t=1, U=4, half filling, 2Sz=0, bond dimension M=512, two-site centre at
the middle bond, compact nearest-neighbour MPO.  Sector dimensions follow a converged chain's spectrum shape and every
quantum-number-allowed block holds random N(0,1)/sqrt(m) entries (synthetic data, see mpsxx_amd/workloads.py).  One
step = one sigma-vector application with inputs resident in HBM.

Prints one JSON line.  `value` = algorithmic TFLOP/s of the sigma vector (reference contraction order, flops counted
by the planner); `roofline` = the kernel with the most device time per sigma (all its launches: algorithmic flops /
device time, HIP events on the launch stream) against its bound; `cpu_baseline` = the oracle (FP64 CPU restatement of
the reference algorithm) on the same inputs on one host core.  Beside it: the same sigma with FP32 block GEMMs, and a
batch of 64 independent chains (L in 16..64, U in {2,4,8}, M in {128,256,512}) applied sequentially in FP64 and as one
grouped launch per step (FP64 and FP32, hipGraph replay, autotuned tiles), and the merged-form sigma at the middle of
a 24-orbital quantum-chemistry chain at M=1024 (synthetic integrals from seed 1, QC MPO with D_w=1226) in FP64, FP32 and
split-BF16x3 with the operators resident; and the sweep half of the metric: device-resident two-site sweeps of the
L=32 chain over M 128x3, 256x3, 512x2 in FP64 and in mixed precision (FP32 sigma vectors + FP64 refinement), with
the energy difference between them; and merged-form sweeps of the 24-orbital chain over M 256, 512, 1024, 1024 in
FP64 and mixed precision.
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
from mpsxx_amd import workloads  # noqa: E402
from mpsxx_amd.batch import (ALL_INPUTS, WFN_ONLY, SigmaBatch, dominant_kernel, last_plan_profile,  # noqa: E402
                             sigma_twosite_profile)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the headline workload from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_sigma_twosite_hubbard_L32_M512.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
            return int(parts[2])
    return None


def sigma_stats(dl, prec):
    flops = ctypes.c_int64()
    nbytes = ctypes.c_int64()
    nsteps = ctypes.c_int32()
    mx._check(mx.lib().mpsxx_sigma_twosite_stats(*[d.h for d in dl], ctypes.c_int(prec), ctypes.byref(flops),
                                                 ctypes.byref(nbytes), ctypes.byref(nsteps)))
    return flops.value, nbytes.value


def time_sigma(dl, prec, steps, warmup):
    """Wall time per sigma application (inputs in HBM, output accumulating as in Davidson's multH)."""
    out = mx.DTensor()
    mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
    for _ in range(max(warmup, 1)):
        mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
    mx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
    mx.synchronize()
    return (time.perf_counter() - t0) / steps


def roofline_of(entries, prec):
    name, agg, allk = dominant_kernel(entries)
    per_kernel = {k: dict(ms=v["ms"], launches=v["launches"], flops=v["flops"], bytes=v["bytes"],
                          avg_launch_us=1e3 * v["ms"] / v["launches"]) for k, v in allk.items()}
    sec = agg["ms"] * 1e-3
    if name == "axpy_grouped_kernel":
        achieved = agg["bytes"] / sec / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS)
    else:
        peak = FP64_PEAK_TFLOPS if prec == mx.FP64 else FP32_PEAK_TFLOPS
        achieved = agg["flops"] / sec / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak)
    roof.update(traffic=None, kernel=name, launches_per_sigma=agg["launches"],
                avg_launch_us=1e3 * agg["ms"] / agg["launches"], algorithmic_flops=agg["flops"],
                algorithmic_bytes=agg["bytes"])
    return roof, per_kernel


def rel_diff(a, b):
    return a.diff_norm(b) / b.norm()


def qdot(a, b):
    """<a|b> over the common blocks of two host block-sparse arrays."""
    return float(sum(np.vdot(a.blocks[k], b.blocks[k]) for k in a.blocks if k in b.blocks))


def batched_leg(n, reps):
    specs = workloads.batch_specs(n, 4)
    probs = [workloads.hubbard_twosite(L=L, M=M, U=U, seed=k) for k, (L, U, M) in enumerate(specs)]
    dev = [tuple(mx.DTensor(x) for x in p) for p in probs]
    # plain FP64 path: one compute_sigma_vector call per instance
    outs = [mx.DTensor() for _ in dev]
    for d, o in zip(dev, outs):
        mx.compute_sigma_vector_twosite(*d, sgv0=o)
    mx.synchronize()
    ref = [o.to_qst() for o in outs]
    t0 = time.perf_counter()
    for _ in range(reps):
        for d, o in zip(dev, outs):
            mx.compute_sigma_vector_twosite(*d, sgv0=o)
    mx.synchronize()
    seq = (time.perf_counter() - t0) / reps
    res = dict(instances=n, sample_seed=4, specs=[list(s) for s in specs], seq_fp64_ms=seq * 1e3)
    for prec, name in ((mx.FP64, "fp64"), (mx.FP32, "fp32")):
        inst = [d + (mx.DTensor(),) for d in dev]
        b = SigmaBatch(inst, precision=prec)
        info = b.info()
        res["flops_per_batch"] = info["flops"]
        res[name + "_default_ms"] = b.time(reps, ALL_INPUTS, True)
        res[name + "_policy"] = b.tune(3)
        res[name + "_graph_ms"] = b.time(reps, ALL_INPUTS, True)
        if prec == mx.FP32:
            res[name + "_graph_wfn_ms"] = b.time(reps, WFN_ONLY, True)
        prof = b.profile(5)
        res[name + "_roofline"], _ = roofline_of(prof, prec)
        # accuracy of one application against the plain FP64 path
        for t in inst:
            t[5].clear()
        b.run(ALL_INPUTS, True)
        mx.synchronize()
        outs_p = [t[5].to_qst() for t in inst]
        res[name + "_max_rel_err"] = max(rel_diff(o, r) for o, r in zip(outs_p, ref))
        if prec == mx.FP32:
            # per-instance energy functional <psi|H psi>/<psi|psi> from this launch vs the plain FP64 path
            # (the environments are random, so E can be near zero: the error is also given on the operator scale
            # ||H psi|| / ||psi||, which bounds |E|)
            errs, errs_op = [], []
            for pr, o, r in zip(probs, outs_p, ref):
                psi_q = pr[4]
                nn = qdot(psi_q, psi_q)
                e32, e64 = qdot(psi_q, o) / nn, qdot(psi_q, r) / nn
                errs.append(abs(e32 - e64) / max(abs(e64), 1e-300))
                errs_op.append(abs(e32 - e64) / max(np.sqrt(qdot(r, r) / nn), 1e-300))
            res["fp32_energy_rel_err_max"] = max(errs)
            res["fp32_energy_rel_err_median"] = float(np.median(errs))
            res["fp32_energy_err_operator_scale_max"] = max(errs_op)
        b.free()
    fl = res["flops_per_batch"]
    for k in list(res):
        if k.endswith("_ms"):
            res[k[:-3] + "_tflops"] = fl / (res[k] * 1e-3) / 1e12
            res[k[:-3] + "_sigma_per_s"] = n / (res[k] * 1e-3)
    res["speedup_fp32_batched_vs_fp64_sequential"] = res["seq_fp64_ms"] / res["fp32_graph_wfn_ms"]
    res["speedup_fp64_batched_vs_fp64_sequential"] = res["seq_fp64_ms"] / res["fp64_graph_ms"]
    return res


def qc_leg(norb, M, reps):
    """Merged-form sigma (src/compute_sigma_vector.hpp:86-96) at the middle of a QC chain, FP64 / FP32 / split-BF16."""
    Lm, Rm, psim, info = workloads.qc_merged_problem(norb, M, 1)
    res = dict(info)
    ref = None
    for prec, name in ((mx.FP64, "fp64"), (mx.FP32, "fp32"), (mx.BF16X3, "bf16x3")):
        out = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=prec)  # converts every input once
        mx.synchronize()
        one = out.to_qst()
        if ref is None:
            ref = one
        ent, pf = last_plan_profile(3)
        roof, _ = roofline_of(ent, prec if prec != mx.BF16X3 else mx.FP32)
        # operators resident (FP32 copies kept): only the wavefunction is refreshed per call, as inside Davidson
        mask = 1 << 2
        t0 = time.perf_counter()
        for _ in range(reps):
            mx.compute_sigma_vector_merged_ex(Lm, Rm, psim, out, prec, mask)
        mx.synchronize()
        sec = (time.perf_counter() - t0) / reps
        res[name] = dict(ms=sec * 1e3, tflops=pf / sec / 1e12, rel_diff_vs_fp64=rel_diff(one, ref), roofline=roof)
        res["flops_per_sigma"] = pf
        del out
    res["speedup_fp32_vs_fp64"] = res["fp64"]["ms"] / res["fp32"]["ms"]
    res["speedup_bf16x3_vs_fp64"] = res["fp64"]["ms"] / res["bf16x3"]["ms"]
    return res


def sweep_leg(L, schedule, precisions):
    """Device-resident two-site sweeps of the open Hubbard chain (t=1, U=4, half filling, 2Sz=0) from the seeded
    random initial MPS over the M schedule; per precision: total and per-sweep seconds, final energy, and for the
    mixed-precision runs |E - E_fp64|."""
    from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
    from mpsxx_amd.mpo import hubbard_mpo
    mpos = hubbard_mpo(L, 1.0, 4.0)
    init = initial_mps(L, (L, 0), schedule[0], 0)
    out = {"L": L, "schedule": schedule, "tol": 1e-10}
    e64 = None
    for prec in precisions:
        name = {mx.FP64: "fp64", mx.FP32: "fp32_mixed", mx.BF16X3: "bf16x3_mixed"}[prec]
        print("sweep leg: %s" % name, file=sys.stderr, flush=True)
        dm = DeviceDMRG(mpos, init, schedule[0])
        secs, e, st = [], None, None
        t0 = time.perf_counter()
        for M in schedule:
            e, st = dm.sweep(M, 1e-10, precision=prec)
            secs.append(round(st["seconds"], 3))
        total = time.perf_counter() - t0
        r = {"seconds_total": total, "seconds_per_sweep": secs, "energy": e, "last_sweep_svd_s": st["svd_s"],
             "last_sweep_davidson_s": st["davidson_s"], "last_sweep_sigma_ms": st["sigma_ms"]}
        if prec == mx.FP64:
            e64 = e
        elif e64 is not None:
            r["abs_dE_vs_fp64"] = abs(e - e64)
        out[name] = r
    return out


def qc_sweep_leg(norb, schedule, precisions, max_final=6):
    """Merged two-site sweeps (src/make_sweep.cpp:112) of the QC chain with synthetic integrals from seed 1 over the
    M schedule, the last M repeated until the sweep energy changes by less than the tolerance (src/dmrg.cpp:44-78,
    at most max_final sweeps at that M): per precision the total and per-sweep seconds, the converged energy and
    |E - E_fp64|."""
    from mpsxx_amd import qc
    from mpsxx_amd.dmrg import SWEEP_MERGED, DeviceDMRG, initial_mps
    h, v = qc.synthetic_integrals(norb, seed=1)
    m = qc.QcMpo(h, v)
    mpos = m.devices()
    out = {"norb": norb, "schedule": schedule, "tol": 1e-10, "D_w_max": max(m.bond_dims())}
    e64 = None
    for prec in precisions:
        name = {mx.FP64: "fp64", mx.FP32: "fp32_mixed", mx.BF16X3: "bf16x3_mixed"}[prec]
        print("qc sweep leg: %s" % name, file=sys.stderr, flush=True)
        dm = DeviceDMRG(mpos, initial_mps(norb, (norb, 0), schedule[0], 0), schedule[0], form=SWEEP_MERGED)
        secs, e = [], None
        t0 = time.perf_counter()
        for M in schedule[:-1]:
            e, st = dm.sweep(M, 1e-10, precision=prec)
            secs.append(round(st["seconds"], 3))
        for k in range(max_final):
            e_prev = e
            e, st = dm.sweep(schedule[-1], 1e-10, precision=prec)
            secs.append(round(st["seconds"], 3))
            if k > 0 and abs(e - e_prev) < 1e-10:
                break
        r = {"seconds_total": time.perf_counter() - t0, "seconds_per_sweep": secs, "energy": e,
             "last_sweep_svd_s": st["svd_s"], "last_sweep_davidson_s": st["davidson_s"],
             "last_sweep_renorm_s": st["renorm_s"]}
        if prec == mx.FP64:
            e64 = e
        elif e64 is not None:
            r["abs_dE_vs_fp64"] = abs(e - e64)
        out[name] = r
        del dm
    return out


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
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched leg (0: skip it)")
    ap.add_argument("--qc-norb", type=int, default=24, help="orbitals of the quantum-chemistry leg (0: skip it)")
    ap.add_argument("--qc-M", type=int, default=1024)
    ap.add_argument("--sweeps", type=int, default=1, help="run the sweep leg (L=32, M 128x3, 256x3, 512x2; 0: skip)")
    ap.add_argument("--qc-sweeps", type=int, default=1,
                    help="run the QC merged-sweep leg (norb --qc-norb, M 256, 512, then 1024 until converged; 0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X: the sigma-vector path runs on a single GPU in this build "
              "(multi-GPU sharding over operator groups is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    prec = mx.FP64 if args.precision == "fp64" else mx.FP32
    lm, rm, le, re_, psi = workloads.hubbard_twosite(args.L, args.M, args.U)
    dl = [mx.DTensor(x) for x in (lm, rm, le, re_, psi)]
    print("headline sigma: L=%d M=%d %s" % (args.L, args.M, args.precision), file=sys.stderr, flush=True)
    flops, nbytes = sigma_stats(dl, prec)
    sec = time_sigma(dl, prec, args.steps, args.warmup)
    tflops = flops / sec / 1e12
    entries = sigma_twosite_profile(*dl, precision=prec, reps=50)
    roof, per_kernel = roofline_of(entries, prec)
    if prec == mx.FP64 and args.L == 32 and args.M == 512 and args.U == 4.0:
        # HBM bytes per launch from the rocprofv3 --pmc runs of this workload (profiles/), when committed
        roof["traffic"] = measured_traffic(roof["kernel"])
        if roof["traffic"] is not None:
            roof["traffic_source"] = "profiles/traffic_sigma_twosite_hubbard_L32_M512.md (rocprofv3 --pmc FETCH_SIZE, " \
                                     "WRITE_SIZE; copy-kernel calibration 2.0 and 1.0)"

    # the other precision on the same inputs
    other = mx.FP32 if prec == mx.FP64 else mx.FP64
    sec_o = time_sigma(dl, other, args.steps, args.warmup)
    one = mx.compute_sigma_vector_twosite(*dl, precision=prec).to_qst()
    one_o = mx.compute_sigma_vector_twosite(*dl, precision=other).to_qst()
    other_leg = dict(precision="fp32" if other == mx.FP32 else "fp64", ms_per_sigma=sec_o * 1e3,
                     tflops=flops / sec_o / 1e12, rel_diff_vs_headline=rel_diff(one_o, one))

    batched = None
    if args.batch > 0:
        print("batched leg: %d instances" % args.batch, file=sys.stderr, flush=True)
        batched = batched_leg(args.batch, 20)

    qcres = None
    if args.qc_norb > 0:
        print("quantum-chemistry leg: norb=%d M=%d" % (args.qc_norb, args.qc_M), file=sys.stderr, flush=True)
        qcres = qc_leg(args.qc_norb, args.qc_M, 5)

    sweeps = None
    if args.sweeps > 0:
        sweeps = sweep_leg(32, [128] * 3 + [256] * 3 + [512] * 2, [mx.FP64, mx.FP32])

    qc_sweeps = None
    if args.qc_sweeps > 0 and args.qc_norb > 0:
        qc_sweeps = qc_sweep_leg(args.qc_norb, [256, 512, 1024], [mx.FP64, mx.FP32])

    cpu = None
    parity = None
    if not args.no_cpu:
        print("cpu baseline", file=sys.stderr, flush=True)
        import oracle
        oracle.set_threads(1)
        reps = 0
        t0 = time.perf_counter()
        while True:
            ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
            reps += 1
            if time.perf_counter() - t0 > args.cpu_seconds or reps >= 200:
                break
        tc = (time.perf_counter() - t0) / reps
        cpu = dict(value=flops / tc / 1e12, unit="TFLOP/s", cores=1, kind="port",
                   sample="%d sigma applications of the headline workload (%.3f s each)" % (reps, tc),
                   sigma_per_s=1.0 / tc)
        scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values())
        parity = dict(same_blocks=one.block_set() == ref.block_set(), max_abs_diff=one.max_abs_diff(ref) / scale)

    res = {
        "metric": "H_psi TFLOP/s (two-site sigma vector, block GEMMs on MFMA)",
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
        "flops_per_sigma": flops,
        "bytes_per_sigma": nbytes,
        "kernels": per_kernel,
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "other_precision": other_leg,
        "batched": batched,
        "qc_merged": qcres,
        "sweeps": sweeps,
        "qc_sweeps": qc_sweeps,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
