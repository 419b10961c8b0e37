#!/usr/bin/env python3
"""Benchmark: H_eff|psi> TFLOP/s and DMRG sweep time on one MI355X, FP32-MFMA mixed precision beside FP64.

Headline workload: the merged-form two-site sigma vector (src/compute_sigma_vector.hpp:86-96) at the middle bond
(sites 11, 12) of a 24-orbital quantum-chemistry chain, N = 24, 2Sz = 0, bond dimension M = 1024.  Integrals are
synthetic from seed 1 (h_pq symmetric, diagonal -2..-0.5 linear, off-diagonal N(0, 0.05); v_pqrs = sum_k
L^k_pq L^k_rs over 48 symmetric L^k with N(0, 0.1) entries), the MPO is the QC generator's (D_w = 1226 at the middle
bond).  Sector dimensions follow a converged chain's spectrum shape and every allowed block holds random entries
(synthetic data, mpsxx_amd/workloads.py); the merged operators L~, R~ are built on the device from the environments
(compute_merged_operator) and stay resident, as inside Davidson.  One step = one sigma application with its inputs
resident in HBM; `value` = algorithmic TFLOP/s of the FP32-MFMA sigma (flops = 2 m n k over the block pairs of the
reference contraction order, counted by the planner).

Beside it, on the same inputs: the FP64 and split-BF16x3 sigma.  The sigma on a real state: the merged operators at
the middle bond of the state the FP64 sweeps below converge to.  The sweep half of the metric: merged two-site
sweeps of the same chain over M 256, 512, then 1024 until the energy changes by less than 1e-10, in FP64 and in
mixed precision (FP32 or split-BF16x3 sigma vectors, FP64 Davidson, refinement and SVD), with sweep times and
|E - E_fp64|; each precision's sweeps start with the library's plan and buffer caches cleared
(mpsxx_release_caches), so that none runs on plans an earlier leg built.  The CPU baseline: the oracle (FP64 restatement of the reference's block contraction) with its block
dgemm in numpy's bundled OpenBLAS, on one core and on all cores given to the process, over a bounded sample of the
headline sigma (the terms of every k-th sector of the MPO bond index; sigma is a sum over that index), compared
with the device's FP64 result on the same sample.  Secondary legs: the two-site sigma of the Hubbard chain L = 32,
M = 512; 64 independent Hubbard chains as one grouped launch; the ground-state problems of the same 64 chains solved
by one batched lockstep Davidson (FP64, and FP32 / split-BF16x3 with FP64 refinement) beside the one-by-one FP64
Davidson, with per-instance |E - E_fp64|; two-site sweeps of the Hubbard chain L = 32.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mpsxx_amd as mx  # noqa: E402
from mpsxx_amd import workloads  # noqa: E402
from mpsxx_amd.batch import (ALL_INPUTS, WFN_ONLY, SigmaBatch, dominant_kernel, last_plan_profile,  # noqa: E402
                             last_plan_tiles, sigma_twosite_profile)
from mpsxx_amd.qst import QST  # noqa: E402

METRIC = "H·ψ TFLOP/s + sweep time, FP32-MFMA vs FP64 GPU at |ΔE|≤1e-8 Eh"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {mx.FP64: 78.6, mx.FP32: 157.3, mx.BF16X3: 2500.0}
PNAME = {mx.FP64: "fp64", mx.FP32: "fp32", mx.BF16X3: "bf16x3"}
HEADLINE_TRAFFIC = "traffic_sigma_merged_qc_norb24_M1024.md"
HUBBARD_TRAFFIC = "traffic_sigma_twosite_hubbard_L32_M512.md"


def measured_traffic(fname, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, scaled
    by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
            return int(parts[2])
    return None


def roofline_of(entries, prec):
    """Roofline of the kernel with the most device time over all its launches of one application: algorithmic
    flops (bytes for the AXPY kernel) / device time from HIP events around each launch on its stream.  For split
    BF16x3 the flops are the bf16 MFMA products issued (three per FP32 product) against the bf16 peak."""
    name, agg, allk = dominant_kernel(entries)
    per_kernel = {k: dict(ms=v["ms"], launches=v["launches"], flops=v["flops"], bytes=v["bytes"],
                          avg_launch_us=1e3 * v["ms"] / v["launches"],
                          tflops=v["flops"] / max(v["ms"], 1e-12) / 1e9) for k, v in allk.items()}
    sec = agg["ms"] * 1e-3
    if name == "axpy_grouped_kernel":
        achieved = agg["bytes"] / sec / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS)
    else:
        mult = 3 if prec == mx.BF16X3 else 1
        achieved = mult * agg["flops"] / sec / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=PEAK_TFLOPS[prec], unit="TFLOP/s",
                    frac=achieved / PEAK_TFLOPS[prec])
        if prec == mx.BF16X3:
            roof["note"] = "bf16 MFMA products issued (3 per FP32 product) against the dense bf16 peak"
    roof.update(traffic=None, kernel=name, launches_per_sigma=agg["launches"],
                avg_launch_us=1e3 * agg["ms"] / agg["launches"], algorithmic_flops=agg["flops"],
                algorithmic_bytes=agg["bytes"], device_ms_per_sigma=sum(v["ms"] for v in allk.values()))
    return roof, per_kernel


def rel_diff(a, b):
    return a.diff_norm(b) / b.norm()


def qdot(a, b):
    """<a|b> over the common blocks of two host block-sparse arrays."""
    return float(sum(np.vdot(a.blocks[k], b.blocks[k]) for k in a.blocks if k in b.blocks))


# -- merged sigma ------------------------------------------------------------------------------------------------
def merged_leg(Lm, Rm, psim, prec, steps, warmup, ref=None):
    """Time `steps` merged sigma applications with the operators resident (their FP32 / split copies made by the
    first call and kept, only the wavefunction refreshed per call, as inside Davidson's multH); profile each kernel
    launch of the plan with HIP events; one fresh application for the accuracy against `ref`."""
    fresh = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=prec)  # converts every input once
    mx.synchronize()
    one = fresh.to_qst()
    del fresh
    out = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=prec)
    mask = 1 << 2
    for _ in range(max(warmup, 1)):
        mx.compute_sigma_vector_merged_ex(Lm, Rm, psim, out, prec, mask)
    mx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mx.compute_sigma_vector_merged_ex(Lm, Rm, psim, out, prec, mask)
    mx.synchronize()
    sec = (time.perf_counter() - t0) / steps
    ent, pf = last_plan_profile(3)
    tiles = last_plan_tiles()
    roof, per_kernel = roofline_of(ent, prec)
    r = dict(ms=sec * 1e3, tflops=pf / sec / 1e12, flops_per_sigma=pf, roofline=roof, kernels=per_kernel,
             tiles=tiles)
    if ref is not None:
        r["rel_diff_vs_fp64"] = rel_diff(one, ref)
    del out
    return r, one, sec, pf


def sample_sectors(Lh, Rh, stride):
    """L~ and R~ restricted to every `stride`-th sector of their shared MPO bond index (index 1): sigma =
    sum over that index of independent terms, so the restriction is the sum over the sample's sectors."""
    keep = set(range(0, len(Lh.qshape[1]), stride))
    out = []
    for X in (Lh, Rh):
        Y = QST(X.q, X.qshape, X.dshape, X.qsym)
        Y.blocks = {k: v for k, v in X.blocks.items() if k[1] in keep}
        out.append(Y)
    return out[0], out[1], len(keep)


def cpu_cores():
    """Cores the process may use: its affinity mask, capped by OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(Lm, Rm, psim, total_flops, cpu_seconds):
    """The oracle's merged sigma (FP64, block dgemm in OpenBLAS) on a sample of the headline sigma sized for about
    `cpu_seconds` on one core (calibrated on a 1/64 sample), timed on one core and on all cores, and the device FP64
    sigma of the same sample against it."""
    import oracle
    Lh, Rh, ph = Lm.to_qst(), Rm.to_qst(), psim.to_qst()
    nw = len(Lh.qshape[1])

    def device_sample(Ls, Rs):
        dL, dR = mx.DTensor(Ls), mx.DTensor(Rs)
        y = mx.compute_sigma_vector_merged(dL, dR, psim, precision=mx.FP64)
        mx.synchronize()
        _, pf = last_plan_profile(1)
        return y.to_qst(), pf

    def timed(Ls, Rs, threads):
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        y = oracle.sigma_merged(Ls, Rs, ph)
        return y, time.perf_counter() - t0

    blas = oracle.use_blas(True, 1)
    try:
        Ls, Rs, _ = sample_sectors(Lh, Rh, 64)
        _, pf_probe = device_sample(Ls, Rs)
        _, t_probe = timed(Ls, Rs, 1)
        stride = max(1, int(math.ceil(total_flops / max(pf_probe / t_probe * cpu_seconds, 1.0))))
        stride = min(stride, nw)
        Ls, Rs, nkeep = sample_sectors(Lh, Rh, stride)
        dev, pf = device_sample(Ls, Rs)
        y1, t1 = timed(Ls, Rs, 1)
        nc = cpu_cores()
        yn, tn = timed(Ls, Rs, nc)
    finally:
        oracle.use_blas(False)
        oracle.set_threads(1)
    scale = max(float(np.max(np.abs(b))) for b in y1.blocks.values() if b.size)
    sample = ("the merged sigma terms of every %d-th of the %d sectors of the MPO bond index (%d sectors, %.3g of "
              "the %.3g flops of one sigma)" % (stride, nw, nkeep, pf, total_flops))
    machine = dict(nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)), cpu_model=cpu_model(),
                   dgemm="numpy's bundled " + blas.strip(), threading="OpenMP over output blocks, 1 BLAS thread each")
    one = dict(value=pf / t1 / 1e12, unit="TFLOP/s", cores=1, kind="port", sample=sample, seconds=t1,
               sigma_seconds_extrapolated=total_flops * t1 / pf)
    allc = dict(value=pf / tn / 1e12, unit="TFLOP/s", cores=nc, kind="port", sample=sample, seconds=tn,
                sigma_seconds_extrapolated=total_flops * tn / pf, single_core=one, **machine)
    parity = dict(same_blocks=dev.block_set() == y1.block_set(), max_abs_diff=dev.max_abs_diff(y1) / scale,
                  all_cores_max_abs_diff=yn.max_abs_diff(y1) / scale, sample_flops=pf)
    return allc, parity


# -- sweeps --------------------------------------------------------------------------------------------------------
def qc_sweep_leg(norb, schedule, precisions, max_final=6, keep_fp64_state=False):
    """Merged two-site sweeps (src/make_sweep.cpp:112) of the QC chain with synthetic integrals from seed 1 over the
    M schedule, the last M repeated until the sweep energy changes by less than 1e-10 (src/dmrg.cpp:44-78, at most
    max_final sweeps at that M): per precision the total and per-sweep seconds, the converged energy and
    |E - E_fp64|.  With keep_fp64_state, also the MPO and site tensors of the converged FP64 state."""
    from mpsxx_amd import qc
    from mpsxx_amd.dmrg import SWEEP_MERGED, DeviceDMRG, initial_mps
    h, v = qc.synthetic_integrals(norb, seed=1)
    m = qc.QcMpo(h, v)
    mpos = m.devices()
    out = {"norb": norb, "schedule": schedule, "tol": 1e-10, "D_w_max": int(max(m.bond_dims()))}
    e64, state = None, None
    for prec in precisions:
        name = {mx.FP64: "fp64", mx.FP32: "fp32_mixed", mx.BF16X3: "bf16x3_mixed"}[prec]
        print("qc sweeps: %s" % name, file=sys.stderr, flush=True)
        mx.release_caches()  # every precision starts without the plans (and pooled buffers) of the one before
        dm = DeviceDMRG(mpos, initial_mps(norb, (norb, 0), schedule[0], 0), schedule[0], form=SWEEP_MERGED)
        secs, e, st = [], None, None
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
        r = {"seconds_total": time.perf_counter() - t0, "seconds_per_sweep": secs, "sweeps": len(secs),
             "energy": e, "last_sweep_svd_s": st["svd_s"], "last_sweep_davidson_s": st["davidson_s"],
             "last_sweep_renorm_s": st["renorm_s"], "max_bond": st["max_bond"]}
        if prec == mx.FP64:
            e64 = e
            if keep_fp64_state:
                state = (mpos, [dm.site_device(k) for k in range(norb)])
        elif e64 is not None:
            r["abs_dE_vs_fp64"] = abs(e - e64)
        out[name] = r
        del dm
    if "fp64" in out:
        for name in ("fp32_mixed", "bf16x3_mixed"):
            if name in out:
                out[name]["speedup_vs_fp64"] = out["fp64"]["seconds_total"] / out[name]["seconds_total"]
    return out, state


def hubbard_sweep_leg(L, schedule, precisions):
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
        print("hubbard sweeps: %s" % name, file=sys.stderr, flush=True)
        mx.release_caches()  # every precision starts without the plans (and pooled buffers) of the one before
        dm = DeviceDMRG(mpos, init, schedule[0])
        secs, e, st = [], None, None
        t0 = time.perf_counter()
        for M in schedule:
            e, st = dm.sweep(M, 1e-10, precision=prec)
            secs.append(round(st["seconds"], 3))
        r = {"seconds_total": time.perf_counter() - t0, "seconds_per_sweep": secs, "energy": e,
             "last_sweep_svd_s": st["svd_s"], "last_sweep_davidson_s": st["davidson_s"],
             "last_sweep_sigma_ms": st["sigma_ms"]}
        if prec == mx.FP64:
            e64 = e
        elif e64 is not None:
            r["abs_dE_vs_fp64"] = abs(e - e64)
        out[name] = r
    return out


# -- secondary sigma legs ------------------------------------------------------------------------------------------
def hubbard_sigma_leg(L, M, U, steps, warmup):
    """Two-site sigma (src/compute_sigma_vector.hpp:54-70) of the open Hubbard chain L, M at the middle bond (compact
    D_w = 6 MPO, synthetic fill), FP64 and FP32, with its per-kernel roofline."""
    lm, rm, le, re_, psi = workloads.hubbard_twosite(L, M, U)
    dl = [mx.DTensor(x) for x in (lm, rm, le, re_, psi)]
    res = dict(L=L, M=M, U=U, psi_blocks=len(psi.blocks), psi_elems=psi.nelem())
    ref = None
    for prec in (mx.FP64, mx.FP32):
        out = mx.DTensor()
        for _ in range(max(warmup, 1) + 1):
            mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
        mx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            mx.compute_sigma_vector_twosite(*dl, sgv0=out, precision=prec)
        mx.synchronize()
        sec = (time.perf_counter() - t0) / steps
        flops = ctypes.c_int64()
        nbytes = ctypes.c_int64()
        nsteps = ctypes.c_int32()
        mx._check(mx.lib().mpsxx_sigma_twosite_stats(*[d.h for d in dl], ctypes.c_int(prec), ctypes.byref(flops),
                                                     ctypes.byref(nbytes), ctypes.byref(nsteps)))
        roof, per_kernel = roofline_of(sigma_twosite_profile(*dl, precision=prec, reps=50), prec)
        if prec == mx.FP64:
            roof["traffic"] = measured_traffic(HUBBARD_TRAFFIC, roof["kernel"])
        one = mx.compute_sigma_vector_twosite(*dl, precision=prec).to_qst()
        r = dict(ms=sec * 1e3, tflops=flops.value / sec / 1e12, flops_per_sigma=flops.value,
                 bytes_per_sigma=nbytes.value, roofline=roof, kernels=per_kernel)
        if ref is None:
            ref = one
        else:
            r["rel_diff_vs_fp64"] = rel_diff(one, ref)
        res[PNAME[prec]] = r
    return res


def batched_leg(n, reps):
    """n independent Hubbard chains (L in 16..64, U in {2,4,8}, M in {128,256,512}) applied sequentially in FP64
    and as one grouped launch per step (FP64 and FP32, hipGraph replay, autotuned tiles)."""
    specs = workloads.batch_specs(n, 4)
    probs = [workloads.hubbard_twosite(L=L, M=M, U=U, seed=k) for k, (L, U, M) in enumerate(specs)]
    dev = [tuple(mx.DTensor(x) for x in p) for p in probs]
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
        for t in inst:
            t[5].clear()
        b.run(ALL_INPUTS, True)
        mx.synchronize()
        outs_p = [t[5].to_qst() for t in inst]
        res[name + "_max_rel_err"] = max(rel_diff(o, r) for o, r in zip(outs_p, ref))
        b.free()
    fl = res["flops_per_batch"]
    for k in list(res):
        if k.endswith("_ms"):
            res[k[:-3] + "_tflops"] = fl / (res[k] * 1e-3) / 1e12
            res[k[:-3] + "_sigma_per_s"] = n / (res[k] * 1e-3)
    res["speedup_fp32_batched_vs_fp64_sequential"] = res["seq_fp64_ms"] / res["fp32_graph_wfn_ms"]
    res["speedup_fp64_batched_vs_fp64_sequential"] = res["seq_fp64_ms"] / res["fp64_graph_ms"]
    return res


def batched_davidson_leg(n, tol=1e-10, max_calls=20):
    """Config 4 as a ground-state problem: the middle two-site problem of each of the n chains of the batched leg,
    environments from isometric configuration-selected MPS (workloads.hubbard_isometric_problem), Davidson to
    ||r||^2 < tol in FP64 instance by instance (mpsxx_davidson_twosite, the plain FP64 GPU path) and as one batched
    lockstep solve (mpsxx_batch_davidson_twosite) in FP64, FP32 + FP64 refinement and split-BF16x3 + FP64 refinement.
    Each solve is repeated from its own result until converged (at most max_calls times; one call is the reference's
    <= 4 x 10 sigma vectors), as successive sweeps restart it; later batched calls take only the instances still above
    the tolerance.  Reports per-instance |E - E_fp64| over the instances converged in both runs and the sigma vectors
    per second of each form (sigma vectors applied to instances still iterating / wall time)."""
    from mpsxx_amd.batch import batch_davidson_twosite
    specs = workloads.batch_specs(n, 4)
    t0 = time.perf_counter()
    probs = [workloads.hubbard_isometric_problem(L, M, U, seed=k) for k, (L, U, M) in enumerate(specs)]
    mx.synchronize()
    setup_s = time.perf_counter() - t0
    flops = []
    for p in probs:
        fl = ctypes.c_int64()
        mx._check(mx.lib().mpsxx_sigma_twosite_stats(*[x.h for x in p[:4]], p[5].h, ctypes.c_int(mx.FP64),
                                                     ctypes.byref(fl), None, None))
        flops.append(fl.value)
    flops = np.array(flops, dtype=np.float64)
    res = dict(instances=n, sample_seed=4, tol=tol, setup_s=setup_s,
               environments="isometric, configuration-selected (fewest doubly occupied sites), renormalized on the "
                            "device; start vector x / (diag - min diag + 1/2)")
    # plain FP64 path, one instance after another
    e_single, ns_single = np.zeros(n), np.zeros(n, dtype=np.int64)
    mx.synchronize()
    t0 = time.perf_counter()
    for k, p in enumerate(probs):
        w = mx.copy(p[5])
        e_prev = None
        for _ in range(max_calls):
            e, ns = mx.davidson_twosite([p[:4]], p[4], w, tol, mx.FP64)
            ns_single[k] += ns
            if e_prev is not None and abs(e - e_prev) < 1e-13:
                break
            e_prev = e
        e_single[k] = e
        del w
    mx.synchronize()
    sec = time.perf_counter() - t0
    res["single_fp64"] = dict(seconds=sec, sigmas=int(ns_single.sum()), sigma_per_s=float(ns_single.sum()) / sec,
                              tflops=float((ns_single * flops).sum()) / sec / 1e12)
    e64 = None
    for prec in (mx.FP64, mx.FP32, mx.BF16X3):
        print("batched davidson: %s" % PNAME[prec], file=sys.stderr, flush=True)
        ws = [mx.copy(p[5]) for p in probs]
        e, rn = np.zeros(n), np.full(n, np.inf)
        ns_tot, nl_tot = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
        act, calls, sig_ms = list(range(n)), 0, 0.0
        mx.synchronize()
        t0 = time.perf_counter()
        while act and calls < max_calls:
            # later calls take only the instances still above the tolerance
            ek, ns, nl, st = batch_davidson_twosite([probs[i][:5] + (ws[i],) for i in act], tol, prec, True)
            for j, i in enumerate(act):
                e[i], rn[i] = ek[j], st["residuals"][j]
                ns_tot[i] += ns[j]
                nl_tot[i] += nl[j]
            calls += 1
            sig_ms += st["sigma_ms"]
            act = [i for i in act if rn[i] >= tol]
        sec = time.perf_counter() - t0
        conv = rn < tol
        r = dict(seconds=sec, calls=calls, converged=int(conv.sum()), sigmas=int(ns_tot.sum()),
                 sigmas_low=int(nl_tot.sum()), sigma_per_s=float(ns_tot.sum()) / sec,
                 tflops=float((ns_tot * flops).sum()) / sec / 1e12, batched_sigma_device_s=sig_ms * 1e-3,
                 speedup_vs_single_fp64=res["single_fp64"]["seconds"] / sec)
        if prec == mx.FP64:
            e64, conv64 = e.copy(), conv.copy()
            both = conv
            r["max_abs_dE_vs_single_fp64"] = float(np.max(np.abs(e - e_single)[both])) if both.any() else None
        else:
            d = np.abs(e - e64)
            both = conv & conv64
            r.update(converged_both=int(both.sum()),
                     max_abs_dE_vs_fp64=float(d[both].max()) if both.any() else None,
                     median_abs_dE_vs_fp64=float(np.median(d[both])) if both.any() else None,
                     instances_within_1e8=int((d[both] <= 1e-8).sum()), abs_dE_vs_fp64=[float(x) for x in d])
        res[{mx.FP64: "batched_fp64", mx.FP32: "batched_fp32_mixed", mx.BF16X3: "batched_bf16x3_mixed"}[prec]] = r
        del ws
    res["energies_fp64"] = [float(x) for x in e64]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed sigma applications of the headline workload")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp32", "bf16x3", "fp64"], default="fp32",
                    help="block-GEMM precision of the headline sigma")
    ap.add_argument("--qc-norb", type=int, default=24)
    ap.add_argument("--qc-M", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target one-core seconds of the CPU sample")
    ap.add_argument("--qc-sweeps", type=int, default=1, help="QC merged sweeps, M 256, 512, 1024.. (0: skip)")
    ap.add_argument("--snapshot", type=int, default=1, help="sigma on the converged FP64 sweep state (0: skip)")
    ap.add_argument("--hubbard", type=int, default=1, help="Hubbard L=32 M=512 two-site sigma leg (0: skip)")
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched leg (0: skip)")
    ap.add_argument("--sweeps", type=int, default=1, help="Hubbard L=32 sweeps, M 128x3, 256x3, 512x2 (0: skip)")
    ap.add_argument("--batch-davidson", type=int, default=64,
                    help="instances of the batched Davidson (ground-state) leg (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X: the sigma-vector path runs on a single GPU in this build "
              "(sharding over operator groups is design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    prec = {"fp32": mx.FP32, "bf16x3": mx.BF16X3, "fp64": mx.FP64}[args.precision]
    print("headline: merged sigma norb=%d M=%d" % (args.qc_norb, args.qc_M), file=sys.stderr, flush=True)
    Lm, Rm, psim, info = workloads.qc_merged_problem(args.qc_norb, args.qc_M, 1)
    beside = {}
    r64, ref, _, _ = merged_leg(Lm, Rm, psim, mx.FP64, args.steps, args.warmup)
    beside["fp64"] = r64
    head = None
    for p in (mx.FP32, mx.BF16X3):
        r, _, sec, pf = merged_leg(Lm, Rm, psim, p, args.steps, args.warmup, ref)
        r["speedup_vs_fp64"] = r64["ms"] / r["ms"]
        beside[PNAME[p]] = r
        if p == prec:
            head = (r, sec, pf)
    if prec == mx.FP64:
        head = (r64, r64["ms"] * 1e-3, r64["flops_per_sigma"])
    hr, sec, flops = head
    roof = dict(hr["roofline"])
    if args.qc_norb == 24 and args.qc_M == 1024:
        roof["traffic"] = measured_traffic(HEADLINE_TRAFFIC, "%s:%s" % (PNAME[prec], roof["kernel"]))
        if roof["traffic"] is not None:
            roof["traffic_source"] = "profiles/%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE, copy-kernel " \
                                     "calibration)" % HEADLINE_TRAFFIC

    cpu, parity = None, None
    if not args.no_cpu:
        print("cpu baseline", file=sys.stderr, flush=True)
        cpu, parity = cpu_leg(Lm, Rm, psim, flops, args.cpu_seconds)
    del Lm, Rm, psim

    qc_sweeps, snapshot = None, None
    if args.qc_sweeps > 0:
        precs = [mx.FP64, mx.FP32, mx.BF16X3]
        qc_sweeps, state = qc_sweep_leg(args.qc_norb, [256, 512, args.qc_M], precs,
                                        keep_fp64_state=args.snapshot > 0)
        if state is not None:
            print("snapshot sigma", file=sys.stderr, flush=True)
            mpos, sites = state
            Ls, Rs, ps, sinfo = workloads.merged_problem_from_state(mpos, sites, args.qc_norb // 2 - 1)
            del sites
            snapshot = dict(sinfo, state="converged FP64 merged-sweep state at M=%d (site tensors as the last sweep "
                                         "left them)" % args.qc_M)
            s64, sref, _, _ = merged_leg(Ls, Rs, ps, mx.FP64, 10, 2)
            snapshot["fp64"] = s64
            for p in (mx.FP32, mx.BF16X3):
                r, _, _, _ = merged_leg(Ls, Rs, ps, p, 10, 2, sref)
                r["speedup_vs_fp64"] = s64["ms"] / r["ms"]
                snapshot[PNAME[p]] = r
            del Ls, Rs, ps

    hub = None
    if args.hubbard > 0:
        print("hubbard sigma", file=sys.stderr, flush=True)
        hub = hubbard_sigma_leg(32, 512, 4.0, 200, 20)
    batched = None
    if args.batch > 0:
        print("batched leg: %d instances" % args.batch, file=sys.stderr, flush=True)
        batched = batched_leg(args.batch, 20)
    bdav = None
    if args.batch_davidson > 0:
        print("batched davidson leg: %d instances" % args.batch_davidson, file=sys.stderr, flush=True)
        bdav = batched_davidson_leg(args.batch_davidson)
    hsweeps = None
    if args.sweeps > 0:
        hsweeps = hubbard_sweep_leg(32, [128] * 3 + [256] * 3 + [512] * 2, [mx.FP64, mx.FP32])

    sweep_time = None
    if qc_sweeps is not None:
        mixed = "fp32_mixed" if prec != mx.BF16X3 else "bf16x3_mixed"
        sweep_time = dict(workload="qc merged sweeps norb=%d, M 256, 512, %d until |dE| < 1e-10" % (
            args.qc_norb, args.qc_M), fp64_s=qc_sweeps["fp64"]["seconds_total"],
            mixed=mixed, mixed_s=qc_sweeps[mixed]["seconds_total"],
            speedup=qc_sweeps[mixed]["speedup_vs_fp64"], abs_dE=qc_sweeps[mixed]["abs_dE_vs_fp64"],
            abs_dE_ok=qc_sweeps[mixed]["abs_dE_vs_fp64"] <= 1e-8)

    res = {
        "metric": METRIC,
        "value": hr["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {mx.FP32: "f32", mx.BF16X3: "bf16x3", mx.FP64: "f64"}[prec],
        "data": "synthetic",
        "config": {"workload": "qc_sigma_merged", "norb": args.qc_norb, "N": args.qc_norb, "twoSz": 0,
                   "M": args.qc_M, "integrals": "synthetic, seed 1", "mpo": "QC generator, D_w=%d" % info["D_w"],
                   "site": info["site"], "psi_blocks": info["psi_blocks"], "psi_elems": info["psi_elems"],
                   "env_elems": info["env_elems"], "precision": args.precision,
                   "operators": "resident (FP32 / split copies kept), wavefunction refreshed per call"},
        "sigma_per_s": 1.0 / sec,
        "flops_per_sigma": flops,
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "sweep_time": sweep_time,
        "headline_kernels": hr["kernels"],
        "precisions": beside,
        "snapshot": snapshot,
        "qc_sweeps": qc_sweeps,
        "hubbard_sigma": hub,
        "batched": batched,
        "batched_davidson": bdav,
        "hubbard_sweeps": hsweeps,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
