"""Batched two-site sigma vectors over the C ABI (``mpsxx_batch_*`` in include/mpsxx_amd.h).

One ``SigmaBatch`` binds n independent (lmpo, rmpo, lopr, ropr, wfn0, sgv0) instances -- different chains, bond
dimensions and block structures -- and applies ``sgv0[i] += H_i wfn0[i]`` for all of them with one grouped launch per
contraction step (optionally replayed from a captured hipGraph).
"""
import ctypes

import numpy as np

from . import FP64, VP, DTensor, _check, lib

ALL_INPUTS = 31   # convert mask: refresh the FP32 copies of every input
WFN_ONLY = 16     # only the wavefunction changed (Davidson iterations)


class SigmaBatch:
    def __init__(self, instances, precision=FP64):
        """instances: sequence of (lmpo, rmpo, lopr, ropr, wfn0, sgv0) DTensors; empty sgv0 are shaped here."""
        self.inst = [tuple(x) for x in instances]  # keeps the tensors alive as long as the batch
        n = len(self.inst)
        arr = [(VP * n)(*[self.inst[i][k].h for i in range(n)]) for k in range(6)]
        self.h = VP()
        _check(lib().mpsxx_batch_sigma_twosite_create(ctypes.c_int32(n), *arr, ctypes.c_int(precision),
                                                      ctypes.byref(self.h), None))
        self.precision = precision

    def run(self, convert_mask=ALL_INPUTS, use_graph=True):
        _check(lib().mpsxx_batch_sigma_twosite_run(self.h, ctypes.c_uint32(convert_mask), ctypes.c_int(int(use_graph)),
                                                   None))

    def time(self, reps, convert_mask=ALL_INPUTS, use_graph=True):
        ms = ctypes.c_double()
        _check(lib().mpsxx_batch_time(self.h, ctypes.c_int32(reps), ctypes.c_uint32(convert_mask),
                                      ctypes.c_int(int(use_graph)), ctypes.byref(ms), None))
        return ms.value

    def profile(self, reps):
        """Per kernel launch: dict(kernel, step, ms, flops, bytes) (outputs redirected to scratch)."""
        return _profile_entries(lambda *a: lib().mpsxx_batch_profile(self.h, ctypes.c_int32(reps), *a, None))

    def tune(self, reps=5):
        n = 16
        pol = np.zeros(4 * n, dtype=np.int32)
        _check(lib().mpsxx_batch_tune(self.h, ctypes.c_int32(reps), ctypes.c_int32(n), pol.ctypes.data_as(VP), None))
        return [tuple(int(v) for v in pol[4 * k:4 * k + 4]) for k in range(self.info()["nsteps"])]

    def set_policy(self, policies):
        """One (small-block limit, split threshold, grid cap[, big-tile split target]) tuple per launch."""
        policies = [tuple(p) + (2048,) if len(p) == 3 else tuple(p) for p in policies]
        pol = np.ascontiguousarray(np.array(policies, dtype=np.int32).reshape(-1))
        _check(lib().mpsxx_batch_set_policy(self.h, ctypes.c_int32(len(policies)), pol.ctypes.data_as(VP)))

    def info(self):
        fl = ctypes.c_int64()
        by = ctypes.c_int64()
        ns = ctypes.c_int32()
        wi = ctypes.c_int64()
        _check(lib().mpsxx_batch_info(self.h, ctypes.byref(fl), ctypes.byref(by), ctypes.byref(ns), ctypes.byref(wi)))
        return {"flops": fl.value, "bytes": by.value, "nsteps": ns.value, "work_items": wi.value}

    def free(self):
        if self.h:
            lib().mpsxx_batch_free(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchDavidsonStats(ctypes.Structure):
    _fields_ = [("batch_sigmas", ctypes.c_int32), ("batch_sigmas_low", ctypes.c_int32),
                ("macro_iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("sigma_ms", ctypes.c_double),
                ("seconds", ctypes.c_double)]


def batch_davidson_twosite(instances, tol, precision=FP64, use_graph=True):
    """Davidson (src/diagonalize.hpp:283-449) on n independent two-site problems in lockstep, one batched sigma launch
    sequence and one set of segmented vector operations per step for all of them (mpsxx_batch_davidson_twosite).
    instances: sequence of (lmpo, rmpo, lopr, ropr, diag, wfnc) DTensors; every wfnc (start vector) is overwritten by
    its normalized eigenvector.  Returns (energies, n_sigma, n_sigma_low, stats dict); stats["residuals"] holds each
    instance's last ||r||^2."""
    inst = [tuple(x) for x in instances]
    n = len(inst)
    arr = [(VP * n)(*[inst[i][k].h for i in range(n)]) for k in range(6)]
    e = np.zeros(n)
    ns = np.zeros(n, dtype=np.int32)
    nl = np.zeros(n, dtype=np.int32)
    rn = np.zeros(n)
    st = BatchDavidsonStats()
    _check(lib().mpsxx_batch_davidson_twosite(ctypes.c_int32(n), *arr, ctypes.c_double(tol), ctypes.c_int(precision),
                                              ctypes.c_int(int(use_graph)), e.ctypes.data_as(VP),
                                              ns.ctypes.data_as(VP), nl.ctypes.data_as(VP), rn.ctypes.data_as(VP),
                                              ctypes.byref(st), None))
    stats = {k: getattr(st, k) for k, _ in BatchDavidsonStats._fields_}
    stats["residuals"] = rn
    return e, ns, nl, stats


KERNELS = ["gemm_grouped_kernel", "gemm_small_kernel", "gemm_split_kernel", "axpy_grouped_kernel", "gemm_x3_kernel",
           "gemm128_kernel", "gemm_x3_128_kernel"]


def _profile_entries(call, cap=256):
    ms = np.zeros(cap)
    fl = np.zeros(cap, dtype=np.int64)
    by = np.zeros(cap, dtype=np.int64)
    kn = np.zeros(cap, dtype=np.int32)
    st = np.zeros(cap, dtype=np.int32)
    n = ctypes.c_int32()
    _check(call(ctypes.c_int32(cap), ms.ctypes.data_as(VP), fl.ctypes.data_as(VP), by.ctypes.data_as(VP),
                kn.ctypes.data_as(VP), st.ctypes.data_as(VP), ctypes.byref(n)))
    return [dict(kernel=KERNELS[kn[i]], step=int(st[i]), ms=float(ms[i]), flops=int(fl[i]), bytes=int(by[i]))
            for i in range(n.value)]


def sigma_twosite_profile(lmpo, rmpo, lopr, ropr, wfn0, precision=FP64, reps=50):
    """Per kernel launch of the single-instance two-site sigma plan (see mpsxx_sigma_twosite_profile)."""
    hs = [t.h for t in (lmpo, rmpo, lopr, ropr, wfn0)]
    return _profile_entries(lambda *a: lib().mpsxx_sigma_twosite_profile(*hs, ctypes.c_int(precision),
                                                                         ctypes.c_int(reps), *a, None))


def dominant_kernel(entries):
    """Aggregate launches per kernel; return (name, dict(ms, flops, bytes, launches)) of the one with most time."""
    agg = {}
    for e in entries:
        a = agg.setdefault(e["kernel"], dict(ms=0.0, flops=0, bytes=0, launches=0))
        a["ms"] += e["ms"]
        a["flops"] += e["flops"]
        a["bytes"] += e["bytes"]
        a["launches"] += 1
    name = max(agg, key=lambda k: agg[k]["ms"])
    return name, agg[name], agg


def last_plan_tiles():
    """Work items of the plan of the most recent hot-path call (mpsxx_last_plan_tiles): 64x64 tiles, 64x64 tiles of
    split reductions, one-wave 32x32 tiles, split 32x32 tiles, split 32x32 tiles cut over work-groups, AXPY chunks,
    128x128 tiles, 128x128 tiles of split reductions."""
    t = (ctypes.c_int64 * 8)()
    _check(lib().mpsxx_last_plan_tiles(t))
    keys = ("big", "big_red", "small", "split", "split_cut", "axpy_chunks", "big128", "big128_red")
    return dict(zip(keys, (int(x) for x in t)))


def last_plan_profile(reps=20):
    """Per kernel launch of the plan of the most recent hot-path call (see mpsxx_last_plan_profile)."""
    pf = ctypes.c_int64()
    ent = _profile_entries(lambda *a: lib().mpsxx_last_plan_profile(ctypes.c_int(reps), *a, ctypes.byref(pf), None))
    return ent, pf.value
