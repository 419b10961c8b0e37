"""The C-ABI library builds, loads and exports every entry point declared in include/mpsxx_amd.h (CPU)."""
import ctypes
import os
import re
import subprocess

import mpsxx_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mpsxx_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpsxx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = mpsxx_amd.lib()
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", mpsxx_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mpsxx_[a-z0-9_]+)", out))
    assert set(syms) <= exported


def test_last_error_is_a_string():
    lib = mpsxx_amd.lib()
    assert isinstance(lib.mpsxx_last_error(), bytes)


def test_descriptor_struct_layout():
    # mpsxx_qst_desc: 4 int32 + 3 pointers + int64 + pointer
    assert ctypes.sizeof(mpsxx_amd.QstDesc) == 4 * 4 + 3 * 8 + 8 + 8


def test_diagnostic_counters_are_host_only_and_consistent():
    """The plan-cache and device-allocation counters (no reference counterpart) are readable without a device: counts
    are non-negative, times finite, and NULL outputs are accepted."""
    pc = mpsxx_amd.plan_cache_stats()
    assert set(pc) == {"hits", "builds", "evictions", "build_ms", "plans", "device_bytes", "releases", "restores",
                       "collisions"}
    assert all(v >= 0 for v in pc.values())
    assert pc["plans"] <= pc["builds"]
    assert pc["restores"] <= pc["releases"] <= pc["builds"] + pc["restores"]
    dm = mpsxx_amd.device_memory_stats()
    assert set(dm) == {"mallocs", "frees", "malloc_ms", "free_ms", "reuses", "pool_syncs", "bytes_live",
                       "cached_bytes"}
    assert all(v >= 0 for v in dm.values())
    lib = mpsxx_amd.lib()
    assert lib.mpsxx_plan_cache_stats(None, None, None, None, None, None, None, None) == 0
    assert lib.mpsxx_device_memory_stats(None, None, None, None, None, None, None, None) == 0


def test_rocsolver_argument_check_on_the_cpu():
    """The host check before every rocSOLVER batched SVD call (the fallback path): the padded layout the library uses
    passes, and each rule, broken on its own, is reported by name (no device involved)."""
    ERR_HIP = 3  # MPSXX_ERR_HIP
    lib = mpsxx_amd.lib()
    lib.mpsxx_last_error.restype = ctypes.c_char_p
    f = lib.mpsxx_check_gesvd_args
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_int64, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int] + [ctypes.c_int64] * 5
    # the library's layout: nb blocks zero-padded to (mmax x nmax), column-major view nmax x mmax, kmax = min
    nb, mmax, nmax = 7, 40, 25
    kmax = min(mmax, nmax)
    sa, su, sv = mmax * nmax, nmax * kmax, kmax * mmax
    good = dict(m=nmax, n=mmax, lda=nmax, sA=sa, sS=kmax, ldu=nmax, sU=su, ldv=kmax, sV=sv, sE=max(kmax, 2),
                batch=nb, bA=8 * nb * sa, bS=8 * nb * kmax, bU=8 * nb * su, bV=8 * nb * sv, bE=8 * nb * max(kmax, 2))

    def call(**kw):
        a = dict(good, **kw)
        return f(a["m"], a["n"], a["lda"], a["sA"], a["sS"], a["ldu"], a["sU"], a["ldv"], a["sV"], a["sE"],
                 a["batch"], a["bA"], a["bS"], a["bU"], a["bV"], a["bE"])

    assert call() == 0
    assert call(bE=-1, sE=0) == 0                      # the dgesvdj retry passes no E array
    assert call(batch=3) == 0                          # a prefix of the slots
    cases = {"lda / strideA": [dict(lda=nmax - 1), dict(sA=sa - 1)],
             "strideS": [dict(sS=kmax - 1)],
             "ldu / strideU": [dict(ldu=nmax - 1), dict(sU=su - 1)],
             "ldv / strideV": [dict(ldv=kmax - 1), dict(sV=sv - 1)],
             "strideE": [dict(sE=kmax - 2)],
             "A buffer": [dict(bA=8 * nb * sa - 8), dict(batch=nb + 1, bS=10 ** 9, bU=10 ** 9, bV=10 ** 9,
                                                          bE=10 ** 9)],
             "S buffer": [dict(bS=8 * nb * kmax - 8)],
             "U buffer": [dict(bU=8 * nb * su - 8)],
             "V buffer": [dict(bV=8 * nb * sv - 8)],
             "E buffer": [dict(bE=8 * ((nb - 1) * max(kmax, 2) + kmax - 1) - 8)],
             "empty problem": [dict(m=0), dict(batch=0)]}
    for rule, variants in cases.items():
        for kw in variants:
            assert call(**kw) == ERR_HIP, (rule, kw)
            assert rule in lib.mpsxx_last_error().decode(), (rule, kw, lib.mpsxx_last_error())
