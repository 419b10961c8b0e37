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
