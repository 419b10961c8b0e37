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
