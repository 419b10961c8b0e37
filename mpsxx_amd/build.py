"""Build the HIP library in-tree: mpsxx_amd/lib/libmpsxx_amd.so (gfx950).

Run as ``python -m mpsxx_amd.build`` or through ``__graft_entry__.build()``.  Objects are cached under
``mpsxx_amd/_build`` keyed by the source's modification time, so a rebuild only recompiles changed files; headers are
tracked conservatively (any header change recompiles everything).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libmpsxx_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + SRC,
          "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
# device code: no packed-FP32 instructions (the compiler is kept from forming v_pk_{add,mul,fma}_f32 from scalar or
# float2 code; the kernels' FP32 arithmetic stays in unpacked instructions)
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
HOST = ["-D__HIP_PLATFORM_AMD__"]
# per-source device options: the FP64 grouped GEMM keeps its MFMA accumulators in VGPRs (kernels_f64.hip)
EXTRA = {"kernels_f64.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
LIBS = ["-lrocsolver", "-lrocblas"]


def _sources():
    out = []
    for f in sorted(os.listdir(SRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(SRC, f))
    return out


def _headers_mtime():
    m = 0.0
    for d in (SRC, os.path.join(ROOT, "include")):
        for f in os.listdir(d):
            if f.endswith((".hpp", ".h")):
                m = max(m, os.path.getmtime(os.path.join(d, f)))
    return m


def _compile(src, force):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), _headers_mtime()):
        return obj
    flags = COMMON + (DEVICE if src.endswith(".hip") else HOST) + EXTRA.get(os.path.basename(src), [])
    cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(force=False, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
