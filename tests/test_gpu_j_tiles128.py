"""The 128x128 tiles of large blocks (gemm128_kernel for FP64 and FP32; gemm_x3_128_kernel for split BF16x3) against
the oracle (needs an MI355X).

Blocks of 100..230 rows and columns make 128x128 tiles with partial edges in both directions, reductions whose k is
not a multiple of the 16- or 32-wide panels, operands stored with unit stride along k and along the rows or columns
(the two-site sigma's permuted operands), and split reductions over several work-groups (the merged QC sigma's second
product).  Bars: FP32 2e-5, split BF16x3 1e-4 and FP64 1e-13 relative Frobenius norm (FP64 also 1e-12 of the output
scale in the 64x64-tile child)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from helpers import random_twosite_problem
from mpsxx_amd import workloads
from mpsxx_amd.batch import last_plan_profile, last_plan_tiles

pytestmark = pytest.mark.gpu


def _two_site(seed, lo=100, hi=230):
    rng = np.random.default_rng(seed)
    return random_twosite_problem(10, 4, rng, lambda q: int(rng.integers(lo, hi + 1)) if abs(q[1]) <= 1 else 0)


@pytest.mark.parametrize("prec,tol", [(mx.FP64, 1e-13), (mx.FP32, 2e-5), (mx.BF16X3, 1e-4)])
def test_two_site_sigma_on_128_tiles(prec, tol):
    print("128 tiles two-site", prec, flush=True)
    lm, rm, le, re_, psi = _two_site(31)
    ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
    d = [mx.DTensor(x) for x in (lm, rm, le, re_, psi)]
    got = mx.compute_sigma_vector_twosite(*d, precision=prec).to_qst()
    tiles = last_plan_tiles()
    kernels = {e["kernel"] for e in last_plan_profile(1)[0]}
    rel = got.diff_norm(ref) / ref.norm()
    print("tiles", tiles, "kernels", sorted(kernels), "rel", rel, flush=True)
    assert tiles["big128"] > 0
    assert ("gemm_x3_128_kernel" if prec == mx.BF16X3 else "gemm128_kernel") in kernels
    assert got.block_set() == ref.block_set()
    assert rel <= tol


_MERGED_CHILD = r"""
import json, sys
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/tests"]
import mpsxx_amd as mx
import oracle
from mpsxx_amd import workloads
from mpsxx_amd.batch import last_plan_tiles
prec = int(sys.argv[2])
Lm, Rm, psim, info = workloads.qc_merged_problem(12, 384, seed=5)
ref = oracle.sigma_merged(Lm.to_qst(), Rm.to_qst(), psim.to_qst())
got = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=prec).to_qst()
t = last_plan_tiles()
print(json.dumps([got.block_set() == ref.block_set(), got.diff_norm(ref) / ref.norm(), t["big128"], t["big128_red"]]),
      flush=True)
"""


@pytest.mark.parametrize("prec,tol", [(mx.FP64, 1e-13), (mx.FP32, 2e-5), (mx.BF16X3, 1e-4)])
def test_merged_qc_sigma_split_reductions_on_128_tiles(prec, tol):
    """norb 12 at M = 384: merged sectors of 100..190 on both sides, on 128x128 tiles from 96 rows on (the FP64 and
    FP32 defaults start at 192, so a child process sets MPSXX_BIG_TILE_MIN(_F64) = 96; read once per process); the
    second product's long reductions are split over work-groups (partial 128x128 tiles summed in slot order)."""
    print("128 tiles merged qc", prec, flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MPSXX_BIG_TILE_MIN="96", MPSXX_BIG_TILE_MIN_F64="96")
    r = subprocess.run([sys.executable, "-c", _MERGED_CHILD, root, str(prec)], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    same, rel, big, red = json.loads(r.stdout.strip().splitlines()[-1])
    print("128 tiles", big, "split", red, "rel", rel, flush=True)
    assert big > 0 and red > 0
    assert same
    assert rel <= tol


_F64_CHILD = r"""
import json, sys
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/tests"]
import numpy as np
import mpsxx_amd as mx
import oracle
from helpers import random_twosite_problem
from mpsxx_amd import workloads
from mpsxx_amd.batch import last_plan_tiles
rng = np.random.default_rng(31)
lm, rm, le, re_, psi = random_twosite_problem(10, 4, rng, lambda q: int(rng.integers(100, 231)) if abs(q[1]) <= 1 else 0)
out = {}
ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
d = [mx.DTensor(x) for x in (lm, rm, le, re_, psi)]
got = mx.compute_sigma_vector_twosite(*d).to_qst()
scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size)
out["two_site"] = [got.block_set() == ref.block_set(), got.max_abs_diff(ref) / scale, last_plan_tiles()["big128"]]
Lm, Rm, psim, info = workloads.qc_merged_problem(12, 384, seed=5)
ref = oracle.sigma_merged(Lm.to_qst(), Rm.to_qst(), psim.to_qst())
got = mx.compute_sigma_vector_merged(Lm, Rm, psim).to_qst()
scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size)
t = last_plan_tiles()
out["merged"] = [got.block_set() == ref.block_set(), got.max_abs_diff(ref) / scale, t["big128"], t["big128_red"]]
print(json.dumps(out), flush=True)
"""


def test_fp64_64_tiles_when_128_tiles_are_off():
    """FP64 runs its large blocks (both sides >= 192) on gemm128_kernel<double> by default; MPSXX_BIG_TILE_MIN_F64=0 (read once per
    process, so in a child) keeps them on 64x64 tiles: same bar, no 128x128 tile."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MPSXX_BIG_TILE_MIN_F64="0")
    print("128 tiles fp64 child", flush=True)
    r = subprocess.run([sys.executable, "-c", _F64_CHILD, root], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res, flush=True)
    assert res["two_site"][0] and res["two_site"][1] <= 1e-12 and res["two_site"][2] == 0
    assert res["merged"][0] and res["merged"][1] <= 1e-12 and res["merged"][2] == 0 and res["merged"][3] == 0
