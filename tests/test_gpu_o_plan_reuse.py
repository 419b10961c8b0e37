"""Low-precision sigma plans made from the cached FP64 plan of the same structures (needs an MI355X).

A mixed-precision Davidson solve builds its FP64 sigma plans first; its FP32 and split-BF16x3 plans are then copies of
those (same task and tile lists, FP32 storage for inputs and intermediates) instead of new symbolic contractions.  The
copies must give exactly the sigma vectors of plans built from scratch: the reference values come from a child
process with MPSXX_NO_PLAN_REUSE set (read once per process), compared bit for bit, for the two-site and merged forms (the one-site form
goes through the same function, plan_from_fp64)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import mpsxx_amd as mx
from helpers import hubbard_problem
from mpsxx_amd import workloads

pytestmark = pytest.mark.gpu

_CHILD = r"""
import json, sys
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/tests"]
import numpy as np
import mpsxx_amd as mx
from test_gpu_o_plan_reuse import problems
out = {}
for name, (kind, args) in problems().items():
    for prec in (mx.FP32, mx.BF16X3):
        f = {"two": mx.compute_sigma_vector_twosite, "merged": mx.compute_sigma_vector_merged}[kind]
        q = f(*args, precision=prec).to_qst()
        out["%s:%d" % (name, prec)] = {repr(k): v.ravel().tolist() for k, v in q.blocks.items()}
print(json.dumps(out), flush=True)
"""


def problems():
    rng = np.random.default_rng(21)
    L, i = 10, 4
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=6)
    two = tuple(mx.DTensor(x) for x in (mpos[i], mpos[i + 1], lenv, renv, ch["psi"]))
    Lm, Rm, psim, _ = workloads.qc_merged_problem(8, 96, seed=4)
    return {"two": ("two", two), "merged": ("merged", (Lm, Rm, psim))}


def test_low_precision_plans_from_fp64_plans_equal_fresh_plans():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MPSXX_NO_PLAN_REUSE="1")
    print("plan reuse: fresh plans in a child", flush=True)
    r = subprocess.run([sys.executable, "-c", _CHILD, root], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = json.loads(r.stdout.strip().splitlines()[-1])
    for name, (kind, args) in problems().items():
        f = {"two": mx.compute_sigma_vector_twosite, "merged": mx.compute_sigma_vector_merged}[kind]
        f(*args, precision=mx.FP64)  # builds the FP64 plan
        n64 = mx.plan_cache_stats()["builds"]
        for prec in (mx.FP32, mx.BF16X3):
            q = f(*args, precision=prec).to_qst()
            want = ref["%s:%d" % (name, prec)]
            assert sorted(repr(k) for k in q.blocks) == sorted(want)
            for k, v in q.blocks.items():
                assert np.array_equal(v.ravel(), np.array(want[repr(k)])), (name, prec, k)
        print(name, "plans built after the FP64 one", mx.plan_cache_stats()["builds"] - n64, flush=True)
