"""The plan cache under a small device-memory budget (needs an MI355X).

With MPSXX_PLAN_CACHE_GB far below what a sweep's plans hold, least recently used plans release their intermediate
buffers (allocated again on their next use) and, once their task lists alone exceed the budget, are evicted and
rebuilt; the 16 most recent plans are kept.  None of this may change a result: QC merged sweeps (FP64 and FP32
mixed) and four-index Hubbard sweeps in a child process with a 4 MB budget give the energies of the same sweeps
under the default budget bit for bit, and the child's counters show that plans were released, restored and
evicted.  The budget is read once per process, hence the child."""
import json
import os
import subprocess
import sys

import pytest

import mpsxx_amd as mx

pytestmark = pytest.mark.gpu

_RUN = r"""
import json, sys
sys.path[:0] = [sys.argv[1]]
import mpsxx_amd as mx
from mpsxx_amd import qc
from mpsxx_amd.dmrg import SWEEP_MERGED, DeviceDMRG, initial_mps
from mpsxx_amd.mpo import hubbard_mpo
energies = []
h, v = qc.synthetic_integrals(10, seed=3)
mpos = qc.QcMpo(h, v).devices()
for prec in (mx.FP64, mx.FP32):
    dq = DeviceDMRG(mpos, initial_mps(10, (10, 0), 64, 2), 64, form=SWEEP_MERGED)
    for M in (64, 128, 128):
        energies.append(dq.sweep(M, 1e-10, precision=prec)[0])
    del dq
dm = DeviceDMRG(hubbard_mpo(12, 1.0, 4.0), initial_mps(12, (12, 0), 32, 0), 32)
for M in (32, 64, 64):
    energies.append(dm.sweep(M, 1e-10)[0])
print(json.dumps({"energies": energies, "stats": mx.plan_cache_stats()}), flush=True)
"""


def _run(env):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _RUN, root], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_small_plan_budget_gives_the_same_energies():
    print("plan budget: default", flush=True)
    env = {k: v for k, v in os.environ.items() if k != "MPSXX_PLAN_CACHE_GB"}
    ref = _run(env)
    print("plan budget: 4 MB", flush=True)
    small = _run(dict(env, MPSXX_PLAN_CACHE_GB="0.004"))
    print("default", ref["stats"], "\nsmall", small["stats"], flush=True)
    assert small["energies"] == ref["energies"]
    st = small["stats"]
    assert st["releases"] > 0 and st["restores"] > 0 and st["evictions"] > 0
    assert ref["stats"]["evictions"] == 0


def test_release_caches_keeps_results_and_frees_the_pool():
    """mpsxx_release_caches between sweeps: no plan and no pooled buffer is left, and the next sweeps (which rebuild
    their plans) give the energies of the same sweeps run without it, bit for bit."""
    from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
    from mpsxx_amd.mpo import hubbard_mpo
    print("release caches", flush=True)
    out = []
    for release in (False, True):
        dm = DeviceDMRG(hubbard_mpo(12, 1.0, 4.0), initial_mps(12, (12, 0), 32, 0), 32)
        es = [dm.sweep(32, 1e-10)[0]]
        if release:
            mx.release_caches()
            assert mx.plan_cache_stats()["plans"] == 0
            assert mx.device_memory_stats()["cached_bytes"] == 0
        for M in (64, 64):
            es.append(dm.sweep(M, 1e-10, precision=mx.FP32)[0])
        out.append(es)
        del dm
    assert out[0] == out[1], out
