"""Sigma plans replayed from hipGraphs inside the Davidson solvers (mpsxx_sigma_graphs) against host-issued launches
(needs an MI355X).

The replayed graph runs the same kernels with their operand bases read from the plan's pointer table, so every
result must equal the host-launched one bit for bit: Davidson solves in every precision (each replaying its sigma
graphs tens of times with eager vector work in between), and whole device sweeps in the 4-index and merged forms,
where every site's plans are captured and replayed again in later sweeps."""
import ctypes

import numpy as np
import pytest

import mpsxx_amd as mx
from mpsxx_amd import qc, workloads
from mpsxx_amd.dmrg import SWEEP_MERGED, DeviceDMRG, initial_mps
from mpsxx_amd.mpo import hubbard_mpo

pytestmark = pytest.mark.gpu


def graphs(on):
    return mx.lib().mpsxx_sigma_graphs(ctypes.c_int(1 if on else 0))


@pytest.fixture(autouse=True)
def restore_graph_setting():
    prev = graphs(True)
    yield
    graphs(prev)


@pytest.mark.parametrize("prec", [mx.FP64, mx.FP32, mx.BF16X3])
def test_davidson_graph_replay_equals_host_launches(prec):
    print("sigma graph davidson", prec, flush=True)
    ps = [workloads.hubbard_isometric_problem(L, M, U, seed=k)
          for k, (L, M, U) in enumerate([(12, 48, 4.0), (16, 96, 2.0)])]
    for p in ps:
        res = []
        for on in (False, True, True):  # the second graph run replays the graphs the first one captured
            graphs(on)
            w = mx.copy(p[5])
            e, ns = mx.davidson_twosite([p[:4]], p[4], w, 1e-12, prec)
            res.append((e, ns, w.to_qst()))
        for e, ns, w in res[1:]:
            assert e == res[0][0] and ns == res[0][1]
            assert w.block_set() == res[0][2].block_set() and w.max_abs_diff(res[0][2]) == 0.0


def test_sweeps_with_graphs_equal_sweeps_without():
    print("sigma graph sweeps", flush=True)
    out = {}
    for on in (False, True):
        graphs(on)
        energies = []
        dm = DeviceDMRG(hubbard_mpo(10, 1.0, 4.0), initial_mps(10, (10, 0), 32, 0), 32)
        for M in (32, 64, 64):
            energies.append(dm.sweep(M, 1e-10)[0])
        h, v = qc.synthetic_integrals(6, seed=5)
        mpos = qc.QcMpo(h, v).devices()
        dq = DeviceDMRG(mpos, initial_mps(6, (6, 0), 64, 1), 64, form=SWEEP_MERGED)
        for prec in (mx.FP64, mx.FP32):
            energies.append(dq.sweep(64, 1e-10, precision=prec)[0])
        out[on] = np.array(energies)
    assert np.array_equal(out[False], out[True]), out
