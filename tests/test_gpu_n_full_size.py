"""The benchmark configurations at their full sizes, through properties that do not need the oracle at that size
(needs an MI355X; the oracle pins the same code paths at reduced sizes in the other GPU tests).

* Config 3 (quantum chemistry, norb 24, M 1024, D_w 1226, the merged sigma of the headline): FP32 within 2e-5 and
  split-BF16x3 within 1e-4 of FP64 (relative Frobenius norm), identical block sets in every precision, and linearity
  sigma(a x + b y) = a sigma(x) + b sigma(y) in FP64 to 1e-12 relative.
* Config 2 (Hubbard L = 32, M 128x3, 256x3, 512x2): every sweep's energy at most the previous one's (1e-9 slack) and
  FP32-mixed sweeps within 1e-8 of FP64.
* Config 4 (64 chains, L 16..64, M 128..512): the batched sigma of all 64 instances equals the single-instance path
  bit for bit under the default tile policies (tuned policies may cut reductions differently, which changes the
  summation order), and its FP32 form stays within 2e-5."""
import numpy as np
import pytest

import mpsxx_amd as mx
from mpsxx_amd import workloads
from mpsxx_amd.batch import SigmaBatch
from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
from mpsxx_amd.mpo import hubbard_mpo

pytestmark = pytest.mark.gpu


def test_config3_merged_sigma_full_size():
    print("config 3 merged sigma norb 24 M 1024", flush=True)
    Lm, Rm, psim, info = workloads.qc_merged_problem(24, 1024, 1)
    assert info["D_w"] == 1226
    ref = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=mx.FP64).to_qst()
    for prec, tol in ((mx.FP32, 2e-5), (mx.BF16X3, 1e-4)):
        got = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=prec).to_qst()
        assert got.block_set() == ref.block_set()
        rel = got.diff_norm(ref) / ref.norm()
        print("precision", prec, "relative difference", rel, flush=True)
        assert rel < tol, (prec, rel)
    # linearity in FP64: sigma(2 x - 0.5 y) = 2 sigma(x) - 0.5 sigma(y), y = a second random fill of x's blocks
    x = psim.to_qst()
    rng = np.random.default_rng(11)
    y = x.empty_like()
    for k, b in x.blocks.items():
        y.blocks[k] = rng.standard_normal(b.shape)
    comb = x.empty_like()
    for k in x.blocks:
        comb.blocks[k] = 2.0 * x.blocks[k] - 0.5 * y.blocks[k]
    sy = mx.compute_sigma_vector_merged(Lm, Rm, mx.DTensor(y), precision=mx.FP64).to_qst()
    sc = mx.compute_sigma_vector_merged(Lm, Rm, mx.DTensor(comb), precision=mx.FP64).to_qst()
    expect = ref.empty_like()
    for k in ref.blocks:
        expect.blocks[k] = 2.0 * ref.blocks[k] - 0.5 * sy.blocks[k]
    assert sc.block_set() == expect.block_set()
    assert sc.diff_norm(expect) <= 1e-12 * expect.norm()


def test_config2_sweeps_full_size():
    print("config 2 sweeps L 32 M 128-512", flush=True)
    sched = [128] * 3 + [256] * 3 + [512] * 2
    final = {}
    for prec in (mx.FP64, mx.FP32):
        dm = DeviceDMRG(hubbard_mpo(32, 1.0, 4.0), initial_mps(32, (32, 0), sched[0], 0), sched[0])
        es = []
        for M in sched:
            es.append(dm.sweep(M, 1e-10, precision=prec)[0])
        print("precision", prec, es, flush=True)
        assert all(b <= a + 1e-9 for a, b in zip(es, es[1:])), es
        final[prec] = es[-1]
        del dm
    assert abs(final[mx.FP32] - final[mx.FP64]) <= 1e-8, final


def test_config4_batched_sigma_full_size():
    print("config 4 batched sigma, 64 chains", flush=True)
    specs = workloads.batch_specs(64, 4)
    probs = [workloads.hubbard_twosite(L=L, M=M, U=U, seed=k) for k, (L, U, M) in enumerate(specs)]
    dev = [tuple(mx.DTensor(x) for x in p) for p in probs]
    ref = [mx.compute_sigma_vector_twosite(*d).to_qst() for d in dev]
    for prec in (mx.FP64, mx.FP32):
        inst = [d + (mx.DTensor(),) for d in dev]
        b = SigmaBatch(inst, precision=prec)
        b.run(use_graph=True)
        worst = 0.0
        for r, t in zip(ref, inst):
            got = t[5].to_qst()
            assert got.block_set() == r.block_set()
            worst = max(worst, got.diff_norm(r) / r.norm())
        print("precision", prec, "largest relative difference", worst, flush=True)
        assert worst == 0.0 if prec == mx.FP64 else worst < 2e-5, worst
        b.free()
