"""Quantum-chemistry MPO on the device path (needs an MI355X).

* sigma two-site with the QC MPO at the middle of a 10-orbital chain (D_w = 232 on both bonds) against the oracle;
* device two-site sweeps with the QC MPO of random 6-orbital integrals against the oracle's sweeps and the FCI energy
  (exact diagonalization, sector dimension 400);
* Hubbard L=12, U=1 read from an FCIDUMP (tests/golden/hubbard12.FCIDUMP, the reference's dmrg/tests/HUBBARD12 model)
  through the QC MPO: E0 = -11.8406372775523305 (dmrg/tests/hub12_dmrg.out:15536)."""
import os

import numpy as np
import pytest

import ed
import mpsxx_amd as mx
import oracle
from helpers import random_env_left, random_env_right
from mpsxx_amd import mps, qc
from mpsxx_amd.dmrg import DeviceDMRG, initial_mps

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sigma_twosite_with_qc_mpo_matches_oracle():
    print("qc sigma K=10", flush=True)
    K, i = 10, 4
    rng = np.random.default_rng(21)
    h, v = qc.synthetic_integrals(K, seed=4)
    m = qc.QcMpo(h, v)
    W1, W2 = m.site_qst(i), m.site_qst(i + 1)
    assert len(W1.qshape[0]) == m.bond_dims()[i] and len(W2.qshape[3]) == m.bond_dims()[i + 2]
    qt = (K, 0)
    lq = mps.left_bond_qns(K, i, qt)
    rq = mps.right_bond_qns(K, i + 2, qt)
    ld = [int(rng.integers(1, 6)) if abs(q[1]) <= 2 else 0 for q in lq]
    rd = [int(rng.integers(1, 6)) if abs(q[1]) <= 2 else 0 for q in rq]
    lenv = random_env_left(lq, ld, W1, rng)
    renv = random_env_right(rq, rd, W2, rng)
    psi = mps.fill_allowed(mps.twosite_wavefunction(lq, ld, rq, rd, qt), rng)
    ref = oracle.sigma_twosite(W1, W2, lenv, renv, psi)
    got = mx.compute_sigma_vector_twosite(*[mx.DTensor(x) for x in (W1, W2, lenv, renv, psi)]).to_qst()
    assert got.block_set() == ref.block_set()
    scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size)
    assert got.max_abs_diff(ref) <= 1e-12 * scale


def test_device_dmrg_with_qc_mpo_matches_oracle_and_fci():
    print("qc dmrg K=6", flush=True)
    K, M = 6, 64
    h, v = qc.synthetic_integrals(K, seed=5, nvec=8, hsig=0.2, lsig=0.25)
    e_fci = ed.ground_energy(h, v, K, K, 0)
    mpos = qc.QcMpo(h, v).qsts()
    init = initial_mps(K, (K, 0), M, 0)
    es, _ = oracle.dmrg(mpos, init, [M] * 4, 1e-10)
    dm = DeviceDMRG(mpos, init, M)
    e = None
    for _ in range(4):
        e, st = dm.sweep(M, 1e-10)
    assert abs(e - es[-1]) < 1e-10, (e, es)
    assert abs(e - e_fci) < 1e-10, (e, e_fci)


@pytest.mark.parametrize("prec", [mx.FP64, mx.FP32, mx.BF16X3])
def test_device_dmrg_merged_form_matches_fourindex_and_fci(prec):
    """Merged two-site sweeps (src/make_sweep.cpp:112, optimize_twosite_merged): merged operators, merged sigma and
    diag, SVD of the merged wavefunction, merged renormalization.  Same inputs as the 4-index test above: the final
    energy equals the oracle's 4-index sweeps and FCI to 1e-10 (FP64) / 1e-8 (mixed precision, FP64 refinement)."""
    from mpsxx_amd.dmrg import SWEEP_MERGED
    print("qc dmrg merged K=6 prec", prec, flush=True)
    K, M = 6, 64
    h, v = qc.synthetic_integrals(K, seed=5, nvec=8, hsig=0.2, lsig=0.25)
    e_fci = ed.ground_energy(h, v, K, K, 0)
    m = qc.QcMpo(h, v)
    init = initial_mps(K, (K, 0), M, 0)
    dm = DeviceDMRG(m.devices(), init, M, form=SWEEP_MERGED)
    e = None
    nlow = 0
    for _ in range(4):
        e, st = dm.sweep(M, 1e-10, precision=prec)
        nlow += st["nsigma_fp32"]
    tol = 1e-10 if prec == mx.FP64 else 1e-8
    assert abs(e - e_fci) < tol, (e, e_fci)
    if prec != mx.FP64:
        # low-precision sigma vectors in the early sweeps; a converged site starts in FP64 and stays there
        assert nlow > 0


def test_device_dmrg_merged_form_hubbard_matches_fourindex():
    """Hubbard L=8, U=4, M=64 (the compact MPO): merged and 4-index sweeps reach the same energy (ED to 1e-9)."""
    from mpsxx_amd.dmrg import SWEEP_MERGED
    from mpsxx_amd.mpo import hubbard_mpo
    print("hubbard merged L=8", flush=True)
    L, M = 8, 64
    mpos = hubbard_mpo(L, 1.0, 4.0)
    init = initial_mps(L, (L, 0), M, 0)
    es = []
    for form in (0, SWEEP_MERGED):
        dm = DeviceDMRG(mpos, init, M, form=form)
        for _ in range(4):
            e, _ = dm.sweep(M, 1e-10)
        es.append(e)
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    e0 = ed.ground_energy(h, v, L, L, 0)
    assert abs(es[0] - es[1]) < 1e-10, es
    assert abs(es[1] - e0) < 1e-9, (es, e0)


def test_hubbard12_from_fcidump_through_qc_mpo_matches_reference_log():
    print("hub12 fcidump", flush=True)
    norb, nelec, ms2, h, v, ecore = qc.read_fcidump(os.path.join(HERE, "golden", "hubbard12.FCIDUMP"))
    assert (norb, nelec, ms2) == (12, 12, 0)
    mpos = qc.QcMpo(h, v, ecore).qsts()
    M = 200
    dm = DeviceDMRG(mpos, initial_mps(norb, (nelec, ms2), M, 0), M)
    e_prev, e = None, None
    for sweep in range(8):
        e, st = dm.sweep(M, 1e-8)
        print("sweep", sweep, e, st["seconds"], st["max_bond"], flush=True)
        if e_prev is not None and abs(e - e_prev) < 1e-10:
            break
        e_prev = e
    assert abs(e - (-11.8406372775523305)) < 2e-7, e


@pytest.mark.parametrize("norb,M", [(8, 48), (12, 128)])
def test_merged_sigma_qc_matches_oracle(norb, M):
    """Merged-form sigma (src/compute_sigma_vector.hpp:86-96) with the merged QC operators at the middle bond: the
    second GEMM reduces over every (bond operator, right sector) pair of an output block, which the planner splits
    over several work-groups (partial tiles summed in a fixed order)."""
    from mpsxx_amd import workloads
    print("qc merged sigma", norb, M, flush=True)
    Lm, Rm, psim, info = workloads.qc_merged_problem(norb, M, seed=3)
    Lq, Rq, pq = Lm.to_qst(), Rm.to_qst(), psim.to_qst()
    ref = oracle.sigma_merged(Lq, Rq, pq)
    got = mx.compute_sigma_vector_merged(Lm, Rm, psim).to_qst()
    assert got.block_set() == ref.block_set()
    scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size)
    assert got.max_abs_diff(ref) <= 1e-12 * scale
    got32 = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=mx.FP32).to_qst()
    assert got32.diff_norm(ref) <= 2e-5 * ref.norm()


_SEGMENTED_CHILD = r"""
import json, sys
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/tests"]
import numpy as np
import mpsxx_amd as mx
import oracle
from mpsxx_amd import workloads
Lm, Rm, psim, info = workloads.qc_merged_problem(8, 48, seed=3)
ref = oracle.sigma_merged(Lm.to_qst(), Rm.to_qst(), psim.to_qst())
scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size)
out = {}
from mpsxx_amd.batch import last_plan_tiles
for name, prec in (("fp64", mx.FP64), ("fp32", mx.FP32)):
    got = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=prec).to_qst()
    out[name] = [got.block_set() == ref.block_set(), got.max_abs_diff(ref) / scale, got.diff_norm(ref) / ref.norm(),
                 last_plan_tiles()["split_cut"]]
print(json.dumps(out), flush=True)
"""


def test_merged_sigma_qc_with_every_split_reduction_cut_matches_oracle():
    """The 32x32 tiles with long reductions cut over work-groups (partial tiles summed in slot order by
    gemm_reduce32_kernel): a child process with a policy that cuts every such tile into single-unit pair ranges
    (MPSXX_SPLIT_SEG=1, MPSXX_SPLIT_SMALL=100000) computes the merged QC sigma; same bars as the default policy."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MPSXX_SPLIT_SEG="1", MPSXX_SPLIT_SMALL="100000")
    print("qc merged sigma, every split reduction cut", flush=True)
    r = subprocess.run([sys.executable, "-c", _SEGMENTED_CHILD, root], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res, flush=True)
    assert res["fp64"][0] and res["fp32"][0]
    assert res["fp64"][3] > 0 and res["fp32"][3] > 0  # the cut path really ran
    assert res["fp64"][1] <= 1e-12
    assert res["fp32"][2] <= 2e-5


def test_merged_sigma_qc_split_bf16_close_to_fp64():
    """MPSXX_BF16X3: the large blocks multiply as split-BF16 pairs (three bf16 MFMAs per product, ~2^-16 relative per
    product); bar 1e-4 relative Frobenius norm against the FP64 device result (itself pinned to the oracle above)."""
    from mpsxx_amd import workloads
    print("qc merged sigma bf16x3", flush=True)
    Lm, Rm, psim, info = workloads.qc_merged_problem(12, 128, seed=3)
    ref = mx.compute_sigma_vector_merged(Lm, Rm, psim).to_qst()
    got = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=mx.BF16X3).to_qst()
    assert got.block_set() == ref.block_set()
    rel = got.diff_norm(ref) / ref.norm()
    print("bf16x3 rel", rel, flush=True)
    assert rel <= 1e-4, rel
    got32 = mx.compute_sigma_vector_merged(Lm, Rm, psim, precision=mx.FP32).to_qst()
    assert rel >= 0.5 * got32.diff_norm(ref) / ref.norm()  # not more accurate than FP32: the split path really ran


def test_hubbard12_excited_states_match_reference_log():
    """Four roots of the reference's Hubbard-12 run (dmrg/tests/twodot.conf: M = 200, nroots 4) from its FCIDUMP
    through the QC MPO, each root projected against the converged lower ones: the energies in
    dmrg/tests/hub12_dmrg.out:15536-15539 to 2e-6 (the reference adds one-site sweeps after the two-site ones)."""
    print("hub12 excited states", flush=True)
    ref = [-11.8406372775523305, -11.4516084815790968, -11.2058172994693876, -11.0278375085199389]
    norb, nelec, ms2, h, v, ecore = qc.read_fcidump(os.path.join(HERE, "golden", "hubbard12.FCIDUMP"))
    mpos = qc.QcMpo(h, v, ecore).devices()
    M = 200
    done, es = [], []
    for k in range(4):
        dm = DeviceDMRG(mpos, initial_mps(norb, (nelec, ms2), M, k), M)
        for r, e_r in zip(done, es):
            dm.add_lower_root(r, e_r)
        e_prev = None
        for sweep in range(12):
            e, st = dm.sweep(M, 1e-8)
            if e_prev is not None and abs(e - e_prev) < 1e-10:
                break
            e_prev = e
        print("root", k, e, "sweeps", sweep + 1, flush=True)
        done.append(dm)
        es.append(e)
    assert np.max(np.abs(np.array(es) - np.array(ref))) < 2e-6, (es, ref)


# Molpro FCI of the reference's C12H14 pi-system FCIDUMP (mpsxx/qcdmrg/benchmark/FCIDUMP; the same integrals as
# dmrg/tests/FCIDUMP12 up to header formatting), total energy with E_core: mpsxx/qcdmrg/benchmark/molpro/
# c12h14.fci.out:979.  Orbital order of the reference's run: mpsxx/qcdmrg/benchmark/reorder.dat (1-based).
E_FCI_C12H14 = -457.052106408795
C12H14_ORDER = [6, 9, 5, 11, 12, 4, 3, 7, 8, 2, 10, 1]


def test_c12h14_merged_sweeps_reach_molpro_fci():
    """The reference's own molecule: the C12H14 FCIDUMP read as the reference reads it (twoint(p,r,q,s) = (pq|rs),
    src/parsing_integral.cpp:139-147) pins the two-electron convention of the FCIDUMP -> QC MPO path, which the
    Hubbard FCIDUMP (only (ii|ii) and hopping terms) cannot.  Merged two-site FP64 sweeps in the reference's orbital
    order, M 250 -> 500 -> 1000 -> 2000 (the centre bond of 12 orbitals holds at most 4^6 = 4096 states), reach the
    Molpro FCI energy to 1e-6; FP32 sigma vectors with FP64 refinement land within 1e-8 of the FP64 sweeps."""
    from mpsxx_amd.dmrg import SWEEP_MERGED
    print("c12h14 fci", flush=True)
    norb, nelec, ms2, h, v, ecore = qc.read_fcidump(os.path.join(HERE, "golden", "c12h14.FCIDUMP"))
    assert (norb, nelec, ms2) == (12, 12, 0) and abs(ecore - -0.4422645349012726e3) < 1e-12
    o = [k - 1 for k in C12H14_ORDER]
    h = h[np.ix_(o, o)]
    v = v[np.ix_(o, o, o, o)]
    mpos = qc.QcMpo(h, v, ecore).devices()
    sched = [250, 500, 1000, 2000, 2000, 2000, 2000]
    res = {}
    for prec in (mx.FP64, mx.FP32):
        dm = DeviceDMRG(mpos, initial_mps(norb, (nelec, ms2), sched[0], 0), sched[0], form=SWEEP_MERGED)
        e_prev = None
        for k, M in enumerate(sched):
            e, st = dm.sweep(M, 1e-9, prec)
            print("prec", prec, "sweep", k, "M", M, e, st["seconds"], flush=True)
            if k >= 4 and e_prev is not None and abs(e - e_prev) < 1e-10:
                break
            e_prev = e
        res[prec] = e
    assert abs(res[mx.FP64] - E_FCI_C12H14) < 1e-6, res
    assert abs(res[mx.FP32] - res[mx.FP64]) < 1e-8, res
