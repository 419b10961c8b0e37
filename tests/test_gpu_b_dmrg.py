"""Device canonicalization, merged operators and two-site sweeps vs the oracle (needs an MI355X)."""
import ctypes

import numpy as np
import pytest

import ed
import mpsxx_amd as mx
import oracle
from mpsxx_amd import mps
from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
from mpsxx_amd.mpo import hubbard_mpo

pytestmark = pytest.mark.gpu


def _canon(forward, psi, M):
    U, V, P = mx.DTensor(), mx.DTensor(), mx.DTensor(psi)  # handles must outlive the call
    d = ctypes.c_double()
    mx._check(mx.lib().mpsxx_canonicalize_twosite(ctypes.c_int(1 if forward else 0), P.h, U.h, V.h,
                                                  ctypes.c_int(M), ctypes.byref(d), None))
    return U.to_qst(), V.to_qst(), d.value


@pytest.mark.parametrize("forward,M", [(True, 0), (False, 0), (True, 12), (False, 12)])
def test_canonicalize_matches_oracle(forward, M):
    print("canonicalize", forward, M, flush=True)
    rng = np.random.default_rng(5)
    ch = mps.random_chain(8, 3, (8, 0), rng, dmax=5)
    psi = ch["psi"]
    U, V, dw = _canon(forward, psi, M)
    Ur, Vr, dwr = oracle.canonicalize_twosite(forward, psi, M)
    # same kept sectors and dimensions, same discarded weight, same reconstructed (truncated) wavefunction
    assert U.qshape == Ur.qshape and U.dshape == Ur.dshape and U.block_set() == Ur.block_set()
    assert V.qshape == Vr.qshape and V.dshape == Vr.dshape and V.block_set() == Vr.block_set()
    assert abs(dw - dwr) < 1e-12 * max(1.0, psi.norm() ** 2)
    rec = oracle.contract(1.0, U, [2], V, [0])
    recr = oracle.contract(1.0, Ur, [2], Vr, [0])
    scale = max(float(np.max(np.abs(b))) for b in psi.blocks.values())
    assert rec.max_abs_diff(recr) < 1e-11 * scale


@pytest.mark.parametrize("solver", ["jacobi", "jacobi_unsorted", "jacobi_noqr", "rocsolver"])
@pytest.mark.parametrize("forward,M", [(True, 0), (False, 0), (True, 150), (False, 150)])
def test_canonicalize_large_blocks_matches_oracle(forward, M, solver, monkeypatch):
    """Merged blocks of up to ~150 vectors (several column groups per Jacobi task, both orientations m < n and
    m > n, a graded spectrum from a rescaled psi) against the oracle's per-block SVD, for the default batched Jacobi
    path (vectors sorted by norm, then QR-preconditioned), the same without the sort (MPSXX_SVD_SORT=0), the Jacobi
    sweeps without the QR step (MPSXX_SVD=jacobi_noqr, the path of blocks too long for the QR kernels) and the
    rocSOLVER path (MPSXX_SVD=rocsolver)."""
    print("canonicalize large", forward, M, solver, flush=True)
    if solver == "jacobi_unsorted":
        monkeypatch.setenv("MPSXX_SVD_SORT", "0")
    else:
        monkeypatch.setenv("MPSXX_SVD", solver)
    rng = np.random.default_rng(11)
    ch = mps.random_chain(8, 3, (8, 0), rng, dmax=40)
    psi = ch["psi"]
    # grade the spectrum: scale the entries of each block by a decaying factor of its row position
    for k, b in psi.blocks.items():
        psi.blocks[k] = b * np.exp(-0.05 * np.arange(b.shape[0]))[:, None, None, None]
    U, V, dw = _canon(forward, psi, M)
    Ur, Vr, dwr = oracle.canonicalize_twosite(forward, psi, M)
    assert U.qshape == Ur.qshape and U.dshape == Ur.dshape and U.block_set() == Ur.block_set()
    assert V.qshape == Vr.qshape and V.dshape == Vr.dshape and V.block_set() == Vr.block_set()
    assert abs(dw - dwr) < 1e-12 * max(1.0, psi.norm() ** 2)
    rec = oracle.contract(1.0, U, [2], V, [0])
    recr = oracle.contract(1.0, Ur, [2], Vr, [0])
    scale = max(float(np.max(np.abs(b))) for b in psi.blocks.values())
    assert rec.max_abs_diff(recr) < 1e-11 * scale
    # the orthonormal factor: U^T U = 1 (forward) or V V^T = 1 (backward) on every kept sector
    if forward:
        g = oracle.contract(1.0, U, [0, 1], U, [0, 1], conj_a=True)
    else:
        g = oracle.contract(1.0, V, [1, 2], V, [1, 2], conj_b=True)
    for (i, j), blk in g.blocks.items():
        assert np.max(np.abs(blk - (np.eye(blk.shape[0]) if i == j else 0.0))) < 1e-12


@pytest.mark.parametrize("solver", ["jacobi", "rocsolver"])
def test_canonicalize_rank_deficient_keeps_exact_rank(solver, monkeypatch):
    """psi = A(l,n1,m) B(m,n2,r) through a middle bond of 2 states per sector has merged blocks of rank <= 2 but
    dimensions up to ~30: with M = 0 (keep every value >= 1e-16) the device keeps exactly the numerical rank (values
    below max(m,n) eps |A|_F are reported as zero, not kept as noise directions), discards no weight, and U s Vt
    reconstructs psi."""
    print("canonicalize rank-deficient", solver, flush=True)
    monkeypatch.setenv("MPSXX_SVD", solver)
    rng = np.random.default_rng(21)
    L, i, qt = 8, 3, (8, 0)
    lq, mq, rq = mps.left_bond_qns(L, i, qt), mps.left_bond_qns(L, i + 1, qt), mps.right_bond_qns(L, i + 2, qt)
    A = mps.fill_allowed(mps.site_tensor(lq, [8] * len(lq), mq, [2] * len(mq)), rng)
    B = mps.fill_allowed(mps.site_tensor(mq, [2] * len(mq), rq, [8] * len(rq), total=qt), rng)
    psi = oracle.contract(1.0, A, [2], B, [0])
    for forward in (True, False):
        U, V, dw = _canon(forward, psi, 0)
        kept = U.dshape[-1]
        assert max(kept) <= 2, kept
        assert abs(dw) < 1e-20 * max(1.0, psi.norm() ** 2)
        rec = oracle.contract(1.0, U, [2], V, [0])
        scale = max(float(np.max(np.abs(b))) for b in psi.blocks.values())
        assert rec.max_abs_diff(psi) < 1e-11 * scale


def test_merged_operator_matches_oracle():
    rng = np.random.default_rng(6)
    L, i = 8, 3
    mpos = hubbard_mpo(L, 1.0, 4.0)
    qt = (L, 0)
    lq = mps.left_bond_qns(L, i, qt)
    ld = [int(rng.integers(1, 8)) for _ in lq]
    from helpers import random_env_left, random_env_right
    le = random_env_left(lq, ld, mpos[i], rng)
    lib = mx.lib()
    out, dw, dl = mx.DTensor(), mx.DTensor(mpos[i]), mx.DTensor(le)
    mx._check(lib.mpsxx_compute_merged_operator(ctypes.c_int(1), dw.h, dl.h, out.h, None))
    got = out.to_qst()
    o, ow, ol = oracle.OTensor(), oracle.OTensor(mpos[i]), oracle.OTensor(le)
    oracle._run(oracle.lib().orc_merged_operator, ctypes.c_int(1), ow.h, ol.h, o.h)
    ref = o.to_qst()
    assert got.qshape == ref.qshape and got.dshape == ref.dshape and got.block_set() == ref.block_set()
    assert got.max_abs_diff(ref) < 1e-13
    rq = mps.right_bond_qns(L, i + 1, qt)
    rd = [int(rng.integers(1, 8)) for _ in rq]
    re_ = random_env_right(rq, rd, mpos[i], rng)
    out2, dr = mx.DTensor(), mx.DTensor(re_)
    mx._check(lib.mpsxx_compute_merged_operator(ctypes.c_int(0), dw.h, dr.h, out2.h, None))
    got2 = out2.to_qst()
    o2, orr = oracle.OTensor(), oracle.OTensor(re_)
    oracle._run(oracle.lib().orc_merged_operator, ctypes.c_int(0), ow.h, orr.h, o2.h)
    ref2 = o2.to_qst()
    assert got2.qshape == ref2.qshape and got2.block_set() == ref2.block_set()
    assert got2.max_abs_diff(ref2) < 1e-13


def test_device_dmrg_config1_matches_oracle_and_ed():
    """Config 1: L=8, t=1, U=4, N=8, Sz=0, M=64, 4 sweeps, tol 1e-10; |dE| GPU vs CPU <= 1e-10."""
    print("dmrg config1", flush=True)
    L, M = 8, 64
    mpos = hubbard_mpo(L, 1.0, 4.0)
    init = initial_mps(L, (L, 0), M, 0)
    dm = DeviceDMRG(mpos, init, M)
    eg = []
    for _ in range(4):
        e, st = dm.sweep(M, 1e-10)
        eg.append(e)
    er, _ = oracle.dmrg(mpos, init, [M] * 4, 1e-10)
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    e0 = ed.ground_energy(h, v, L, L, 0)
    assert abs(eg[-1] - er[-1]) <= 1e-10, (eg, er)
    assert abs(eg[-1] - e0) <= 1e-9, (eg, e0)


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_sweeps_match_fp64(prec):
    """Mixed precision (FP32 or split-BF16 sigma vectors, then FP64 refinement; FP64 Gram-Schmidt, subspace problem,
    residuals and SVD): the energy after 4 sweeps of the open Hubbard chain L=16, M=128 agrees with the FP64 sweeps to
    1e-8 (the bar stated for mixed precision)."""
    print("mixed precision sweeps", prec, flush=True)
    L, M = 16, 128
    mpos = hubbard_mpo(L, 1.0, 4.0)
    init = initial_mps(L, (L, 0), M, 0)
    e64 = e_mx = None
    d64, dmx = DeviceDMRG(mpos, init, M), DeviceDMRG(mpos, init, M)
    nfp = 0
    for _ in range(4):
        e64, _ = d64.sweep(M, 1e-10, 0)
        e_mx, st = dmx.sweep(M, 1e-10, prec)
        nfp += st["nsigma_fp32"]
    print("E fp64", e64, "mixed", e_mx, "low-precision sigma", nfp, flush=True)
    assert nfp > 0
    assert abs(e_mx - e64) <= 1e-8, (e_mx, e64)


def _roots(mpos, L, qt, M, nroots, sweeps, tol):
    """Roots 0..nroots-1 one after another, each projected against the converged lower ones (dmrg/dmrg.cpp:36-72)."""
    done, energies = [], []
    for k in range(nroots):
        dm = DeviceDMRG(mpos, initial_mps(L, qt, M, k), M)
        for r, e_r in zip(done, energies):
            dm.add_lower_root(r, e_r)
        e_prev = None
        for _ in range(sweeps):
            e, _ = dm.sweep(M, tol)
            if e_prev is not None and abs(e - e_prev) < tol:
                break
            e_prev = e
        done.append(dm)
        energies.append(e)
    return energies


def test_excited_states_match_ed():
    """Excited states by sequential projected sweeps: Hubbard L = 6, U = 4, half filling, 2Sz = 0, M = 64 (exact at
    this size): roots 0..3 equal the four lowest eigenvalues of the sector Hamiltonian (dimension 400) to 1e-8."""
    print("excited states L=6", flush=True)
    L, M = 6, 64
    mpos = hubbard_mpo(L, 1.0, 4.0)
    es = _roots(mpos, L, (L, 0), M, 4, 10, 1e-11)
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    H = ed.hamiltonian(h, v, L, 0.0, ed.sector_states(L, L, 0)).toarray()
    w = np.linalg.eigvalsh(H)[:4]
    print("dmrg", es, "ed", w, flush=True)
    assert np.max(np.abs(np.array(es) - w)) < 1e-8, (es, w)


def test_onesite_refinement_sweeps_match_ed():
    """dmrg/dmrg.cpp:76-114: after the two-site sweeps, one-site sweeps (optimize_onesite, one-site sigma vector,
    diagonal and SVD) refine the state.  Hubbard L = 6, U = 4 at M = 64 (exact bond dimension here): the one-site
    energy equals ED to 1e-9, for the ground state and for root 1 projected against it.  L = 8 at M = 64 (truncated):
    a one-site sweep measures the truncated MPS itself, so its energy lies above the two-site value (which is taken
    before truncation) by the truncation error (< 1e-6 here) and does not rise from one one-site sweep to the next."""
    from mpsxx_amd.dmrg import SWEEP_ONESITE
    print("onesite sweeps", flush=True)
    L, M = 6, 64
    mpos = hubbard_mpo(L, 1.0, 4.0)
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    w = np.linalg.eigvalsh(ed.hamiltonian(h, v, L, 0.0, ed.sector_states(L, L, 0)).toarray())
    r0 = DeviceDMRG(mpos, initial_mps(L, (L, 0), M, 0), M)
    for _ in range(4):
        ea, _ = r0.sweep(M, 1e-11)
    r0.set_form(SWEEP_ONESITE)
    for _ in range(2):
        ea1, st = r0.sweep(M, 1e-11)
    assert abs(ea1 - w[0]) < 1e-9 and abs(ea - w[0]) < 1e-9, (ea, ea1, w[0])
    assert st["max_bond"] <= M
    r1 = DeviceDMRG(mpos, initial_mps(L, (L, 0), M, 1), M)
    r1.add_lower_root(r0, ea1)
    for _ in range(4):
        eb, _ = r1.sweep(M, 1e-11)
    r1.set_form(SWEEP_ONESITE)
    for _ in range(2):
        eb1, _ = r1.sweep(M, 1e-11)
    assert abs(eb1 - w[1]) < 1e-8, (eb, eb1, w[:3])
    # truncated case
    L, M = 8, 64
    mpos = hubbard_mpo(L, 1.0, 4.0)
    dm = DeviceDMRG(mpos, initial_mps(L, (L, 0), M, 0), M)
    for _ in range(4):
        e2, _ = dm.sweep(M, 1e-10)
    dm.set_form(SWEEP_ONESITE)
    e1a, _ = dm.sweep(M, 1e-10)
    e1b, _ = dm.sweep(M, 1e-10)
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    e0 = ed.ground_energy(h, v, L, L, 0)
    assert e2 - 1e-10 <= e1b <= e1a + 1e-10 and e1b - e0 < 1e-6, (e2, e1a, e1b, e0)
