"""Quantum-chemistry inputs (CPU): FCIDUMP I/O, synthetic integrals and the QC MPO generator, pinned by exact
diagonalization (tests/ed.py builds H from the integrals by second quantization on bit strings, independently of the
MPO code; its two-body convention is pinned against the Slater-Condon two-electron matrix below).

The generator is host code in the library (mpsxx_qc_mpo_*), callable without a GPU."""
import itertools
import os

import numpy as np
import pytest

import ed
import oracle
from mpsxx_amd import qc
from mpsxx_amd.dmrg import initial_mps
from mpsxx_amd.mpo import hubbard_mpo, mpo_dense

HERE = os.path.dirname(os.path.abspath(__file__))


def test_ed_two_body_convention_slater_condon():
    """<p_a q_b|H|r_a s_b> = h_pr d_qs + h_qs d_pr + (pr|qs) for one alpha and one beta electron."""
    K = 3
    h, v = qc.synthetic_integrals(K, seed=5, nvec=3, hsig=0.4, lsig=0.5)
    Hsc = np.zeros((K * K, K * K))
    for p, q, r, s in itertools.product(range(K), repeat=4):
        Hsc[p * K + q, r * K + s] = h[p, r] * (q == s) + h[q, s] * (p == r) + v[p, r, q, s]
    assert abs(np.linalg.eigvalsh(Hsc)[0] - ed.ground_energy(h, v, K, 2, 0)) < 1e-12


@pytest.mark.parametrize("K", [1, 2, 3, 4])
@pytest.mark.parametrize("swap", [True, False])
def test_qc_mpo_equals_exact_hamiltonian(K, swap):
    h, v = qc.synthetic_integrals(K, seed=10 + K, nvec=3, hsig=0.3, lsig=0.3)
    m = qc.QcMpo(h, v, ecore=0.37, swap_sweep=swap)
    H = mpo_dense(m.qsts())
    Hed = ed.hamiltonian(h, v, K, ecore=0.37).toarray()
    assert np.max(np.abs(H - Hed)) < 1e-12


def test_qc_mpo_dense_random_integrals_without_symmetry_in_values():
    """Integrals with every 8-fold class independent (as the reference's test generator writes them)."""
    K = 4
    rng = np.random.default_rng(7)
    v = np.zeros((K, K, K, K))
    for i, j, k, l in itertools.product(range(K), repeat=4):
        if v[i, j, k, l] == 0.0:
            x = rng.uniform(-1, 1)
            for (a, b, c, d) in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                                 (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                v[a, b, c, d] = x
    h = rng.uniform(-1, 1, size=(K, K))
    h = h + h.T
    H = mpo_dense(qc.QcMpo(h, v).qsts())
    assert np.max(np.abs(H - ed.hamiltonian(h, v, K).toarray())) < 1e-12


@pytest.mark.parametrize("K", [6, 10, 16, 24])
def test_qc_mpo_bond_dimensions_follow_the_reference_operator_sets(K):
    """D_w = 2 + 4K + 8n^2 - 2n, n = min(L, R) with swap sweeps (src/boundary_opinfo.cpp:58,65).  Without swap sweeps
    the pairs stay on the left (n = L); that count is the reference's estimate before it drops bond operators no term
    uses (boundary_opinfo::clean, :143-154), so near the right end, where one right site cannot carry every pair
    type, the generator's count is at most the estimate."""
    h, v = qc.synthetic_integrals(K, seed=1)
    for swap in (True, False):
        if K == 24 and not swap:
            continue
        d = qc.QcMpo(h, v, swap_sweep=swap).bond_dims()
        assert d[0] == 1 and d[K] == 1
        for b in range(K - 1):
            L, R = b + 1, K - b - 1
            n = (L if L < R else R) if swap else L
            est = 2 + 4 * K + 8 * n * n - 2 * n
            if swap or R >= 2:
                assert d[b + 1] == est, (b, d)
            else:
                assert d[b + 1] <= est, (b, d)
    if K == 24:
        assert max(d) == 1226


def test_qc_mpo_of_hubbard_integrals_equals_compact_hubbard_mpo():
    L, t, U = 4, 1.0, 3.0
    h, v = ed.hubbard_integrals(L, t, U)
    Hq = mpo_dense(qc.QcMpo(h, v).qsts())
    Hc = mpo_dense(hubbard_mpo(L, t, U))
    assert np.max(np.abs(Hq - Hc)) < 1e-12


def test_fcidump_round_trip(tmp_path):
    h, v = qc.synthetic_integrals(5, seed=3)
    p = tmp_path / "FCIDUMP"
    qc.write_fcidump(str(p), h, v, ecore=1.25, nelec=4, ms2=2)
    norb, nelec, ms2, h2, v2, e2 = qc.read_fcidump(str(p))
    assert (norb, nelec, ms2) == (5, 4, 2)
    assert abs(e2 - 1.25) < 1e-15
    assert np.max(np.abs(h2 - h)) < 1e-15 and np.max(np.abs(v2 - v)) < 1e-15


def test_regenerated_test_integrals_fixture():
    """tests/golden/intgen6.FCIDUMP (regenerated by tests/golden/make_intgen_fcidump.cpp): 231 two-electron lines of
    6 orbitals (1 + 5*... classes per orbital), 21 one-electron lines, the core energy; 8-fold symmetric on read."""
    norb, nelec, ms2, h, v, ec = qc.read_fcidump(os.path.join(HERE, "golden", "intgen6.FCIDUMP"))
    assert (norb, nelec, ms2, ec) == (6, 6, 0, 0.0)
    assert np.allclose(v, v.transpose(1, 0, 2, 3)) and np.allclose(v, v.transpose(2, 3, 0, 1))
    assert np.allclose(h, h.T)
    assert abs(v[0, 0, 0, 0] - 1.3547700429678050e-01) == 0.0
    assert abs(v[1, 1, 1, 0] - 4.8443388556211570e-01) == 0.0


@pytest.mark.parametrize("seed", [5, 7])
def test_oracle_dmrg_with_qc_mpo_reaches_fci(seed):
    """Two-site sweeps of the oracle with the QC MPO of random 6-orbital integrals: FCI energy of the (6, 2Sz=0)
    sector (dimension 400) to 1e-10 (integral seeds whose first excitation gap is >= 0.4, so the Davidson residual
    tolerance 1e-11 bounds the energy error by ~1e-11)."""
    K = 6
    h, v = qc.synthetic_integrals(K, seed=seed, nvec=8, hsig=0.2, lsig=0.25)
    e_fci = ed.ground_energy(h, v, K, K, 0)
    mpos = qc.QcMpo(h, v).qsts()
    es, ns = oracle.dmrg(mpos, initial_mps(K, (K, 0), 64, 0), [64] * 4, 1e-10)
    assert abs(es[-1] - e_fci) < 1e-10, (es, e_fci)
