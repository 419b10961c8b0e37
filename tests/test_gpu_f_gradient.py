"""Energy derivatives through the adjoint of the sigma-vector contraction (needs an MI355X).

The MPO gradient G_k = <psi|dH/dW_k|psi> of every site (mpsxx_dmrg_mpo_gradient) is mapped back through the QC MPO
generator's term records to dE/dh_pq and dE/dlambda (v -> lambda v).  Checks, for random 6-orbital integrals with a
well-separated ground state (two-site sweeps reach the FCI energy):
  * the rebuilt <H> equals the sweep energy;
  * dE/dh_pq equals the exact 1-RDM sum_s <a+_ps a_qs> and dE/dlambda the exact <V> of the FCI ground state (exact
    diagonalization, tests/ed.py);
  * Hellmann-Feynman through the expectation routine (src/expectation.hpp:55-114) on one-term MPOs (h = e_pq + e_qp);
  * central finite differences of the sweep energy, step 1e-4."""
import numpy as np
import pytest

import ed
from mpsxx_amd import qc
from mpsxx_amd.dmrg import DeviceDMRG, expectation, initial_mps

pytestmark = pytest.mark.gpu

K, M, SEED = 6, 64, 5


def integrals():
    return qc.synthetic_integrals(K, seed=SEED, nvec=8, hsig=0.2, lsig=0.25)


def converged(h, v, sweeps=4):
    m = qc.QcMpo(h, v)
    dm = DeviceDMRG(m.qsts(), initial_mps(K, (K, 0), M, 0), M)
    e = None
    for _ in range(sweeps):
        e, _ = dm.sweep(M, 1e-12)
    return m, dm, e


def test_gradient_matches_exact_rdm_and_expectation():
    print("gradient K=6", flush=True)
    h, v = integrals()
    m, dm, e = converged(h, v)
    grads = dm.mpo_gradient()
    dh, dl, e_rebuilt = m.gradient(grads)
    e0, vec, st = ed.ground_state(h, v, K, K, 0)
    assert abs(e - e0) < 1e-10
    assert abs(e_rebuilt - e) < 1e-9, (e_rebuilt, e)
    gamma = np.zeros((K, K))
    for p in range(K):
        for q in range(K):
            one = np.zeros((K, K))
            one[p, q] = 1.0
            gamma[p, q] = ed.expect(one, None, K, vec, st)
    # the state's own error is first order here: Davidson stops at ||r||^2 < 1e-13, i.e. |psi - psi_exact| ~ 3e-7 / gap
    # (gap 1.1), while the energy error is second order (< 1e-10)
    assert np.max(np.abs(dh - gamma)) < 1e-6, np.max(np.abs(dh - gamma))
    assert abs(dl - ed.expect(np.zeros((K, K)), v, K, vec, st)) < 1e-6
    # Hellmann-Feynman through the expectation routine on one-term MPOs, on the state the sweeps left
    mps = [dm.site_device(i) for i in range(K)]
    for (p, q) in [(0, 0), (1, 3), (2, 5), (4, 4)]:
        one = np.zeros((K, K))
        one[p, q] += 1.0
        one[q, p] += 1.0
        op = qc.QcMpo(one, np.zeros((K, K, K, K)))
        val = expectation(op.devices(), mps)
        want = dh[p, q] + (dh[q, p] if p != q else 0.0)
        assert abs(val - (want if p != q else 2 * dh[p, p])) < 1e-9, (p, q, val, want)
    # the gradient pass leaves the state usable: another sweep keeps the energy
    e2, _ = dm.sweep(M, 1e-12)
    assert abs(e2 - e) < 1e-10


@pytest.mark.parametrize("p,q", [(0, 1), (3, 3)])
def test_gradient_matches_central_differences(p, q):
    print("finite differences", p, q, flush=True)
    h, v = integrals()
    m, dm, e = converged(h, v)
    dh, dl, _ = m.gradient(dm.mpo_gradient())
    d = 1e-4
    sym = np.zeros((K, K))
    sym[p, q] = sym[q, p] = 1.0
    ep = converged(h + d * sym, v)[2]
    em = converged(h - d * sym, v)[2]
    fd = (ep - em) / (2 * d)
    an = dh[p, q] + dh[q, p] if p != q else dh[p, p]
    assert abs(fd - an) < 1e-6, (fd, an)
    lp = converged(h, (1 + d) * v)[2]
    lm = converged(h, (1 - d) * v)[2]
    assert abs((lp - lm) / (2 * d) - dl) < 1e-6
