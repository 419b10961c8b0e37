"""Pins of the oracle (CPU): H_eff = P^T H P from exact diagonalization, energies, dense brute force.

The environments come from the oracle's own renormalize chains, so each check pins renormalize (both directions)
together with the sigma / diagonal contraction it feeds.
"""
import numpy as np
import pytest

import ed
import oracle
from helpers import chain_overlaps, hubbard_problem, projectors
from mpsxx_amd import mps
from mpsxx_amd.mpo import SITE_Q, hubbard_mpo


def _heff_apply(H, PL, PR, psid):
    X, Y = PL.shape[0], PR.shape[1]
    Psi = np.einsum("xa,anmb,by->xnmy", PL, psid, PR).reshape(-1)
    HPsi = (H @ Psi).reshape(X, 4, 4, Y)
    return np.einsum("xa,xnmy,by->anmb", PL, HPsi, PR), Psi


@pytest.mark.parametrize("L,i,seed", [(4, 1, 0), (5, 2, 1), (6, 2, 2), (6, 0, 3), (6, 4, 4)])
def test_sigma_twosite_equals_projected_hamiltonian(L, i, seed):
    rng = np.random.default_rng(seed)
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=2, U=4.0)
    sig = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, ch["psi"])
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    H = ed.hamiltonian(h, v, L)
    PL, PR = projectors(ch, i, L)
    ref, Psi = _heff_apply(H, PL, PR, ch["psi"].dense())
    got = sig.dense()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))
    # energy functional <Psi|H|Psi> = <psi|sigma>
    e_full = Psi @ (H @ Psi)
    e_loc = float(np.sum(ch["psi"].dense() * got))
    assert abs(e_full - e_loc) < 1e-9 * max(1.0, abs(e_full))
    # sigma blocks respect the quantum numbers of psi
    for idx in sig.blocks:
        assert ch["psi"].allowed(idx)


@pytest.mark.parametrize("L,i,seed", [(5, 1, 5), (6, 3, 6)])
def test_diagonal_twosite_equals_projected_diagonal(L, i, seed):
    rng = np.random.default_rng(seed)
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=2, U=3.0)
    diag = ch["psi"].empty_like()
    d = oracle.diag_twosite(mpos[i], mpos[i + 1], lenv, renv, diag)
    h, v = ed.hubbard_integrals(L, 1.0, 3.0)
    H = ed.hamiltonian(h, v, L).toarray()
    PL, PR = projectors(ch, i, L)
    # diag of P^T H P over the (a, n1, n2, b) basis
    X, Y = PL.shape[0], PR.shape[1]
    Hr = H.reshape(X, 4, 4, Y, X, 4, 4, Y)
    Hd = np.einsum("xnmyXnmY,xa,Xa,by,bY->anmb", Hr, PL, PL, PR, PR, optimize=True)
    got = d.dense()
    # compare on the entries of QN-allowed blocks of psi
    mask = np.zeros_like(got, dtype=bool)
    psi = ch["psi"]
    offs = [psi.offsets(k) for k in range(4)]
    import itertools
    for idx in itertools.product(*[range(psi.nsec(k)) for k in range(4)]):
        if psi.allowed(idx):
            sl = tuple(slice(offs[k][j], offs[k][j] + psi.dshape[k][j]) for k, j in enumerate(idx))
            mask[sl] = True
    assert np.max(np.abs((got - Hd)[mask])) < 1e-10 * max(1.0, np.max(np.abs(Hd)))


def test_sigma_onesite_equals_projected_hamiltonian():
    L, i = 5, 2
    rng = np.random.default_rng(7)
    mpos = hubbard_mpo(L, 1.0, 2.0)
    qt = (L, 1)
    ch = mps.random_chain(L, i, qt, rng, dmax=2)
    # one-site centre at site i: left env from A_0..A_{i-1}, right env from B_{i+1}..: build B_{i+1} too
    rq_i1 = mps.right_bond_qns(L, i + 1, qt)
    rd_i1 = [int(rng.integers(1, 3)) for _ in rq_i1]
    B = dict(ch["B"])
    B[i + 1] = mps.fill_allowed(mps.site_tensor(rq_i1, rd_i1, ch["rq"][i + 2], ch["rd"][i + 2]), rng)
    wf = mps.fill_allowed(mps.site_tensor(ch["lq"][i], ch["ld"][i], rq_i1, rd_i1, total=qt), rng)
    lenv = mps.boundary_env()
    for k in range(i):
        lenv = oracle.renormalize(True, mpos[k], lenv, ch["A"][k], ch["A"][k])
    renv = mps.boundary_env()
    for k in range(L - 1, i, -1):
        renv = oracle.renormalize(False, mpos[k], renv, B[k], B[k])
    sig = oracle.sigma_onesite(mpos[i], lenv, renv, wf)
    # dense reference
    PL = np.ones((1, 1))
    for k in range(i):
        x = np.einsum("xa,anb->xnb", PL, ch["A"][k].dense())
        PL = x.reshape(-1, x.shape[2])
    PR = np.ones((1, 1))
    for k in range(L - 1, i, -1):
        x = np.einsum("anb,by->any", B[k].dense(), PR)
        PR = x.reshape(x.shape[0], -1)
    h, v = ed.hubbard_integrals(L, 1.0, 2.0)
    H = ed.hamiltonian(h, v, L)
    Psi = np.einsum("xa,anb,by->xny", PL, wf.dense(), PR).reshape(-1)
    HPsi = (H @ Psi).reshape(PL.shape[0], 4, PR.shape[1])
    ref = np.einsum("xa,xny,by->anb", PL, HPsi, PR)
    assert np.max(np.abs(sig.dense() - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))


def test_overlap_sigma_equals_projected_identity():
    L, i = 6, 2
    rng = np.random.default_rng(8)
    ch = mps.random_chain(L, i, (L, L % 2), rng, dmax=2)
    lsov, rsov = chain_overlaps(ch, i, L)
    s = oracle.sigma_overlap(lsov, rsov, ch["psi"])
    PL, PR = projectors(ch, i, L)
    ref = np.einsum("xa,xc,cnmd,yd,yb->anmb", PL, PL, ch["psi"].dense(), PR.T, PR.T)
    assert np.max(np.abs(s.dense() - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))


def test_renormalize_chain_energy_both_directions():
    """<Psi|H|Psi> from a full left sweep of renormalize equals the right sweep and ED."""
    L = 5
    rng = np.random.default_rng(9)
    mpos = hubbard_mpo(L, 1.0, 4.0)
    qt = (L, 1)
    # a full left-canonical-shaped chain: A_0..A_{L-1} (last tensor carries the total)
    lq = [mps.left_bond_qns(L, k, qt) for k in range(L + 1)]
    lq[L] = [qt]
    ld = [[1]] + [[int(rng.integers(1, 3)) for _ in lq[k]] for k in range(1, L)] + [[1]]
    A = [mps.fill_allowed(mps.site_tensor(lq[k], ld[k], lq[k + 1], ld[k + 1]), rng) for k in range(L)]
    env = mps.boundary_env()
    for k in range(L):
        env = oracle.renormalize(True, mpos[k], env, A[k], A[k])
    # env is (r_bra, w, r_ket) on the single right-boundary sector
    e_left = float(env.dense().reshape(-1)[0])
    # right chain on the same tensors: shift total to the first site by reading the chain backwards
    # (quantum numbers: B_k(l,n,r) with q_l = q_r - q_n, boundary (qt) on the right)
    Psi = np.ones((1, 1))
    for k in range(L):
        x = np.einsum("xa,anb->xnb", Psi, A[k].dense())
        Psi = x.reshape(-1, x.shape[2])
    Psi = Psi[:, 0]
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    H = ed.hamiltonian(h, v, L)
    assert abs(e_left - Psi @ (H @ Psi)) < 1e-9 * max(1.0, abs(e_left))
