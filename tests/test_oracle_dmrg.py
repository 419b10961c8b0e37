"""Pins of the oracle's canonicalize and two-site sweeps (CPU): exact diagonalization and algebraic identities."""
import numpy as np
import pytest

import ed
import oracle
from mpsxx_amd import mps
from mpsxx_amd.dmrg import initial_mps
from mpsxx_amd.mpo import hubbard_mpo


@pytest.mark.parametrize("L,U,M", [(4, 4.0, 16), (6, 4.0, 64), (8, 4.0, 64), (6, 1.0, 64)])
def test_oracle_dmrg_matches_ed(L, U, M):
    h, v = ed.hubbard_integrals(L, 1.0, U)
    e0 = ed.ground_energy(h, v, L, L, 0)
    es, ns = oracle.dmrg(hubbard_mpo(L, 1.0, U), initial_mps(L, (L, 0), M, 0), [M] * 4, 1e-10)
    assert abs(es[-1] - e0) < 1e-9, (es, e0)
    assert ns > 0


def test_oracle_dmrg_free_fermions_l8():
    es, _ = oracle.dmrg(hubbard_mpo(8, 1.0, 0.0), initial_mps(8, (8, 0), 64, 1), [64] * 4, 1e-10)
    assert abs(es[-1] - (-9.51754097)) < 1e-7


@pytest.mark.parametrize("forward", [True, False])
def test_oracle_canonicalize_reconstructs_and_is_isometric(forward):
    rng = np.random.default_rng(3)
    ch = mps.random_chain(6, 2, (6, 0), rng, dmax=3)
    psi = ch["psi"]
    U, V, dw = oracle.canonicalize_twosite(forward, psi, 0)
    assert dw < 1e-20
    rec = oracle.contract(1.0, U, [2], V, [0])
    assert rec.block_set() <= psi.block_set() | rec.block_set()
    scale = max(float(np.max(np.abs(b))) for b in psi.blocks.values())
    assert rec.max_abs_diff(psi) < 1e-12 * scale
    iso = U if forward else V
    # sum over the two physical/bond legs gives the identity on the new bond
    if forward:
        g = oracle.contract(1.0, iso, [0, 1], iso, [0, 1], conj_a=True)
    else:
        g = oracle.contract(1.0, iso, [1, 2], iso, [1, 2], conj_a=True)
    gd = g.dense()
    assert np.max(np.abs(gd - np.eye(gd.shape[0]))) < 1e-12


def test_oracle_truncation_keeps_m_largest():
    rng = np.random.default_rng(4)
    ch = mps.random_chain(6, 2, (6, 0), rng, dmax=4)
    psi = ch["psi"]
    _, _, d0 = oracle.canonicalize_twosite(True, psi, 0)
    U, V, dw = oracle.canonicalize_twosite(True, psi, 5)
    assert sum(U.dshape[2]) >= 5
    # discarded weight + kept weight = ||psi||^2
    kept = V.norm() ** 2
    assert abs(kept + dw - psi.norm() ** 2) < 1e-10 * psi.norm() ** 2
