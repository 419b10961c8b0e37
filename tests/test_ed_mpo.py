"""Pins of the test-side exact diagonalization and of the Hubbard MPO (CPU)."""
import numpy as np
import pytest

import ed
from mpsxx_amd.mpo import hubbard_mpo, mpo_dense


def test_two_site_hubbard_closed_form():
    # E0 = (U - sqrt(U^2 + 16 t^2)) / 2 for 2 sites, 2 electrons, Sz = 0
    for U in (1.0, 4.0, 8.0):
        h, v = ed.hubbard_integrals(2, 1.0, U)
        e = ed.ground_energy(h, v, 2, 2, 0)
        assert abs(e - (U - np.sqrt(U * U + 16.0)) / 2) < 1e-12


@pytest.mark.parametrize("L,ref", [(8, -9.51754097)])
def test_free_fermions(L, ref):
    h, v = ed.hubbard_integrals(L, 1.0, 0.0)
    e = ed.ground_energy(h, v, L, L, 0)
    eps = sorted(2.0 * np.cos(np.pi * k / (L + 1)) for k in range(1, L + 1))
    assert abs(e - 2 * sum(eps[: L // 2])) < 1e-9
    assert abs(e - ref) < 1e-8


def test_atomic_limit():
    # t = 0 at half filling: every site singly occupied, E0 = 0
    h, v = ed.hubbard_integrals(4, 0.0, 4.0)
    assert abs(ed.ground_energy(h, v, 4, 4, 0)) < 1e-12


@pytest.mark.parametrize("L,U", [(2, 4.0), (3, 2.0), (4, 4.0)])
def test_hubbard_mpo_equals_ed(L, U):
    h, v = ed.hubbard_integrals(L, 1.0, U)
    H = ed.hamiltonian(h, v, L).toarray()
    M = mpo_dense(hubbard_mpo(L, 1.0, U))
    assert np.max(np.abs(H - M)) < 1e-12


def test_hubbard_l8_ed_energy():
    # ED of the config-1 system (L=8, t=1, U=4, N=8, Sz=0; sector dimension 4900)
    h, v = ed.hubbard_integrals(8, 1.0, 4.0)
    st = ed.sector_states(8, 8, 0)
    assert len(st) == 4900
    e = ed.ground_energy(h, v, 8, 8, 0)
    assert -9.51754097 < e < -3.0
