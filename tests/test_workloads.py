"""Host side of the synthetic workloads (CPU): the isometric chains behind the batched Davidson problems."""
import numpy as np
import pytest

import oracle
from mpsxx_amd import mps, workloads


@pytest.mark.parametrize("L,M,seed", [(12, 32, 0), (16, 64, 1), (8, 4096, 2)])
def test_isometric_chain_tensors_are_isometries(L, M, seed):
    A, B, psi = workloads.isometric_chain(L, M, seed)
    i = L // 2 - 1
    assert len(A) == i and sorted(B) == list(range(i + 2, L))
    for a in A:
        d = a.dense()
        m = d.reshape(-1, d.shape[2])
        assert np.allclose(m.T @ m, np.eye(m.shape[1]), atol=1e-12)
    for b in B.values():
        d = b.dense()
        m = d.reshape(d.shape[0], -1)
        assert np.allclose(m @ m.T, np.eye(m.shape[0]), atol=1e-12)
    # bond dimensions never exceed M (up to the capacity rule) and match between neighbours
    for k in range(1, i):
        assert A[k].dshape[0] == A[k - 1].dshape[2]
    assert psi.dshape[0] == A[i - 1].dshape[2] and psi.dshape[3] == B[i + 2].dshape[0]


def test_isometric_chain_full_space_is_the_whole_sector():
    """With M above every bond's capacity the middle bonds hold all states of their side: the left and right blocks
    of the half-filled 2Sz = 0 sector, so sum over psi's allowed blocks = the sector dimension C(L, L/2)^2."""
    from math import comb
    L = 8
    _, _, psi = workloads.isometric_chain(L, 4096, 0)
    assert psi.nelem() == comb(L, L // 2) ** 2


def test_isometric_chain_overlap_environments_are_identities():
    """The oracle's overlap renormalization (src/renormalize.hpp:36-53) of the isometric tensors gives identity
    environments on both sides, so the two-site problem's overlap is the identity and H_eff = P^T H P."""
    L = 10
    A, B, psi = workloads.isometric_chain(L, 12, 3)
    i = L // 2 - 1
    lsov = mps.boundary_overlap()
    for k in range(i):
        lsov = oracle.renormalize_overlap(True, lsov, A[k], A[k])
    rsov = mps.boundary_overlap()
    for k in range(L - 1, i + 1, -1):
        rsov = oracle.renormalize_overlap(False, rsov, B[k], B[k])
    for e in (lsov, rsov):
        d = e.dense()
        assert np.allclose(d, np.eye(d.shape[0]), atol=1e-12)
