"""End-to-end energies against values the reference's own logs hold (needs an MI355X).

Hubbard L=12, t=1, U=1, N=12, Sz=0 (reference input dmrg/tests/HUBBARD12: U on (ii|ii), t=+1 on (i+1,i)),
M=200: E0 = -11.8406372775523305 (dmrg/tests/hub12_dmrg.out:15536; one-site restart -11.8406372775523234,
dmrg/tests/onedot.out:1045).  The reference ran two-site then one-site sweeps with tolerance 1e-8; the tolerance on
our two-site result covers the M=200 truncation and the Davidson tolerance.
"""
import pytest

from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
from mpsxx_amd.mpo import hubbard_mpo

pytestmark = pytest.mark.gpu

E_HUB12 = -11.8406372775523305


def test_hubbard12_ground_energy_matches_reference_log():
    print("hub12", flush=True)
    L, M = 12, 200
    dm = DeviceDMRG(hubbard_mpo(L, 1.0, 1.0), initial_mps(L, (L, 0), M, 0), M)
    e_prev, e = None, None
    for sweep in range(8):
        e, st = dm.sweep(M, 1e-8)
        print("sweep", sweep, e, st["seconds"], st["max_bond"], flush=True)
        if e_prev is not None and abs(e - e_prev) < 1e-10:
            break
        e_prev = e
    assert abs(e - E_HUB12) < 2e-7, e
