"""The oracle with its block dgemm routed through numpy's bundled OpenBLAS (oracle.use_blas, used for the CPU
baseline timings) gives the values of its own loops to rounding, for the two-site and merged sigma vectors and
with several OpenMP threads (CPU)."""
import numpy as np
import pytest

import oracle
from helpers import chain_envs
from mpsxx_amd import mps
from mpsxx_amd.mpo import hubbard_mpo


def _close(a, b, tol=1e-13):
    assert a.block_set() == b.block_set()
    scale = max(1.0, max(float(np.max(np.abs(v))) for v in b.blocks.values() if v.size))
    assert a.max_abs_diff(b) <= tol * scale


@pytest.mark.parametrize("threads", [1, 4])
def test_blas_dgemm_matches_the_oracle_loops(threads):
    L, i = 8, 3
    rng = np.random.default_rng(7)
    mpos = hubbard_mpo(L, 1.0, 4.0)
    ch = mps.random_chain(L, i, (L, 0), rng, dmax=6)
    lenv, renv = chain_envs(mpos, ch, i, L)
    psi = ch["psi"]
    Lt = oracle.merged_operator(True, mpos[i], lenv)
    Rt = oracle.merged_operator(False, mpos[i + 1], renv)
    psim = oracle.qst_merge(psi, 2, 2)
    ref4 = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, psi)
    refm = oracle.sigma_merged(Lt, Rt, psim)
    try:
        cfg = oracle.use_blas(True, 1)
        assert "OpenBLAS" in cfg
        oracle.set_threads(threads)
        _close(oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, psi), ref4)
        _close(oracle.sigma_merged(Lt, Rt, psim), refm)
    finally:
        oracle.use_blas(False)
        oracle.set_threads(1)
