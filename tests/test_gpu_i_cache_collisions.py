"""Plan caches under structure-hash collisions (needs an MI355X).

Cached plans are keyed by 64-bit structure hashes of their operands; a hit must still be taken only for operands of
the structures the plan was built for.  With the hashes cut to one or two bits (mpsxx_set_structure_hash_bits),
distinct structures share keys all the time: every result must stay the one computed with full hashes (the oracle's
for single operations, bit-identical energies for a sweep) and the caches must report the collisions they resolved."""
import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from helpers import hubbard_problem

pytestmark = pytest.mark.gpu


def _assert_same(gpu, ref, tol=1e-12):
    assert gpu.block_set() == ref.block_set()
    scale = max(1.0, max(float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size))
    assert gpu.max_abs_diff(ref) <= tol * scale


@pytest.fixture
def few_hash_bits():
    yield
    mx.set_structure_hash_bits(64)


def test_sigma_and_diagonal_under_one_bit_hashes(few_hash_bits):
    print("cache collisions: sigma", flush=True)
    mx.set_structure_hash_bits(1)
    c0 = mx.plan_cache_stats()["collisions"]
    cases = []
    for k, (L, i, dmax) in enumerate(((6, 2, 2), (6, 2, 3), (8, 3, 3), (8, 3, 4), (8, 1, 4))):
        mpos, ch, lenv, renv = hubbard_problem(L, i, np.random.default_rng(40 + k), dmax=dmax)
        cases.append((mpos[i], mpos[i + 1], lenv, renv, ch["psi"]))
    for rep in range(2):
        for lm, rm, le, re_, psi in cases:
            d = [mx.DTensor(x) for x in (lm, rm, le, re_, psi)]
            _assert_same(mx.compute_sigma_vector_twosite(*d).to_qst(), oracle.sigma_twosite(lm, rm, le, re_, psi))
            dg = mx.compute_diagonal_elements_twosite(*d[:4], mx.DTensor(psi.empty_like()))
            _assert_same(dg.to_qst(), oracle.diag_twosite(lm, rm, le, re_, psi.empty_like()))
    assert mx.plan_cache_stats()["collisions"] > c0


def test_sweep_energies_unchanged_under_two_bit_hashes(few_hash_bits):
    from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
    from mpsxx_amd.mpo import hubbard_mpo
    print("cache collisions: sweeps", flush=True)
    L, M = 8, 32
    mpos = hubbard_mpo(L, 1.0, 4.0)

    def run():
        dm = DeviceDMRG(mpos, initial_mps(L, (L, 0), M, 0), M)
        return [dm.sweep(M, 1e-10)[0] for _ in range(3)]

    ref = run()
    mx.set_structure_hash_bits(2)
    c0 = mx.plan_cache_stats()["collisions"]
    got = run()
    print("full hashes", ref, "two-bit hashes", got, flush=True)
    assert got == ref
    assert mx.plan_cache_stats()["collisions"] > c0
