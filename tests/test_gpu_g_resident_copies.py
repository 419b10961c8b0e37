"""Reduced-precision plans keep FP32 copies of their operands between calls (the `_ex` entry points and the Davidson
driver).  These tests pin when such a copy must be made again although the caller's convert mask does not ask for
it: a plan built fresh for a non-empty output, an operand tensor other than the one the copy was made from, and
operator groups of equal structure in one Davidson solve (needs an MI355X).

Bars: FP32 sigma within 2e-5 relative Frobenius norm of the oracle; Davidson energies within 1e-10 of FP64."""
import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from helpers import hubbard_problem

pytestmark = pytest.mark.gpu

WFN_MERGED = 1 << 2   # convert_mask bit of the wavefunction in the merged form (argument order lopr, ropr, wfn0)
ALL = 0xFFFFFFFF


def _scaled(q, a):
    out = q.copy()
    for k in out.blocks:
        out.blocks[k] = a * out.blocks[k]
    return out


def test_merged_ex_reconverts_fresh_plans_and_other_operands():
    from mpsxx_amd import workloads
    print("merged _ex resident copies", flush=True)
    Lm, Rm, psim, info = workloads.qc_merged_problem(8, 48, seed=3)
    ref = oracle.sigma_merged(Lm.to_qst(), Rm.to_qst(), psim.to_qst())
    tol = 2e-5 * ref.norm()
    # first call: every copy converted
    s1 = mx.compute_sigma_vector_merged_ex(Lm, Rm, psim, None, mx.FP32, ALL)
    assert s1.to_qst().diff_norm(ref) <= tol
    # a second operator of the same structure with other contents: only the wavefunction bit is set, but the plan's
    # copy was made from Lm, so it must be converted from Lm2
    Lm2 = mx.copy(Lm)
    mx.scal(0.5, Lm2)
    s2 = mx.compute_sigma_vector_merged_ex(Lm2, Rm, psim, None, mx.FP32, WFN_MERGED)
    assert s2.to_qst().diff_norm(_scaled(ref, 0.5)) <= 0.5 * tol
    # accumulating into the non-empty s1 selects a plan of its own (the output structure is part of its key), built
    # fresh: its copies hold nothing until converted, whatever the mask says
    s3 = mx.compute_sigma_vector_merged_ex(Lm, Rm, psim, s1, mx.FP32, WFN_MERGED)
    assert s3.to_qst().diff_norm(_scaled(ref, 2.0)) <= 2 * tol
    # and back to Lm2 through that plan
    s4 = mx.compute_sigma_vector_merged_ex(Lm2, Rm, psim, s2, mx.FP32, WFN_MERGED)
    assert s4.to_qst().diff_norm(ref) <= tol


def test_twosite_ex_reconverts_other_operands():
    print("two-site _ex resident copies", flush=True)
    rng = np.random.default_rng(7)
    L, i = 8, 3
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=4)
    ref = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, ch["psi"])
    d = [mx.DTensor(x) for x in (mpos[i], mpos[i + 1], lenv, renv, ch["psi"])]
    tol = 2e-5 * ref.norm()
    s1 = mx.compute_sigma_vector_twosite_ex(*d, None, mx.FP32, ALL)
    assert s1.to_qst().diff_norm(ref) <= tol
    le2 = mx.DTensor(_scaled(lenv, 3.0))
    s2 = mx.compute_sigma_vector_twosite_ex(d[0], d[1], le2, d[3], d[4], None, mx.FP32, 1 << 4)
    assert s2.to_qst().diff_norm(_scaled(ref, 3.0)) <= 3 * tol


def test_davidson_fp32_two_groups_of_equal_structure():
    """H = H_a + H_b with H_a, H_b built from the same MPO and right environment and a random elementwise split of
    the left environment (equal block structures, different contents), on an isometric chain (bounded spectrum).
    The split must not change the FP64 energy, and the mixed-precision solve must run its FP32 phase on H itself:
    at tolerance 1e-6 the FP32 phase converges to ||r||^2 < 1e-6 and the FP64 refinement from that vector then
    needs at most two sigma vectors, so the two-group solve takes no more sigma vectors than the one-group solve
    plus two (an FP32 phase on the wrong operator would leave a large FP64 residual)."""
    from helpers import chain_envs, isometric_chain
    from mpsxx_amd.mpo import hubbard_mpo
    print("davidson fp32 equal-structure groups", flush=True)
    rng = np.random.default_rng(113)
    L, i = 10, 4
    mpos = hubbard_mpo(L, 1.0, 4.0)
    ch = isometric_chain(L, i, (L, 0), rng)
    lenv, renv = chain_envs(mpos, ch, i, L)
    psi = ch["psi"]
    la = lenv.copy()
    lb = lenv.copy()
    for k in lenv.blocks:
        r = rng.uniform(0.0, 1.0, size=lenv.blocks[k].shape)
        la.blocks[k] = r * lenv.blocks[k]
        lb.blocks[k] = lenv.blocks[k] - la.blocks[k]
    diag = oracle.diag_twosite(mpos[i], mpos[i + 1], lenv, renv, psi.empty_like())
    D = mx.DTensor
    one = [(D(mpos[i]), D(mpos[i + 1]), D(lenv), D(renv))]
    two = [(D(mpos[i]), D(mpos[i + 1]), D(la), D(renv)), (D(mpos[i]), D(mpos[i + 1]), D(lb), D(renv))]
    e64, _ = mx.davidson_twosite(one, D(diag), D(psi), 1e-10)
    e2_64, _ = mx.davidson_twosite(two, D(diag), D(psi), 1e-10)
    print("fp64 one group", e64, "two groups", e2_64, flush=True)
    assert abs(e2_64 - e64) < 1e-10 * max(1.0, abs(e64))
    e1_32, n1 = mx.davidson_twosite(one, D(diag), D(psi), 1e-6, precision=mx.FP32)
    e2_32, n2 = mx.davidson_twosite(two, D(diag), D(psi), 1e-6, precision=mx.FP32)
    print("fp32 one group", e1_32, n1, "two groups", e2_32, n2, flush=True)
    assert abs(e2_32 - e64) < 1e-6 * max(1.0, abs(e64))
    assert n2 <= n1 + 2, (n1, n2)
