"""The contraction planner on the CPU (mpsxx_contract_structure: host only, no device work).

For random block-sparse pairs (ranks 3-5, one or two contracted indices, conjugated views) the planner's output blocks
and per-block pair lists equal a brute-force enumeration of the reference rule (include/btas/SPARSE/STBLAS.h:316-356:
an output block exists iff it is quantum-number allowed and at least one (A-block, B-block) pair matches on the
contracted sectors; pairs accumulate in contracted-tag order), the output block set equals the oracle's Contract, and
the plan is the same for 1 and 4 planner threads."""
import ctypes
import itertools

import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from mpsxx_amd.qst import QST

VP = ctypes.c_void_p


def _neg(qs):
    return [(-a, -b) for (a, b) in qs]


def random_pair(rng, rank_a, rank_b, k, conj_a, conj_b, density=0.7):
    """A and B with k contracted indices at random positions; B's contracted sectors are the conjugates of A's.
    Dense dims > 1 only on one free index of each side and on the first contracted index (the planner expresses
    permuted blocks as strided 2-D views, which covers every index order of the hot path, where MPO and site sectors
    have dimension 1)."""
    ca = sorted(rng.choice(rank_a, size=k, replace=False).tolist())
    ca = [int(x) for x in rng.permutation(ca)]
    cb = sorted(rng.choice(rank_b, size=k, replace=False).tolist())
    cb = [int(x) for x in rng.permutation(cb)]

    def index():
        # few distinct quantum numbers (repeats allowed, as sectors of equal QN on a bond), so that a good share of
        # the index tuples is allowed
        n = int(rng.integers(3, 7))
        return [(int(rng.integers(-1, 2)), int(rng.choice([-1, 0, 0, 0, 0, 1]))) for _ in range(n)], \
               [int(rng.integers(1, 4)) for _ in range(n)]

    aq, ad = zip(*[index() for _ in range(rank_a)])
    aq, ad = list(aq), list(ad)
    bq, bd = zip(*[index() for _ in range(rank_b)])
    bq, bd = list(bq), list(bd)
    fa = int(rng.choice([i for i in range(rank_a) if i not in ca]))
    fb = int(rng.choice([i for i in range(rank_b) if i not in cb]))
    for i in range(rank_a):
        if i not in (fa, ca[0]):
            ad[i] = [1] * len(ad[i])
    for i in range(rank_b):
        if i != fb:
            bd[i] = [1] * len(bd[i])
    for x, y in zip(ca, cb):
        a_eff = _neg(aq[x]) if conj_a else aq[x]
        b_eff = _neg(a_eff)
        bq[y] = _neg(b_eff) if conj_b else b_eff
        bd[y] = list(ad[x])

    def fill(qs, ds):
        # total quantum number of one random product, so that some blocks are allowed
        pick = [int(rng.integers(0, len(q))) for q in qs]
        q = (sum(qs[i][p][0] for i, p in enumerate(pick)), sum(qs[i][p][1] for i, p in enumerate(pick)))
        t = QST(q, qs, ds)
        for idx in itertools.product(*[range(len(x)) for x in qs]):
            if t.allowed(idx) and (rng.random() < density or idx == tuple(pick)):
                t.blocks[idx] = rng.standard_normal(t.block_shape(idx))
        return t

    return fill(aq, ad), fill(bq, bd), ca, cb


def brute_force(a, b, ca, cb, conj_a, conj_b):
    ua = [i for i in range(a.rank) if i not in ca]
    ub = [i for i in range(b.rank) if i not in cb]
    sa = -1 if conj_a else 1
    sb = -1 if conj_b else 1
    cq = (sa * a.q[0] + sb * b.q[0], sa * a.q[1] + sb * b.q[1])
    akeys, bkeys = a.sorted_keys(), b.sorted_keys()
    nsec_c = [a.nsec(i) for i in ua] + [b.nsec(i) for i in ub]

    def ctag(idx):
        t = 0
        for n, x in zip(nsec_c, idx):
            t = t * n + x
        return t

    def key(ai):
        t = 0
        for x in ca:
            t = t * a.nsec(x) + ai[x]
        return t

    out = {}
    for ia, ai in enumerate(akeys):
        for ib, bi in enumerate(bkeys):
            if any(ai[x] != bi[y] for x, y in zip(ca, cb)):
                continue
            cidx = tuple(ai[i] for i in ua) + tuple(bi[i] for i in ub)
            p = sum(sa * a.qshape[i][ai[i]][0] for i in ua) + sum(sb * b.qshape[i][bi[i]][0] for i in ub)
            s = sum(sa * a.qshape[i][ai[i]][1] for i in ua) + sum(sb * b.qshape[i][bi[i]][1] for i in ub)
            if (p, s) != cq:
                continue
            out.setdefault(cidx, []).append((key(ai), ia, ib))
    blocks = sorted(out, key=ctag)
    return blocks, [[(ia, ib) for _, ia, ib in sorted(out[c])] for c in blocks]


def structure(a, b, ca, cb, conj_a, conj_b):
    da, ka = a.to_desc()
    db, kb = b.to_desc()
    ca_ = np.array(ca, dtype=np.int32)
    cb_ = np.array(cb, dtype=np.int32)
    nb, npairs = ctypes.c_int64(), ctypes.c_int64()
    f = mx.lib().mpsxx_contract_structure
    args = [ctypes.byref(da), ctypes.c_int(int(conj_a)), ca_.ctypes.data_as(VP), ctypes.byref(db),
            ctypes.c_int(int(conj_b)), cb_.ctypes.data_as(VP), ctypes.c_int32(len(ca)), ctypes.byref(nb),
            ctypes.byref(npairs)]
    mx._check(f(*args, None, None, None, None))
    rank_c = a.rank + b.rank - 2 * len(ca)
    idx = np.zeros(max(nb.value * rank_c, 1), dtype=np.int32)
    pb = np.zeros(nb.value + 1, dtype=np.int64)
    pa = np.zeros(max(npairs.value, 1), dtype=np.int64)
    pbb = np.zeros(max(npairs.value, 1), dtype=np.int64)
    mx._check(f(*args, idx.ctypes.data_as(VP), pb.ctypes.data_as(VP), pa.ctypes.data_as(VP), pbb.ctypes.data_as(VP)))
    blocks = [tuple(int(x) for x in idx[i * rank_c:(i + 1) * rank_c]) for i in range(nb.value)]
    pairs = [[(int(pa[p]), int(pbb[p])) for p in range(pb[i], pb[i + 1])] for i in range(nb.value)]
    return blocks, pairs


def set_threads(n, grain):
    mx._check(mx.lib().mpsxx_planner_threads(ctypes.c_int(n), ctypes.c_int64(grain)))


CASES = [(seed, ra, rb, k, cja, cjb) for seed, (ra, rb, k), (cja, cjb) in zip(
    range(12),
    [(3, 3, 1), (4, 3, 1), (4, 4, 2), (5, 4, 2), (3, 4, 1), (4, 4, 1), (5, 3, 1), (3, 5, 2), (4, 5, 2), (5, 5, 2),
     (4, 4, 3), (3, 3, 2)],
    [(False, False), (True, False), (False, True), (False, False), (True, False), (False, False), (False, True),
     (True, False), (False, False), (False, False), (False, False), (True, False)])]


@pytest.mark.parametrize("seed,ra,rb,k,cja,cjb", CASES)
def test_contract_structure_matches_brute_force_and_oracle(seed, ra, rb, k, cja, cjb):
    rng = np.random.default_rng(100 + seed)
    a, b, ca, cb = random_pair(rng, ra, rb, k, cja, cjb)
    exp_blocks, exp_pairs = brute_force(a, b, ca, cb, cja, cjb)
    try:
        set_threads(1, 1)
        got1 = structure(a, b, ca, cb, cja, cjb)
        set_threads(4, 1)
        got4 = structure(a, b, ca, cb, cja, cjb)
    finally:
        set_threads(0, 200000)
    assert got1 == got4
    assert got1[0] == exp_blocks
    assert got1[1] == exp_pairs
    ref = oracle.contract(1.0, a, ca, b, cb, conj_a=cja, conj_b=cjb)
    assert set(got1[0]) == ref.block_set()


def test_contract_structure_on_the_qc_sigma_path():
    """The first two Contract steps of the two-site sigma vector (compute_sigma_vector.hpp:64-66) with the QC MPO of
    a random 6-orbital problem at the middle bond: output blocks equal the oracle's, pair lists equal the brute force,
    and 1 and 4 planner threads give the same plan."""
    from helpers import random_env_left
    from mpsxx_amd import mps, qc
    K, i = 6, 2
    rng = np.random.default_rng(5)
    h, v = qc.synthetic_integrals(K, seed=3)
    W1 = qc.QcMpo(h, v).site_qst(i)
    qt = (K, 0)
    lq = mps.left_bond_qns(K, i, qt)
    rq = mps.right_bond_qns(K, i + 2, qt)
    ld = [int(rng.integers(1, 4)) if abs(q[1]) <= 2 else 0 for q in lq]
    rd = [1 if abs(q[1]) <= 2 else 0 for q in rq]
    lenv = random_env_left(lq, ld, W1, rng)
    psi = mps.fill_allowed(mps.twosite_wavefunction(lq, ld, rq, rd, qt), rng)
    t1 = oracle.contract(1.0, lenv, [2], psi, [0])
    for step, (a, ca, b, cb) in enumerate(((lenv, [2], psi, [0]), (t1, [1, 2], W1, [0, 2]))):
        # zero-size sectors carry no pairs in the numbering: drop them from this check's inputs
        a = _without_empty(a)
        b = _without_empty(b)
        exp = brute_force(a, b, ca, cb, False, False)
        try:
            set_threads(1, 1)
            g1 = structure(a, b, ca, cb, False, False)
            set_threads(4, 1)
            g4 = structure(a, b, ca, cb, False, False)
        finally:
            set_threads(0, 200000)
        assert g1 == g4 and g1 == exp
        assert set(g1[0]) == oracle.contract(1.0, a, ca, b, cb).block_set()
        if step == 1:
            assert sum(len(p) for p in g1[1]) > 2 * len(g1[0])  # the MPO step sums several pairs per block


def _without_empty(x):
    y = x.empty_like()
    y.blocks = {k: blk for k, blk in x.blocks.items() if blk.size}
    return y
