"""Device path vs the oracle on identical inputs (needs an MI355X).

Bar: identical block sets (exact), FP64 values within 1e-12 of the output scale; FP32 mode within 2e-5 (relative
Frobenius norm, written per test)."""
import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from helpers import (chain_overlaps, hubbard_problem, random_env_left, random_env_right, random_twosite_problem)
from mpsxx_amd import mps
from mpsxx_amd.qst import QST

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def D(q):
    return mx.DTensor(q)


def assert_same(gpu, ref, tol=TOL64):
    assert gpu is not None and ref is not None
    assert gpu.block_set() == ref.block_set(), "block sets differ"
    assert gpu.q == ref.q and gpu.qshape == ref.qshape and gpu.dshape == ref.dshape
    scale = max(1.0, max((float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size), default=1.0))
    err = gpu.max_abs_diff(ref)
    assert err <= tol * scale, "max |diff| %.3e (scale %.3e)" % (err, scale)


@pytest.mark.parametrize("L,i,seed,dmax", [(4, 1, 0, 2), (6, 2, 1, 3), (8, 3, 2, 4), (8, 0, 3, 3), (8, 6, 4, 3)])
def test_sigma_twosite_chain(L, i, seed, dmax):
    print("sigma_twosite_chain", L, i, seed, dmax, flush=True)
    rng = np.random.default_rng(seed)
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=dmax)
    ref = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, ch["psi"])
    got = mx.compute_sigma_vector_twosite(D(mpos[i]), D(mpos[i + 1]), D(lenv), D(renv), D(ch["psi"])).to_qst()
    assert_same(got, ref)


@pytest.mark.parametrize("seed,dlo,dhi", [(10, 1, 20), (11, 30, 90), (12, 60, 150)])
def test_sigma_twosite_large_blocks(seed, dlo, dhi):
    print("sigma_twosite_large", seed, dlo, dhi, flush=True)
    rng = np.random.default_rng(seed)
    lm, rm, le, re_, psi = random_twosite_problem(10, 4, rng, lambda q: int(rng.integers(dlo, dhi + 1)) if abs(q[1]) <= 2 else 0)
    ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
    got = mx.compute_sigma_vector_twosite(D(lm), D(rm), D(le), D(re_), D(psi)).to_qst()
    assert_same(got, ref)


def test_sigma_twosite_accumulates():
    rng = np.random.default_rng(20)
    mpos, ch, lenv, renv = hubbard_problem(6, 2, rng, dmax=3)
    ref = oracle.sigma_twosite(mpos[2], mpos[3], lenv, renv, ch["psi"])
    args = [D(x) for x in (mpos[2], mpos[3], lenv, renv, ch["psi"])]
    out = mx.compute_sigma_vector_twosite(*args)
    out = mx.compute_sigma_vector_twosite(*args, sgv0=out)
    got = out.to_qst()
    ref2 = ref.copy()
    for k in ref2.blocks:
        ref2.blocks[k] = 2 * ref2.blocks[k]
    assert_same(got, ref2)


def test_sigma_twosite_fp32():
    rng = np.random.default_rng(21)
    lm, rm, le, re_, psi = random_twosite_problem(10, 4, rng, lambda q: int(rng.integers(5, 40)))
    ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
    got = mx.compute_sigma_vector_twosite(D(lm), D(rm), D(le), D(re_), D(psi), precision=mx.FP32).to_qst()
    assert got.block_set() == ref.block_set()
    diff = np.sqrt(sum(float(np.sum((got.blocks[k] - ref.blocks[k]) ** 2)) for k in ref.blocks))
    assert diff <= 2e-5 * ref.norm()


def test_sigma_qn_mismatch_raises():
    rng = np.random.default_rng(22)
    mpos, ch, lenv, renv = hubbard_problem(6, 2, rng, dmax=2)
    wrong = ch["psi"].empty_like()
    wrong.q = (wrong.q[0] + 2, wrong.q[1])
    wrong.insert(next(iter(k for k in _allowed(wrong))), np.zeros(wrong.block_shape(next(iter(_allowed(wrong))))))
    with pytest.raises(mx.MpsxxError) as ei:
        mx.compute_sigma_vector_twosite(D(mpos[2]), D(mpos[3]), D(lenv), D(renv), D(ch["psi"]), sgv0=D(wrong))
    assert "quantum number" in str(ei.value)


def _allowed(a):
    import itertools
    for idx in itertools.product(*[range(a.nsec(k)) for k in range(a.rank)]):
        if a.allowed(idx):
            yield idx


def test_sigma_onesite_and_diag_onesite():
    L, i = 6, 2
    rng = np.random.default_rng(30)
    mpos = mx.mpo.hubbard_mpo(L, 1.0, 4.0)
    qt = (L, 0)
    lq, rq = mps.left_bond_qns(L, i, qt), mps.right_bond_qns(L, i + 1, qt)
    ld = [int(rng.integers(1, 6)) for _ in lq]
    rd = [int(rng.integers(1, 6)) for _ in rq]
    le = random_env_left(lq, ld, mpos[i], rng)
    re_ = random_env_right(rq, rd, mpos[i], rng)
    wf = mps.fill_allowed(mps.site_tensor(lq, ld, rq, rd, total=qt), rng)
    ref = oracle.sigma_onesite(mpos[i], le, re_, wf)
    got = mx.compute_sigma_vector_onesite(D(mpos[i]), D(le), D(re_), D(wf)).to_qst()
    assert_same(got, ref)
    dref = oracle.diag_onesite(mpos[i], le, re_, wf.empty_like())
    dg = mx.compute_diagonal_elements_onesite(D(mpos[i]), D(le), D(re_), D(wf.empty_like())).to_qst()
    assert_same(dg, dref)


@pytest.mark.parametrize("seed", [40, 41])
def test_diag_twosite(seed):
    rng = np.random.default_rng(seed)
    lm, rm, le, re_, psi = random_twosite_problem(8, 3, rng, lambda q: int(rng.integers(1, 30)))
    dref = oracle.diag_twosite(lm, rm, le, re_, psi.empty_like())
    dg = mx.compute_diagonal_elements_twosite(D(lm), D(rm), D(le), D(re_), D(psi.empty_like())).to_qst()
    assert_same(dg, dref)


def _merged_problem(rng, nsec=5, dmax=40):
    """Random merged operators L~(Lb,w,Lk), R~(Rb,w,Rk) and psi~(Lk,Rk) over shared operator sectors."""
    qL = [(p, s) for p in range(2, 7) for s in (-1, 1) if (p + s) % 2 == 0][:nsec] or [(2, 0)]
    qL = [(p, s) for p in range(0, 6) for s in range(-2, 3) if (p + s) % 2 == 0][:nsec * 2]
    qt = (6, 0)
    qR = [(p - qt[0], s - qt[1]) for (p, s) in qL]  # psi total: qL + (-(-qR)) ... q_l - q_r = qt
    dL = [int(rng.integers(1, dmax)) for _ in qL]
    dR = [int(rng.integers(1, dmax)) for _ in qR]
    ops = [(0, 0), (1, 1), (1, -1), (-1, -1), (-1, 1), (0, 0), (2, 0)]
    Lm = mps.fill_allowed(QST((0, 0), [qL, [(-a, -b) for a, b in ops], [(-a, -b) for a, b in qL]],
                              [dL, [1] * len(ops), dL]), rng)
    Rm = mps.fill_allowed(QST((0, 0), [[(-a, -b) for a, b in qR], ops, qR], [dR, [1] * len(ops), dR]), rng)
    psi = mps.fill_allowed(QST(qt, [qL, [(-a, -b) for a, b in qR]], [dL, dR]), rng)
    return Lm, Rm, psi


def test_sigma_and_diag_merged():
    rng = np.random.default_rng(50)
    Lm, Rm, psi = _merged_problem(rng)
    assert psi.blocks
    ref = oracle.sigma_merged(Lm, Rm, psi)
    got = mx.compute_sigma_vector_merged(D(Lm), D(Rm), D(psi)).to_qst()
    assert_same(got, ref)
    dref = oracle.diag_merged(Lm, Rm, psi.empty_like())
    dg = mx.compute_diagonal_elements_merged(D(Lm), D(Rm), D(psi.empty_like())).to_qst()
    assert_same(dg, dref)


@pytest.mark.parametrize("forward", [True, False])
def test_renormalize(forward):
    rng = np.random.default_rng(60 + forward)
    L, i = 8, 3
    mpos = mx.mpo.hubbard_mpo(L, 1.0, 4.0)
    qt = (L, 0)
    if forward:
        lq, nq = mps.left_bond_qns(L, i, qt), mps.left_bond_qns(L, i + 1, qt)
        ld = [int(rng.integers(1, 40)) for _ in lq]
        nd = [int(rng.integers(1, 40)) for _ in nq]
        env = random_env_left(lq, ld, mpos[i], rng)
        A = mps.fill_allowed(mps.site_tensor(lq, ld, nq, nd), rng)
    else:
        lq, nq = mps.right_bond_qns(L, i, qt), mps.right_bond_qns(L, i + 1, qt)
        ld = [int(rng.integers(1, 40)) for _ in lq]
        nd = [int(rng.integers(1, 40)) for _ in nq]
        env = random_env_right(nq, nd, mpos[i], rng)
        A = mps.fill_allowed(mps.site_tensor(lq, ld, nq, nd), rng)
    ref = oracle.renormalize(forward, mpos[i], env, A, A)
    got = mx.renormalize(forward, D(mpos[i]), D(env), D(A), D(A)).to_qst()
    assert_same(got, ref)


@pytest.mark.parametrize("forward", [True, False])
def test_renormalize_overlap_and_overlap_sigma(forward):
    rng = np.random.default_rng(70)
    L, i = 6, 2
    ch = mps.random_chain(L, i, (L, 0), rng, dmax=4)
    lsov, rsov = chain_overlaps(ch, i, L)
    ref = oracle.sigma_overlap(lsov, rsov, ch["psi"])
    got = mx.compute_sigma_vector_overlap(D(lsov), D(rsov), D(ch["psi"])).to_qst()
    assert_same(got, ref)
    if forward:
        A = ch["A"][1]
        s0 = oracle.renormalize_overlap(True, mps.boundary_overlap(), ch["A"][0], ch["A"][0])
        r1 = oracle.renormalize_overlap(True, s0, A, A)
        g1 = mx.renormalize_overlap(True, D(s0), D(A), D(A)).to_qst()
    else:
        B = ch["B"][4]
        s0 = oracle.renormalize_overlap(False, mps.boundary_overlap(), ch["B"][5], ch["B"][5])
        r1 = oracle.renormalize_overlap(False, s0, B, B)
        g1 = mx.renormalize_overlap(False, D(s0), D(B), D(B)).to_qst()
    assert_same(g1, r1)


@pytest.mark.parametrize("forward", [True, False])
def test_renormalize_merged(forward):
    rng = np.random.default_rng(80 + forward)
    Lm, Rm, psi = _merged_problem(rng, dmax=30)
    # bra/ket: a 2-index matrix mapping the operator's bond onto a new bond
    if forward:
        op = Lm
        qb = op.qshape[0]
        qn = [(p, s) for (p, s) in qb][:4]
        U = mps.fill_allowed(QST((0, 0), [qb, [(-a, -b) for a, b in qn]], [op.dshape[0], [int(rng.integers(1, 20)) for _ in qn]]), rng)
    else:
        op = Rm
        qb = op.qshape[0]
        qn = [(-a, -b) for a, b in qb][:4]
        U = mps.fill_allowed(QST((0, 0), [qn, qb], [[int(rng.integers(1, 20)) for _ in qn], op.dshape[0]]), rng)
    ref = oracle.renormalize_merged(forward, op, U, U)
    got = mx.renormalize_merged(forward, D(op), D(U), D(U)).to_qst()
    assert_same(got, ref)


def test_contract_generic_permutations():
    rng = np.random.default_rng(90)
    lm, rm, le, re_, psi = random_twosite_problem(8, 3, rng, lambda q: int(rng.integers(1, 25)))
    cases = [((le, [2], psi, [0]), False, False),
             ((psi, [3], re_, [2]), False, False),
             ((psi, [0, 3], psi, [0, 3]), False, True),
             ((le, [0, 1], le, [0, 1]), False, True)]
    for (a, ca, b, cb), cja, cjb in cases:
        try:
            ref = oracle.contract(1.0, a, ca, b, cb, conj_a=cja, conj_b=cjb)
        except RuntimeError:
            with pytest.raises(mx.MpsxxError):
                mx.contract(1.0, D(a), ca, D(b), cb, conj_a=cja, conj_b=cjb)
            continue
        got = mx.contract(1.0, D(a), ca, D(b), cb, conj_a=cja, conj_b=cjb).to_qst()
        assert_same(got, ref)


def test_blas1():
    rng = np.random.default_rng(100)
    lm, rm, le, re_, psi = random_twosite_problem(8, 3, rng, lambda q: int(rng.integers(1, 25)))
    x = psi
    y = psi.empty_like()
    keys = sorted(psi.blocks, key=psi.tag)
    for k in keys[::2]:
        y.blocks[k] = rng.standard_normal(psi.blocks[k].shape)
    dx, dy = D(x), D(y)
    ref = sum(float(np.sum(x.blocks[k] * y.blocks[k])) for k in y.blocks)
    assert abs(mx.dotc(dx, dy) - ref) <= 1e-12 * max(1.0, abs(ref))
    assert abs(mx.nrm2(dx) - x.norm()) <= 1e-12 * x.norm()
    mx.axpy(0.5, dx, dy)
    got = dy.to_qst()
    assert got.block_set() == x.block_set()
    for k in x.blocks:
        exp = 0.5 * x.blocks[k] + y.blocks.get(k, 0)
        assert np.max(np.abs(got.blocks[k] - exp)) < 1e-13


def test_davidson_matches_oracle_and_dense():
    rng = np.random.default_rng(110)
    L, i = 6, 2
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=3)
    psi = ch["psi"]
    diag = oracle.diag_twosite(mpos[i], mpos[i + 1], lenv, renv, psi.empty_like())
    e_ref, w_ref, ns_ref = oracle.davidson_twosite([(mpos[i], mpos[i + 1], lenv, renv)], diag, psi, 1e-11)
    import ctypes
    lib = mx.lib()
    g = [D(mpos[i]), D(mpos[i + 1]), D(lenv), D(renv)]
    arr = [(ctypes.c_void_p * 1)(x.h) for x in g]
    dd, dw = D(diag), D(psi)
    e = ctypes.c_double()
    ns = ctypes.c_int32()
    rc = lib.mpsxx_davidson_twosite(1, arr[0], arr[1], arr[2], arr[3], dd.h, dw.h, ctypes.c_double(1e-11),
                                    0, ctypes.byref(e), ctypes.byref(ns), None)
    assert rc == 0, lib.mpsxx_last_error()
    assert abs(e.value - e_ref) < 1e-10 * max(1.0, abs(e_ref))
    # the Rayleigh quotient of the returned vector equals the eigenvalue
    w = dw.to_qst()
    s = mx.compute_sigma_vector_twosite(g[0], g[1], g[2], g[3], dw).to_qst()
    rq = sum(float(np.sum(w.blocks[k] * s.blocks.get(k, 0))) for k in w.blocks)
    assert abs(rq - e.value) < 1e-8 * max(1.0, abs(e.value))


@pytest.mark.parametrize("seed,dlo,dhi", [(12, 60, 150)])
def test_sigma_twosite_split_bf16(seed, dlo, dhi):
    """MPSXX_BF16X3 on blocks above the small-tile size: 1e-4 relative Frobenius norm against the oracle."""
    rng = np.random.default_rng(seed)
    lm, rm, le, re_, psi = random_twosite_problem(10, 4, rng, lambda q: int(rng.integers(dlo, dhi + 1)) if abs(q[1]) <= 2 else 0)
    ref = oracle.sigma_twosite(lm, rm, le, re_, psi)
    got = mx.compute_sigma_vector_twosite(D(lm), D(rm), D(le), D(re_), D(psi), precision=mx.BF16X3).to_qst()
    assert got.block_set() == ref.block_set()
    assert got.diff_norm(ref) <= 1e-4 * ref.norm()


def _dense_heff(mpos, i, lenv, renv, psi):
    """H_eff on psi's blocks (all allowed blocks), column by column from the oracle's sigma of unit vectors."""
    keys = psi.sorted_keys()
    sizes = [psi.blocks[k].size for k in keys]
    n = sum(sizes)
    H = np.zeros((n, n))
    col = 0
    for k, sz in zip(keys, sizes):
        for e in range(sz):
            u = psi.empty_like()
            for k2 in keys:
                u.blocks[k2] = np.zeros(psi.blocks[k2].shape)
            u.blocks[k].flat[e] = 1.0
            s = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, u)
            H[:, col] = np.concatenate([s.blocks[k2].ravel() if k2 in s.blocks else np.zeros(psi.blocks[k2].size)
                                        for k2 in keys])
            col += 1
    return H, keys


def _from_flat(psi, keys, v):
    out = psi.empty_like()
    off = 0
    for k in keys:
        sz = psi.blocks[k].size
        out.blocks[k] = v[off:off + sz].reshape(psi.blocks[k].shape).copy()
        off += sz
    return out


def _orthonormal_problem(L, i, seed):
    """Two-site problem whose left and right blocks are single sites with identity site tensors (every sector of
    dimension 1, every block 1.0), so that H_eff is the chain Hamiltonian itself in the (N, 2Sz) sector."""
    from helpers import chain_envs
    from mpsxx_amd.mpo import hubbard_mpo
    rng = np.random.default_rng(seed)
    mpos = hubbard_mpo(L, 1.0, 4.0)
    ch = mps.random_chain(L, i, (L, L % 2), rng, dims=lambda q: 1)
    for t in list(ch["A"]) + list(ch["B"].values()):
        for k in t.blocks:
            t.blocks[k] = np.ones(t.blocks[k].shape)
    lenv, renv = chain_envs(mpos, ch, i, L)
    return mpos, ch, lenv, renv


def test_davidson_lower_root_projection_matches_dense():
    """Excited states (src/diagonalize.hpp:36-89): with the exact ground state v0 as lower root (H|0> = E0 v0,
    S|0> = v0), the device Davidson converges to the lowest eigenvalue of H - E0 v0 v0^T, computed densely from the
    oracle's H_eff (here the L = 4 Hubbard Hamiltonian at half filling), and returns a normalized vector orthogonal to
    v0."""
    print("davidson lower root", flush=True)
    L, i = 4, 1
    mpos, ch, lenv, renv = _orthonormal_problem(L, i, 111)
    psi = ch["psi"]
    H, keys = _dense_heff(mpos, i, lenv, renv, psi)
    assert np.max(np.abs(H - H.T)) < 1e-12
    w, V = np.linalg.eigh(H)
    v0, e0 = V[:, 0], w[0]
    expected = np.linalg.eigvalsh(H - e0 * np.outer(v0, v0))[0]
    g = (D(mpos[i]), D(mpos[i + 1]), D(lenv), D(renv))
    diag = oracle.diag_twosite(mpos[i], mpos[i + 1], lenv, renv, psi.empty_like())
    lower = [(D(_from_flat(psi, keys, H @ v0)), D(_from_flat(psi, keys, v0)), e0)]
    dw = D(psi)
    e, ns = mx.davidson_twosite([g], D(diag), dw, 1e-14, lower=lower)
    assert abs(e - expected) < 1e-10 * max(1.0, abs(expected)), (e, expected, w[:3])
    got = dw.to_qst()
    x = np.concatenate([got.blocks[k].ravel() if k in got.blocks else np.zeros(psi.blocks[k].size) for k in keys])
    assert abs(np.dot(x, v0)) < 1e-8 and abs(np.linalg.norm(x) - 1.0) < 1e-12
    # without the lower root the same call finds E0
    e_gs, _ = mx.davidson_twosite([g], D(diag), D(psi), 1e-14)
    assert abs(e_gs - e0) < 1e-10 * max(1.0, abs(e0))


def test_davidson_merged_groups_match_fourindex():
    """Merged groups (optimize_twosite_merged): Davidson on the merged sigma vector of the same problem reaches the
    4-index Davidson's eigenvalue."""
    print("davidson merged groups", flush=True)
    rng = np.random.default_rng(112)
    L, i = 6, 2
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=3)
    psi = ch["psi"]
    g = (D(mpos[i]), D(mpos[i + 1]), D(lenv), D(renv))
    diag = oracle.diag_twosite(mpos[i], mpos[i + 1], lenv, renv, psi.empty_like())
    e4, _ = mx.davidson_twosite([g], D(diag), D(psi), 1e-12)
    lm = mx.compute_merged_operator(True, g[0], g[2])
    rm = mx.compute_merged_operator(False, g[1], g[3])
    wm = mx.qst_merge(D(psi), 2, 2)
    dm = mx.DTensor(wm.to_qst().empty_like())
    mx.compute_diagonal_elements_merged(lm, rm, dm)
    em, _ = mx.davidson_twosite([(None, None, lm, rm)], dm, wm, 1e-12)
    assert abs(em - e4) < 1e-10 * max(1.0, abs(e4)), (em, e4)
