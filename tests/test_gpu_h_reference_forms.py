"""Device forms of the reference's conjugate sigma vector, make_noise and canonicalize with noise and either truncation
rule, against the oracle (needs an MI355X).

Bars: block sets exact; sigma within 1e-12 of the output scale; noise within 1e-15; canonicalize: same kept sectors
and dimensions, discarded weight within 1e-12, reconstructions within 1e-11 of the input scale."""
import ctypes

import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from mpsxx_amd import mps
from mpsxx_amd.qst import QST
from test_oracle_pins_merged import _onesite_problem, prescribed_wavefunction

pytestmark = pytest.mark.gpu


def test_sigma_conj_matches_oracle():
    print("sigma conj", flush=True)
    W, lenv, renv, wf, PL, PR, H = _onesite_problem(11)
    wt = QST(wf.q, [wf.qshape[1], wf.qshape[2], wf.qshape[0]], [wf.dshape[1], wf.dshape[2], wf.dshape[0]])
    for (a, c, d), blk in wf.blocks.items():
        wt.blocks[(c, d, a)] = blk.transpose(1, 2, 0).copy()
    ref = oracle.sigma_conj(W, lenv, renv, wt)
    D = mx.DTensor
    out = mx.compute_sigma_vector_conj(D(W), D(lenv), D(renv), D(wt))
    got = out.to_qst()
    assert got.q == ref.q and got.qshape == ref.qshape and got.block_set() == ref.block_set()
    scale = max(float(np.max(np.abs(b))) for b in ref.blocks.values())
    assert got.max_abs_diff(ref) <= 1e-12 * scale
    # accumulation into the non-empty output
    got2 = mx.compute_sigma_vector_conj(D(W), D(lenv), D(renv), D(wt), out).to_qst()
    assert got2.max_abs_diff(ref.scaled(2.0)) <= 2e-12 * scale


@pytest.mark.parametrize("noise", [5e-9, 1e-3])
def test_make_noise_matches_oracle(noise):
    print("make_noise", noise, flush=True)
    rng = np.random.default_rng(12)
    ch = mps.random_chain(8, 3, (8, 0), rng, dmax=6)
    psi = ch["psi"]
    for k in psi.sorted_keys()[::4]:
        del psi.blocks[k]
    ref = oracle.make_noise(psi, noise)
    got = mx.make_noise(mx.DTensor(psi), noise).to_qst()
    assert got.block_set() == ref.block_set()
    assert got.max_abs_diff(ref) <= 1e-15 * max(1.0, max(float(np.max(np.abs(b))) for b in ref.blocks.values()))


def _check_canon(forward, psi, M, noise, rule):
    U, V, dw = mx.canonicalize_twosite(forward, mx.DTensor(psi), M, noise, rule)
    U, V = U.to_qst(), V.to_qst()
    Ur, Vr, dwr = oracle.canonicalize_twosite_ex(forward, psi, M, noise, rule)
    assert U.qshape == Ur.qshape and U.dshape == Ur.dshape and U.block_set() == Ur.block_set()
    assert V.qshape == Vr.qshape and V.dshape == Vr.dshape and V.block_set() == Vr.block_set()
    assert abs(dw - dwr) < 1e-12 * max(1.0, psi.norm() ** 2)
    rec = oracle.contract(1.0, U, [2], V, [0])
    recr = oracle.contract(1.0, Ur, [2], Vr, [0])
    scale = max(float(np.max(np.abs(b))) for b in psi.blocks.values())
    assert rec.max_abs_diff(recr) < 1e-11 * scale
    return U


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("forward", [True, False])
def test_canonicalize_rules_on_a_near_tie_match_oracle(rule, forward):
    """Merged blocks with singular values {1.0, 0.5} and {0.5 (1 - 1e-13), 0.2}, M = 2: the reference rule keeps
    two values, the stable rule three, on the device as in the oracle."""
    print("canonicalize near tie", rule, forward, flush=True)
    psi, _ = prescribed_wavefunction(np.random.default_rng(13), [[1.0, 0.5], [0.5 * (1.0 - 1e-13), 0.2]])
    U = _check_canon(forward, psi, 2, 0.0, rule)
    assert sum(U.dshape[2]) == (2 if rule == 0 else 3)


@pytest.mark.parametrize("rule", [0, 1])
def test_canonicalize_with_noise_matches_oracle(rule):
    """canonicalize(forward, wfnc, lmps, rmps, M, noise): make_noise on wfnc (in place, as the reference), then the
    truncated SVD; random psi with some blocks absent (the noise inserts them), M = 12, noise 1e-3."""
    print("canonicalize noise", rule, flush=True)
    rng = np.random.default_rng(5)
    ch = mps.random_chain(8, 3, (8, 0), rng, dmax=5)
    psi = ch["psi"]
    for k in psi.sorted_keys()[::5]:
        del psi.blocks[k]
    for forward in (True, False):
        _check_canon(forward, psi, 12, 1e-3, rule)
    w = mx.DTensor(psi)
    mx.canonicalize_twosite(True, w, 12, 1e-3, rule)
    assert w.to_qst().max_abs_diff(oracle.make_noise(psi, 1e-3)) <= 1e-15 * 10  # wfnc carries the noise afterwards


def test_sweeps_with_noise_reach_the_ground_state():
    """DMRGInput::noise: two sweeps with noise 1e-4 before every canonicalize, then noise-free sweeps, reach the
    exact ground energy of the L = 8 Hubbard chain (config 1 model) to 1e-9."""
    import ed
    from mpsxx_amd.dmrg import DeviceDMRG, initial_mps
    from mpsxx_amd.mpo import hubbard_mpo
    print("sweeps with noise", flush=True)
    L, M = 8, 64
    dm = DeviceDMRG(hubbard_mpo(L, 1.0, 4.0), initial_mps(L, (L, 0), M, 0), M)
    dm.set_truncation(1e-4, 1)
    for _ in range(2):
        dm.sweep(M, 1e-10)
    dm.set_truncation(0.0, 1)
    e = None
    for _ in range(4):
        e, st = dm.sweep(M, 1e-10)
    h, v = ed.hubbard_integrals(L, 1.0, 4.0)
    e0 = ed.ground_energy(h, v, L, L, 0)
    assert abs(e - e0) <= 1e-9, (e, e0)


def _svd_discarded(blocks, M):
    """Discarded weight of the reference rule (QSTLAPACK.h:209-222) over the singular values of the given blocks."""
    s = np.sort(np.concatenate([np.linalg.svd(b, compute_uv=False) for b in blocks]))[::-1]
    cut = s[M - 1] if 0 < M < s.size else 1e-16
    return float(np.sum(s[s < cut] ** 2))


@pytest.mark.parametrize("forward", [True, False])
def test_canonicalize_onesite_reference_semantics(forward):
    """One-site canonicalize (canonicalize.hpp:49-85): forward lmps = U, rmps <- s Vt rmps; backward rmps = Vt,
    lmps <- lmps U s.  With M = 0 the product lmps rmps equals wfnc rmps (lmps wfnc) and the new canonical factor is
    an isometry; with M = 6 the discarded weight is the reference rule's over numpy's singular values of the merged
    (l, n) x r (l x (n, r)) blocks."""
    from test_oracle_pins_merged import np_merge
    print("canonicalize onesite", forward, flush=True)
    rng = np.random.default_rng(31)
    L, i, qt = 8, 3, (8, 0)
    ch = mps.random_chain(L, i, qt, rng, dmax=4)
    rq = mps.right_bond_qns(L, i + 1, qt)
    rd = [int(rng.integers(1, 5)) for _ in rq]
    wf = mps.fill_allowed(mps.site_tensor(ch["lq"][i], ch["ld"][i], rq, rd, total=qt), rng)
    nb = mps.fill_allowed(mps.site_tensor(rq, rd, ch["rq"][i + 2], ch["rd"][i + 2]), rng)  # right neighbour
    pb = ch["A"][i - 1]  # left neighbour
    D = mx.DTensor
    for M in (0, 6):
        w, l, r = D(wf), D(pb), D(nb)
        d = ctypes.c_double()
        rc = mx.lib().mpsxx_canonicalize_onesite_ex(int(forward), w.h, l.h, r.h, M, ctypes.c_double(0.0), 0,
                                                    ctypes.byref(d), None)
        assert rc == 0, mx.lib().mpsxx_last_error()
        lq, rq_ = l.to_qst(), r.to_qst()
        if M == 0:
            rec = oracle.contract(1.0, lq, [2], rq_, [0])
            ref = oracle.contract(1.0, wf, [2], nb, [0]) if forward else oracle.contract(1.0, pb, [2], wf, [0])
            assert rec.max_abs_diff(ref) < 1e-11 * max(float(np.max(np.abs(b))) for b in ref.blocks.values())
            iso = oracle.contract(1.0, lq, [0, 1], lq, [0, 1], conj_a=True) if forward else \
                oracle.contract(1.0, rq_, [1, 2], rq_, [1, 2], conj_b=True)
            for (a, b), blk in iso.blocks.items():
                assert np.max(np.abs(blk - (np.eye(blk.shape[0]) if a == b else 0.0))) < 1e-12
        mat = np_merge(wf, 2, 1) if forward else np_merge(wf, 1, 2)
        assert abs(d.value - _svd_discarded(list(mat.blocks.values()), M)) < 1e-12 * wf.norm() ** 2


@pytest.mark.parametrize("forward", [True, False])
def test_canonicalize_merged_equals_twosite_on_the_expanded_wavefunction(forward):
    """Merged canonicalize (canonicalize.hpp:88-106) of merge(psi) keeps the spectrum of the two-site form of psi:
    same discarded weight, and lmat rmat equals merge(U V) of the oracle's two-site canonicalize (reference rule)."""
    from test_oracle_pins_merged import np_merge
    print("canonicalize merged", forward, flush=True)
    rng = np.random.default_rng(32)
    ch = mps.random_chain(8, 3, (8, 0), rng, dmax=5)
    psi = ch["psi"]
    M = 10
    D = mx.DTensor
    w, l, r = D(np_merge(psi, 2, 2)), D(), D()
    d = ctypes.c_double()
    rc = mx.lib().mpsxx_canonicalize_merged_ex(int(forward), w.h, l.h, r.h, M, ctypes.c_double(0.0), 0,
                                               ctypes.byref(d), None)
    assert rc == 0, mx.lib().mpsxx_last_error()
    Ur, Vr, dwr = oracle.canonicalize_twosite_ex(forward, psi, M, 0.0, 0)
    assert abs(d.value - dwr) < 1e-12 * max(1.0, psi.norm() ** 2)
    rec = oracle.contract(1.0, l.to_qst(), [1], r.to_qst(), [0])
    ref = np_merge(oracle.contract(1.0, Ur, [2], Vr, [0]), 2, 2)
    assert rec.max_abs_diff(ref) < 1e-11 * max(float(np.max(np.abs(b))) for b in psi.blocks.values())
