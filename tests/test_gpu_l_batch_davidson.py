"""Batched Davidson (mpsxx_batch_davidson_twosite) against the single-instance device Davidson and exact
diagonalization (needs an MI355X).

Instances are two-site problems of open Hubbard chains whose environments come from random isometric MPS
(workloads.hubbard_isometric_problem), so each H_eff = P^T H P is Hermitian with a variational lowest eigenvalue, and
start vectors near the lowest eigenvector, as in a sweep.

Bar: FP64 batched energies within 1e-10 (absolute, energies of order 10) of mpsxx_davidson_twosite on the same inputs
and eigenvectors with |<a|b>| >= 1 - 1e-7; graph replay bit-identical to direct launches; FP32 and split-BF16x3
mixed precision within 1e-8 of FP64; with bond spaces that hold every state of their side (P unitary) the energy
equals the exact ground energy of the chain to 1e-9."""
import numpy as np
import pytest

import ed
import mpsxx_amd as mx
from mpsxx_amd import workloads
from mpsxx_amd.batch import batch_davidson_twosite

pytestmark = pytest.mark.gpu

TOL = 1e-12  # ||r||^2 bar of every solve


def problems(specs):
    return [workloads.hubbard_isometric_problem(L, M, U, seed=k) for k, (L, M, U) in enumerate(specs)]


def single(p, prec=mx.FP64, calls=1):
    """mpsxx_davidson_twosite, called again from its own result up to `calls` times until the energy is stationary
    (each call is the reference's at most 4 x 10 sigma vectors, as successive sweeps restart it)."""
    w = mx.copy(p[5])
    e_prev = None
    for _ in range(calls):
        e, ns = mx.davidson_twosite([(p[0], p[1], p[2], p[3])], p[4], w, TOL, prec)
        if e_prev is not None and abs(e - e_prev) < 1e-13:
            break
        e_prev = e
    return e, w


def batched(ps, prec, use_graph=True, calls=1):
    """The batched solve, called again from its results until every instance converged (at most `calls` times)."""
    ws = [mx.copy(p[5]) for p in ps]
    ns_tot = np.zeros(len(ps), dtype=np.int64)
    nl_tot = np.zeros(len(ps), dtype=np.int64)
    for _ in range(calls):
        e, ns, nl, st = batch_davidson_twosite([p[:5] + (w,) for p, w in zip(ps, ws)], TOL, prec, use_graph)
        ns_tot += ns
        nl_tot += nl
        if st["converged"] == len(ps):
            break
    return e, ns_tot, nl_tot, st, ws


def overlap(a, b):
    qa, qb = a.to_qst(), b.to_qst()
    return sum(float(np.vdot(qa.blocks[k], qb.blocks[k])) for k in qa.blocks if k in qb.blocks)


SPECS = [(8, 16, 4.0), (10, 24, 2.0), (12, 32, 8.0), (8, 8, 2.0), (12, 48, 4.0), (10, 12, 8.0), (14, 40, 2.0)]


def test_batch_fp64_matches_single_path_and_graph_is_bitwise():
    print("batch davidson fp64", flush=True)
    ps = problems(SPECS)
    ref = [single(p, calls=10) for p in ps]
    e, ns, nl, st, ws = batched(ps, mx.FP64, True, calls=10)
    assert st["converged"] == len(ps) and st["batch_sigmas_low"] == 0 and (nl == 0).all()
    for k, (er, wr) in enumerate(ref):
        assert abs(e[k] - er) < 1e-10, (k, e[k], er)
        ov = abs(overlap(ws[k], wr))
        assert ov > 1 - 1e-7, (k, ov)
        assert abs(overlap(ws[k], ws[k]) - 1.0) < 1e-12
        assert ns[k] >= 2
    # one call from the same start: graph replay and direct launches give the same bits
    e1, ns1, _, _, ws1 = batched(ps, mx.FP64, True)
    e2, ns2, _, _, ws2 = batched(ps, mx.FP64, False)
    assert np.array_equal(e1, e2) and np.array_equal(ns1, ns2)
    for a, b in zip(ws1, ws2):
        assert a.to_qst().max_abs_diff(b.to_qst()) == 0.0
    # one call of the batch follows each instance's single-instance trajectory (same algorithm, same bar on ||r||^2)
    for k, p in enumerate(ps):
        es, _ = single(p)
        assert abs(e1[k] - es) < 1e-9 * max(1.0, abs(es)), (k, e1[k], es)


@pytest.mark.parametrize("prec", [mx.FP32, mx.BF16X3])
def test_batch_mixed_precision_within_1e8_of_fp64(prec):
    print("batch davidson mixed", prec, flush=True)
    ps = problems(SPECS)
    e64, _, _, st64, _ = batched(ps, mx.FP64, True, calls=10)
    assert st64["converged"] == len(ps)
    e, ns, nl, st, ws = batched(ps, prec, True, calls=10)
    assert st["converged"] == len(ps)
    assert np.max(np.abs(e - e64)) < 1e-8, np.abs(e - e64)
    assert (nl > 0).any()


def test_batch_full_space_equals_exact_ground_energy():
    """Bond spaces holding every state of their side make P unitary: the two-site problem is the whole chain and its
    lowest eigenvalue is the exact ground energy (exact diagonalization in the N = L, 2Sz = 0 sector)."""
    print("batch davidson full space", flush=True)
    specs = [(6, 4096, 4.0), (8, 4096, 2.0), (8, 4096, 8.0)]
    ps = problems(specs)
    e, _, _, st, _ = batched(ps, mx.FP64, True, calls=10)
    assert st["converged"] == len(ps)
    for (L, _, U), ek in zip(specs, e):
        h, v = ed.hubbard_integrals(L, 1.0, U)
        e0 = ed.ground_energy(h, v, L, L, 0)
        assert abs(ek - e0) < 1e-9, (L, U, ek, e0)


def test_batch_rejects_bad_instances():
    ps = problems(SPECS[:2])
    bad = list(ps[0][:5]) + [mx.DTensor()]
    with pytest.raises(mx.MpsxxError):
        batch_davidson_twosite([tuple(bad)], TOL, mx.FP64)
    with pytest.raises(mx.MpsxxError):
        batch_davidson_twosite([ps[1][:5] + (mx.copy(ps[1][5]),)], TOL, 7)
