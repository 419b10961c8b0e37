"""MPS quantum-number bookkeeping and random MPS / wavefunction / boundary-environment builders.

Conventions follow the reference's sweep (src/make_random_mpss.cpp:20-94, 150-171, 204-263; src/canonicalize.hpp):
  * left-canonical site tensor A(l, n, r): quantum shape (q_l, q_n, -q_r), total 0, so q_r = q_l + q_n
    (bond quantum number = particles and 2Sz accumulated from the left);
  * right-canonical site tensor B(l, n, r): same shape rule with total 0 and right-boundary bond (0,0), so a bond
    to the right of the centre carries minus the quantum numbers of everything to its right;
  * two-site wavefunction psi(l, n1, n2, r): quantum shape (q_l, q_n, q_n, -q_r) with total q_t;
  * boundary environments: one (0,0,0) block equal to 1 (make_random_mpss.cpp:258-263).
"""
import numpy as np

from .mpo import SITE_Q
from .qst import QST


def _feasible(p, s, nsites):
    """Can `nsites` sites hold p particles with 2Sz = s?"""
    if p < 0 or p > 2 * nsites:
        return False
    if (p + s) % 2:
        return False
    na = (p + s) // 2
    nb = (p - s) // 2
    return 0 <= na <= nsites and 0 <= nb <= nsites


def left_bond_qns(L, k, qt):
    """Sectors of the bond after k sites (left block), consistent with total qt on L sites; sorted (p, then s)."""
    out = []
    for p in range(0, 2 * k + 1):
        for s in range(-p, p + 1):
            if _feasible(p, s, k) and _feasible(qt[0] - p, qt[1] - s, L - k):
                out.append((p, s))
    return out


def right_bond_qns(L, k, qt):
    """Sectors of the bond before site k when everything right of it is right-canonical: minus the quantum numbers
    of sites k..L-1."""
    return [(-p, -s) for (p, s) in left_bond_qns(L, L - k, (qt[0], qt[1]))[::-1]
            if _feasible(qt[0] - p, qt[1] - s, k)]


def boundary_env():
    e = QST((0, 0), [[(0, 0)]] * 3, [[1]] * 3)
    e.insert((0, 0, 0), np.ones((1, 1, 1)))
    return e


def boundary_overlap():
    e = QST((0, 0), [[(0, 0)]] * 2, [[1]] * 2)
    e.insert((0, 0), np.ones((1, 1)))
    return e


def site_tensor(lq, ld, rq, rd, total=(0, 0)):
    """Empty A(l, n, r) with quantum shape (lq, site, -rq)."""
    return QST(total, [list(lq), SITE_Q, [(-a, -b) for (a, b) in rq]], [list(ld), [1] * 4, list(rd)])


def fill_allowed(a, rng, scale=1.0):
    import itertools
    for idx in itertools.product(*[range(a.nsec(i)) for i in range(a.rank)]):
        if a.allowed(idx):
            a.blocks[tuple(idx)] = rng.standard_normal(a.block_shape(idx)) * scale
    return a


def twosite_wavefunction(lq, ld, rq, rd, qt):
    """Empty psi(l, n1, n2, r) with quantum shape (lq, site, site, -rq) and total qt."""
    return QST(qt, [list(lq), SITE_Q, SITE_Q, [(-a, -b) for (a, b) in rq]], [list(ld), [1] * 4, [1] * 4, list(rd)])


def random_chain(L, i, qt, rng, dmax=3, dims=None):
    """Random MPS around the two-site centre (i, i+1): left tensors A_0..A_{i-1}, right tensors B_{i+2}..B_{L-1}
    and psi.  Returns dict with 'A', 'B', 'psi', 'lq', 'rq' (bond sector lists) and dims."""
    lbonds = [left_bond_qns(L, k, qt) for k in range(0, i + 1)]
    rbonds = {k: right_bond_qns(L, k, qt) for k in range(i + 2, L + 1)}
    rbonds[L] = [(0, 0)]

    def dims_for(qs):
        if dims is not None:
            return [dims(q) for q in qs]
        return [int(rng.integers(1, dmax + 1)) for _ in qs]

    ld = [[1]] + [dims_for(lbonds[k]) for k in range(1, i + 1)]
    rd = {k: dims_for(rbonds[k]) for k in rbonds}
    rd[L] = [1]
    A = []
    for k in range(i):
        a = site_tensor(lbonds[k], ld[k], lbonds[k + 1], ld[k + 1])
        A.append(fill_allowed(a, rng))
    B = {}
    for k in range(i + 2, L):
        b = site_tensor(rbonds[k], rd[k], rbonds[k + 1], rd[k + 1])
        B[k] = fill_allowed(b, rng)
    psi = fill_allowed(twosite_wavefunction(lbonds[i], ld[i], rbonds[i + 2], rd[i + 2], qt), rng)
    return dict(A=A, B=B, psi=psi, lq=lbonds, rq=rbonds, ld=ld, rd=rd)
