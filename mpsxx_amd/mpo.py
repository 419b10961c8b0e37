"""Site basis, local fermion operators and MPO builders (hot-path inputs).

Conventions (shared with the exact-diagonalization checks in tests/):
  * site basis {0: vacuum, 1: alpha, 2: beta, 3: alpha-beta = c+_a c+_b |0>} with quantum numbers
    (particle, 2Sz) = (0,0), (1,+1), (1,-1), (2,0)  (reference src/mpsite.h:21-36, src/fermion.h);
  * many-body states are ordered with site 0 leftmost in the product of creation operators (alpha before beta on
    a site); in the tensor-product basis the fermion sign of c_{i,s} is the parity string P_0 ... P_{i-1};
  * an MPO tensor is W(wl, n_bra, n_ket, wr) with quantum shape (q_wl, q_n, -q_n, -q_wr) and total 0, one
    quantum-number sector of dimension 1 per bond operator (reference src/gen_qc_operators.cpp:82-84,
    src/prime_operators.h:177-184): the right bond carries the accumulated operator quantum number.
"""
import numpy as np

from .qst import QST

SITE_Q = [(0, 0), (1, 1), (1, -1), (2, 0)]

# local operators <bra|O|ket>
I4 = np.eye(4)
P4 = np.diag([1.0, -1.0, -1.0, 1.0])          # (-1)^n
CA_DAG = np.zeros((4, 4))                     # c+_alpha
CA_DAG[1, 0] = 1.0
CA_DAG[3, 2] = 1.0
CB_DAG = np.zeros((4, 4))                     # c+_beta (passes the alpha of the same site)
CB_DAG[2, 0] = 1.0
CB_DAG[3, 1] = -1.0
CA = CA_DAG.T.copy()
CB = CB_DAG.T.copy()
NA = np.diag([0.0, 1.0, 0.0, 1.0])
NB = np.diag([0.0, 0.0, 1.0, 1.0])


def _op_q(op):
    """Quantum number carried by a local operator (from any non-zero element)."""
    nz = np.argwhere(np.abs(op) > 0)
    if len(nz) == 0:
        return (0, 0)
    b, k = nz[0]
    return (SITE_Q[b][0] - SITE_Q[k][0], SITE_Q[b][1] - SITE_Q[k][1])


def mpo_tensor(lops, rops, entries):
    """Build W(wl, nb, nk, wr): ``lops``/``rops`` are bond-operator QN lists, ``entries`` maps (l, r) -> 4x4 matrix.

    Each non-zero matrix element becomes its own 1x1x1x1 block (reference prime_operators.h:177-184)."""
    qn = SITE_Q
    W = QST((0, 0), [list(lops), qn, [(-a, -b) for (a, b) in qn], [(-a, -b) for (a, b) in rops]],
            [[1] * len(lops), [1] * 4, [1] * 4, [1] * len(rops)])
    for (l, r), op in entries.items():
        for b in range(4):
            for k in range(4):
                v = op[b, k]
                if abs(v) < 1e-16:
                    continue
                W.insert((l, b, k, r), np.full((1, 1, 1, 1), v))
    return W


def hubbard_mpo(L, t=1.0, U=4.0):
    """Compact MPO (bond dimension 6) of the open Hubbard chain
    H = t sum_{i,s} (c+_{i,s} c_{i+1,s} + c+_{i+1,s} c_{i,s}) + U sum_i n_{i,a} n_{i,b}.

    Bond operators: 0 = identity (nothing placed), 1/2 = c+_s P placed, 3/4 = P c_s placed, 5 = complete term.
    c+_i c_{i+1} = (c+ P)_i (x) c_{i+1} and c+_{i+1} c_i = (P c)_i (x) c+_{i+1} in the tensor-product basis."""
    ops = [(0, 0), (1, 1), (1, -1), (-1, -1), (-1, 1), (0, 0)]
    onsite = U * (NA @ NB)
    mpos = []
    for i in range(L):
        lops = [(0, 0)] if i == 0 else ops
        rops = [(0, 0)] if i == L - 1 else ops
        li = {0: 0} if i == 0 else {k: k for k in range(6)}
        ri = {5: 0} if i == L - 1 else {k: k for k in range(6)}
        ent = {}

        def put(l, r, op):
            if l in li and r in ri:
                ent[(li[l], ri[r])] = ent.get((li[l], ri[r]), 0) + op

        put(0, 0, I4)
        put(5, 5, I4)
        put(0, 5, onsite)
        put(0, 1, CA_DAG @ P4)
        put(0, 2, CB_DAG @ P4)
        put(0, 3, P4 @ CA)
        put(0, 4, P4 @ CB)
        put(1, 5, t * CA)
        put(2, 5, t * CB)
        put(3, 5, t * CA_DAG)
        put(4, 5, t * CB_DAG)
        mpos.append(mpo_tensor(lops, rops, ent))
    return mpos


def mpo_dense(mpos):
    """Contract an MPO into the dense 4^L x 4^L matrix (tests only; small L)."""
    acc = None
    for W in mpos:
        Wd = W.dense()  # (wl, nb, nk, wr)
        if acc is None:
            assert Wd.shape[0] == 1
            acc = Wd[0]  # (nb, nk, wr)
        else:
            # acc (B, K, w) x W (w, nb, nk, wr) -> (B nb, K nk, wr)
            x = np.einsum("BKw,wbkr->BbKkr", acc, Wd)
            s = x.shape
            acc = x.reshape(s[0] * s[1], s[2] * s[3], s[4])
    assert acc.shape[2] == 1
    return acc[:, :, 0]
