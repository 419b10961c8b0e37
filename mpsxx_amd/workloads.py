"""Synthetic hot-path workloads (inputs already shaped like a converged DMRG state).

``hubbard_twosite(L, M, U)``: the two-site sigma-vector problem at the middle bond of the open Hubbard chain at half
filling and 2Sz = 0, with bond dimension M distributed over the quantum-number sectors of the bond like the reduced
density matrix spectrum of a converged chain (Gaussian in particle-number and spin deviation from the mean, width
sigma = 1.5; no sector holds more states than the sites on its side can carry), and pattern-preserving random entries
N(0,1)/sqrt(m) in every quantum-number-allowed block.

``qc_merged_problem(norb, M, seed)``: the merged-form sigma problem at the middle of a quantum-chemistry chain
(synthetic integrals, QC MPO with swap-sweep pair operators): left/right environments with the bond operators of the
two middle MPO sites, sector dimensions as above (centre N*k/norb, capacities of k sites), random entries; merged
operators and wavefunction built on the device.

``hubbard_isometric_twosite(L, M, U)``: the same middle two-site problem with the environments of a random MPS whose
tensors left of the centre are left isometries and right of it right isometries (bond sectors as above, capped so that
every tensor can be an isometry), renormalized on the device: H_eff = P^T H P for an isometry P, a Hermitian problem
whose lowest eigenvalue is a variational energy of the chain, as every Davidson solve of a sweep sees it;
``hubbard_isometric_problem`` adds the diagonal and a start vector near the lowest eigenvector.

``batch_specs(n, seed)``: the (L, U, M) sample of a batch of independent chains: L uniform over the even lengths
16..64 (half filling with 2Sz = 0 needs N = L even), U from {2, 4, 8}, M from {128, 256, 512}.
"""
import math

import numpy as np

from . import mps
from .mpo import hubbard_mpo
from .qst import QST


def sector_capacity(nsites, q):
    """Number of states of `nsites` Hubbard sites with particle number q[0] and 2Sz q[1]."""
    p, s = q
    if (p + s) % 2 or p < 0:
        return 0
    nu, nd = (p + s) // 2, (p - s) // 2
    if nu < 0 or nd < 0 or nu > nsites or nd > nsites:
        return 0
    return math.comb(nsites, nu) * math.comb(nsites, nd)


def _cap(dims, caps, w):
    """Clip dims to the sector capacities and hand the excess to the heaviest sectors with room left."""
    dims = list(dims)
    excess = 0
    for j, c in enumerate(caps):
        if dims[j] > c:
            excess += dims[j] - c
            dims[j] = c
    for j in np.argsort(-np.asarray(w), kind="stable"):
        if excess == 0:
            break
        room = caps[j] - dims[j]
        if room > 0 and w[j] > 0:
            take = min(room, excess)
            dims[j] += take
            excess -= take
    return dims


def gaussian_dims(qns, center, M, width=1.5, min_weight=0.02, caps=None):
    """Distribute M states over the sectors `qns` with weight exp(-(dp^2 + ds^2) / (2 width^2)), at most caps[j]
    states in sector j when capacities are given."""
    w = []
    for (p, s) in qns:
        dp = p - center[0]
        ds = s - center[1]
        w.append(math.exp(-(dp * dp + ds * ds) / (2.0 * width * width)))
    w = np.array(w)
    keep = w >= min_weight * w.max()
    w = np.where(keep, w, 0.0)
    raw = w / w.sum() * M
    dims = np.floor(raw).astype(int)
    dims = np.where(keep & (dims == 0), 1, dims)
    # hand the remainder to the largest fractional parts
    rem = M - int(dims.sum())
    order = np.argsort(-(raw - np.floor(raw)))
    k = 0
    while rem > 0:
        j = order[k % len(order)]
        if keep[j]:
            dims[j] += 1
            rem -= 1
        k += 1
    while rem < 0:
        j = int(np.argmax(dims))
        dims[j] -= 1
        rem += 1
    if caps is not None:
        dims = _cap(dims, caps, w)
    return [int(d) for d in dims]


def _prune(qs, ds):
    q2 = [q for q, d in zip(qs, ds) if d > 0]
    d2 = [d for d in ds if d > 0]
    return q2, d2


def hubbard_twosite(L=32, M=512, U=4.0, t=1.0, seed=0, width=1.5):
    """(lmpo, rmpo, lenv, renv, psi) QSTs of the middle two-site problem; see the module docstring."""
    rng = np.random.default_rng(seed)
    qt = (L, 0)
    i = L // 2 - 1  # two-site centre (i, i+1)
    mpos = hubbard_mpo(L, t, U)
    lq = mps.left_bond_qns(L, i, qt)
    rq = mps.right_bond_qns(L, i + 2, qt)
    nl = i             # sites on the left
    nr = L - (i + 2)   # sites on the right
    ld = gaussian_dims(lq, (nl * qt[0] / L, 0), M, width, caps=[sector_capacity(nl, q) for q in lq])
    rd = gaussian_dims(rq, (-nr * qt[0] / L, 0), M, width, caps=[sector_capacity(nr, (-q[0], -q[1])) for q in rq])
    lq, ld = _prune(lq, ld)
    rq, rd = _prune(rq, rd)

    def fill(a):
        import itertools
        for idx in itertools.product(*[range(a.nsec(k)) for k in range(a.rank)]):
            if a.allowed(idx):
                shp = a.block_shape(idx)
                m = max(1, max(shp))
                a.blocks[tuple(idx)] = rng.standard_normal(shp) / math.sqrt(m)
        return a

    qw = mpos[i].qshape[0]
    lenv = fill(QST((0, 0), [lq, [(-a, -b) for a, b in qw], [(-a, -b) for a, b in lq]], [ld, [1] * len(qw), ld]))
    qwr = [(-a, -b) for a, b in mpos[i + 1].qshape[3]]
    renv = fill(QST((0, 0), [[(-a, -b) for a, b in rq], qwr, rq], [rd, [1] * len(qwr), rd]))
    psi = fill(mps.twosite_wavefunction(lq, ld, rq, rd, qt))
    return mpos[i], mpos[i + 1], lenv, renv, psi


def _isometry_columns(nr, d, rng, score):
    """An nr x d matrix with orthonormal columns: the QR factor of a Gaussian matrix (score None), or d distinct unit
    columns at the rows of lowest score (ties in random order), i.e. a selection of basis states.  Returns (Q, the
    scores of the d columns)."""
    if score is None:
        return np.linalg.qr(rng.standard_normal((nr, d)))[0], None
    order = sorted(range(nr), key=lambda r: (score[r], rng.random()))[:d]
    q = np.zeros((nr, d))
    q[order, np.arange(d)] = 1.0
    return q, [score[r] for r in order]


def _left_isometry(lq, ld, rq, rd, rng, scores=None):
    """Left-isometric A(l, n, r) (quantum shape (lq, site, -rq)): for every right sector the stacked (l, n) rows of
    its blocks have orthonormal columns.  With scores (per left sector, per state: the doubly occupied sites of the
    configuration it stands for), the columns select the configurations with the fewest doubly occupied sites; the
    new states' scores are returned with the tensor."""
    from .mpo import SITE_Q
    a = mps.site_tensor(lq, ld, rq, rd)
    out = []
    for b, qr in enumerate(rq):
        rows = [(x, n) for x, ql in enumerate(lq) for n, qn in enumerate(SITE_Q)
                if (ql[0] + qn[0], ql[1] + qn[1]) == qr]
        nr = sum(ld[x] for x, _ in rows)
        sc = None if scores is None else [scores[x][j] + (n == 3) for x, n in rows for j in range(ld[x])]
        q, s = _isometry_columns(nr, rd[b], rng, sc)
        out.append(s)
        r0 = 0
        for x, n in rows:
            a.blocks[(x, n, b)] = q[r0:r0 + ld[x]].reshape(ld[x], 1, rd[b])
            r0 += ld[x]
    return a, out


def _right_isometry(lq, ld, rq, rd, rng, scores=None):
    """Right-isometric B(l, n, r) (quantum shape (lq, site, -rq)): for every left sector the (n, r) columns of its
    blocks have orthonormal rows; scores per right sector as in _left_isometry."""
    from .mpo import SITE_Q
    bt = mps.site_tensor(lq, ld, rq, rd)
    out = []
    for x, ql in enumerate(lq):
        cols = [(n, b) for n, qn in enumerate(SITE_Q) for b, qr in enumerate(rq)
                if (ql[0] + qn[0], ql[1] + qn[1]) == qr]
        nc = sum(rd[b] for _, b in cols)
        sc = None if scores is None else [scores[b][j] + (n == 3) for n, b in cols for j in range(rd[b])]
        q, s = _isometry_columns(nc, ld[x], rng, sc)
        out.append(s)
        c0 = 0
        for n, b in cols:
            bt.blocks[(x, n, b)] = q[c0:c0 + rd[b]].T.reshape(ld[x], 1, rd[b])
            c0 += rd[b]
    return bt, out


def isometric_chain(L=32, M=512, seed=0, width=1.5, basis="configurations"):
    """Host tensors of an MPS around the middle two-site centre (i, i+1), i = L/2 - 1, at half filling and 2Sz = 0:
    left isometries A_0..A_{i-1}, right isometries B_{i+2}..B_{L-1} (dict by site) and a random two-site psi.  Bond
    sectors as hubbard_twosite's, each capped by the states the bond's neighbour can supply, so that every tensor can
    be an isometry.  basis "configurations": every bond state is one occupation configuration of its side, those with
    the fewest doubly occupied sites kept (P^T H P is then H on a selected set of configurations, whose diagonal is a
    working Davidson preconditioner); "random": Gaussian isometries.  Returns (A, B, psi)."""
    if basis not in ("configurations", "random"):
        raise ValueError("basis: 'configurations' or 'random'")
    from .mpo import SITE_Q
    rng = np.random.default_rng(seed)
    qt = (L, 0)
    i = L // 2 - 1
    lq, ld = [[(0, 0)]], [[1]]
    for k in range(1, i + 1):
        q = mps.left_bond_qns(L, k, qt)
        want = gaussian_dims(q, (k * qt[0] / L, 0), M, width, caps=[sector_capacity(k, x) for x in q])
        avail = {}
        for ql, d in zip(lq[-1], ld[-1]):
            for qn in SITE_Q:
                key = (ql[0] + qn[0], ql[1] + qn[1])
                avail[key] = avail.get(key, 0) + d
        q2, d2 = _prune(q, [min(w, avail.get(x, 0)) for x, w in zip(q, want)])
        lq.append(q2)
        ld.append(d2)
    rq, rd = {L: [(0, 0)]}, {L: [1]}
    for k in range(L - 1, i + 1, -1):
        q = mps.right_bond_qns(L, k, qt)
        want = gaussian_dims(q, (-(L - k) * qt[0] / L, 0), M, width,
                             caps=[sector_capacity(L - k, (-x[0], -x[1])) for x in q])
        avail = {}
        for qr, d in zip(rq[k + 1], rd[k + 1]):
            for qn in SITE_Q:
                key = (qr[0] - qn[0], qr[1] - qn[1])
                avail[key] = avail.get(key, 0) + d
        rq[k], rd[k] = _prune(q, [min(w, avail.get(x, 0)) for x, w in zip(q, want)])
    sel = basis == "configurations"
    A, sc = [], [[0]] if sel else None
    for k in range(i):
        a, sc = _left_isometry(lq[k], ld[k], lq[k + 1], ld[k + 1], rng, sc)
        A.append(a)
    B, sc = {}, [[0]] if sel else None
    for k in range(L - 1, i + 1, -1):
        B[k], sc = _right_isometry(rq[k], rd[k], rq[k + 1], rd[k + 1], rng, sc)
    psi = mps.fill_allowed(mps.twosite_wavefunction(lq[i], ld[i], rq[i + 2], rd[i + 2], qt), rng)
    return A, B, psi


def hubbard_isometric_twosite(L=32, M=512, U=4.0, t=1.0, seed=0, width=1.5, basis="configurations"):
    """Device tensors (lmpo, rmpo, lenv, renv, psi) of the middle two-site problem of the open Hubbard chain with the
    environments of isometric_chain renormalized on the device (src/renormalize.hpp:10-32); see the module
    docstring."""
    from . import DTensor, renormalize
    i = L // 2 - 1
    mpos = hubbard_mpo(L, t, U)
    A, B, psi = isometric_chain(L, M, seed, width, basis)
    lenv = DTensor(mps.boundary_env())
    for k in range(i):
        a = DTensor(A[k])
        lenv = renormalize(True, DTensor(mpos[k]), lenv, a, a)
    renv = DTensor(mps.boundary_env())
    for k in range(L - 1, i + 1, -1):
        b = DTensor(B[k])
        renv = renormalize(False, DTensor(mpos[k]), renv, b, b)
    return DTensor(mpos[i]), DTensor(mpos[i + 1]), lenv, renv, DTensor(psi)


def hubbard_isometric_problem(L=32, M=512, U=4.0, t=1.0, seed=0, width=1.5):
    """hubbard_isometric_twosite plus what a Davidson solve takes: (lmpo, rmpo, lenv, renv, diag, psi) with the
    diagonal of H_eff (src/compute_diagonal_elements.hpp:35-61) and a start vector in the basin of the lowest
    eigenvector, psi = x / (diag - min(diag) + 1/2) for the random x, the role the previous step's wavefunction plays
    in a sweep (the reference's Davidson, restarting from a single Ritz vector with a diagonal preconditioner, can
    stall on a random start in these random subspaces)."""
    from . import DTensor, compute_diagonal_elements_twosite
    lm, rm, le, re_, psi = hubbard_isometric_twosite(L, M, U, t, seed, width)
    x = psi.to_qst()
    diag = compute_diagonal_elements_twosite(lm, rm, le, re_, DTensor(x.empty_like()))
    d = diag.to_qst()
    dmin = min([0.0] + [float(np.min(b)) for b in d.blocks.values() if b.size])
    for k, b in x.blocks.items():
        dk = d.blocks.get(k)  # a block absent from the diagonal has a zero diagonal
        x.blocks[k] = b / ((dk if dk is not None else 0.0) - dmin + 0.5)
    return lm, rm, le, re_, diag, DTensor(x)


def batch_specs(n=64, seed=4):
    """(L, U, M) of n independent half-filled chains; see the module docstring."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = 2 * int(rng.integers(8, 33))
        U = float(rng.choice([2.0, 4.0, 8.0]))
        M = int(rng.choice([128, 256, 512]))
        out.append((L, U, M))
    return out


def _env_blocks(rows_q, rows_d, wq, cols_q, cols_d, rng, left):
    """Random environment QST over (rows, w, cols) with every quantum-number-allowed block filled."""
    if left:   # L(l_bra, w, l_ket): quantum shape (lq, -q_w, -lq), total 0
        qs = [list(rows_q), [(-a, -b) for (a, b) in wq], [(-a, -b) for (a, b) in cols_q]]
    else:      # R(r_bra, w, r_ket): quantum shape (-rq, q_w, rq)
        qs = [[(-a, -b) for (a, b) in rows_q], list(wq), list(cols_q)]
    e = QST((0, 0), qs, [list(rows_d), [1] * len(wq), list(cols_d)])
    index = {}
    for a, q in enumerate(qs[0]):
        index.setdefault(q, []).append(a)
    for w, qw in enumerate(qs[1]):
        for c, qc_ in enumerate(qs[2]):
            need = (-(qw[0] + qc_[0]), -(qw[1] + qc_[1]))
            for a in index.get(need, []):
                shp = (rows_d[a], 1, cols_d[c])
                e.blocks[(a, w, c)] = rng.standard_normal(shp) / math.sqrt(max(shp))
    return e


def qc_merged_problem(norb=24, M=1024, seed=1, width=1.5):
    """Device tensors (Lm, Rm, psim) of the merged sigma at the middle of a norb-orbital chain; see the docstring."""
    from . import DTensor, compute_merged_operator, qst_merge
    from .qc import QcMpo, synthetic_integrals
    rng = np.random.default_rng(seed)
    h, v = synthetic_integrals(norb, seed)
    m = QcMpo(h, v, swap_sweep=True)
    K = norb
    qt = (K, 0)
    i = K // 2 - 1
    lq = mps.left_bond_qns(K, i, qt)
    rq = mps.right_bond_qns(K, i + 2, qt)
    nl, nr = i, K - (i + 2)
    ld = gaussian_dims(lq, (nl * qt[0] / K, 0), M, width, caps=[sector_capacity(nl, q) for q in lq])
    rd = gaussian_dims(rq, (-nr * qt[0] / K, 0), M, width, caps=[sector_capacity(nr, (-q[0], -q[1])) for q in rq])
    lq, ld = _prune(lq, ld)
    rq, rd = _prune(rq, rd)
    wl = m.site_entries(i)[0]        # left bond operators of site i
    wr = m.site_entries(i + 1)[1]    # right bond operators of site i+1
    lenv = _env_blocks(lq, ld, wl, lq, ld, rng, True)
    renv = _env_blocks(rq, rd, wr, rq, rd, rng, False)
    psi = mps.twosite_wavefunction(lq, ld, rq, rd, qt)
    psi = mps.fill_allowed(psi, rng, 1.0)
    W1, W2 = m.site_device(i), m.site_device(i + 1)
    dl, dr, dpsi = DTensor(lenv), DTensor(renv), DTensor(psi)
    Lm = compute_merged_operator(True, W1, dl)
    Rm = compute_merged_operator(False, W2, dr)
    psim = qst_merge(dpsi, 2, 2)
    info = dict(norb=norb, M=M, seed=seed, site=i, D_w=m.bond_dims()[i + 1], D_w_left=len(wl), D_w_right=len(wr),
                left_sectors=len(lq), right_sectors=len(rq), psi_blocks=len(psi.blocks), psi_elems=psi.nelem(),
                env_elems=lenv.nelem() + renv.nelem())
    return Lm, Rm, psim, info


def merged_problem_from_state(mpos, sites, i):
    """Device tensors (Lm, Rm, psim, info) of the merged two-site sigma at sites (i, i+1) of an MPS held on the device
    (`sites` = DTensors A(l, n, r), e.g. the site tensors of a converged sweep): the left environment renormalized
    over sites 0..i-1 and the right one over sites L-1..i+2 (src/renormalize.hpp:10-32), the merged operators of
    sites i and i+1 (src/optimize.hpp:273-368) and the two-site wavefunction A_i A_{i+1} merged to (l n, n r)."""
    from . import DTensor, compute_merged_operator, contract, qst_merge, renormalize
    L = len(sites)
    lenv = DTensor(mps.boundary_env())
    for k in range(i):
        lenv = renormalize(True, mpos[k], lenv, sites[k], sites[k])
    renv = DTensor(mps.boundary_env())
    for k in range(L - 1, i + 1, -1):
        renv = renormalize(False, mpos[k], renv, sites[k], sites[k])
    psi = contract(1.0, sites[i], [2], sites[i + 1], [0])
    Lm = compute_merged_operator(True, mpos[i], lenv)
    Rm = compute_merged_operator(False, mpos[i + 1], renv)
    psim = qst_merge(psi, 2, 2)
    ps = psi.to_qst()
    info = dict(site=i, psi_blocks=len(ps.blocks), psi_elems=ps.nelem(),
                left_sectors=len(ps.qshape[0]), right_sectors=len(ps.qshape[3]),
                bond_left=int(sum(ps.dshape[0])), bond_right=int(sum(ps.dshape[3])))
    return Lm, Rm, psim, info
