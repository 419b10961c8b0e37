"""Exact diagonalization helpers for tests (numpy/scipy), independent of the MPO code.

Many-body basis: tensor product of site states {0, a, b, ab}, site 0 most significant; spin-orbital order for the
fermion sign is (0a, 0b, 1a, 1b, ...).  H = sum_{pq,s} h[p,q] c+_{ps} c_{qs}
          + 1/2 sum_{pqrs, s t} v[p,q,r,s] c+_{ps} c+_{rt} c_{st} c_{qs}   (chemist notation (pq|rs)) + ecore.
"""
import itertools

import numpy as np
import scipy.sparse as sp


def _occ(state, L):
    """Occupation bits (2L spin-orbitals) of a tensor-product index."""
    bits = []
    for i in range(L):
        loc = (state // 4 ** (L - 1 - i)) % 4
        bits.append(loc & 1)
        bits.append((loc >> 1) & 1)
    return bits


def _index(bits, L):
    s = 0
    for i in range(L):
        loc = bits[2 * i] + 2 * bits[2 * i + 1]
        s = s * 4 + loc
    return s


def _apply(bits, ops):
    """Apply a product of (kind, so) right-to-left; kind 1 = create, 0 = destroy. Returns (sign, bits) or None."""
    b = list(bits)
    sign = 1
    for kind, so in reversed(ops):
        if kind == 0:
            if not b[so]:
                return None
        else:
            if b[so]:
                return None
        if sum(b[:so]) % 2:
            sign = -sign
        b[so] ^= 1
    return sign, b


def hamiltonian(h, v, L, ecore=0.0, states=None):
    """Sparse H on the given tensor-product states (default: all 4^L)."""
    if states is None:
        states = list(range(4 ** L))
    pos = {s: k for k, s in enumerate(states)}
    rows, cols, vals = [], [], []
    n = len(states)
    terms = []
    for p, q in itertools.product(range(L), repeat=2):
        if abs(h[p, q]) > 0:
            for s in (0, 1):
                terms.append((h[p, q], [(1, 2 * p + s), (0, 2 * q + s)]))
    if v is not None:
        for p, q, r, s_ in itertools.product(range(L), repeat=4):
            val = v[p, q, r, s_]
            if abs(val) == 0:
                continue
            for a, b in itertools.product((0, 1), repeat=2):
                terms.append((0.5 * val, [(1, 2 * p + a), (1, 2 * r + b), (0, 2 * s_ + b), (0, 2 * q + a)]))
    for k, st in enumerate(states):
        bits = _occ(st, L)
        if ecore != 0.0:
            rows.append(k)
            cols.append(k)
            vals.append(ecore)
        for c, ops in terms:
            r = _apply(bits, ops)
            if r is None:
                continue
            sign, nb = r
            j = pos.get(_index(nb, L))
            if j is None:
                continue
            rows.append(j)
            cols.append(k)
            vals.append(c * sign)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def sector_states(L, nelec, twosz):
    out = []
    for st in range(4 ** L):
        bits = _occ(st, L)
        na = sum(bits[0::2])
        nb = sum(bits[1::2])
        if na + nb == nelec and na - nb == twosz:
            out.append(st)
    return out


def hubbard_integrals(L, t=1.0, U=4.0):
    """Nearest-neighbour open chain: h[i,i+1] = h[i+1,i] = t, (ii|ii) = U."""
    h = np.zeros((L, L))
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = t
    v = np.zeros((L, L, L, L))
    for i in range(L):
        v[i, i, i, i] = U
    return h, v


def ground_energy(h, v, L, nelec, twosz, ecore=0.0):
    st = sector_states(L, nelec, twosz)
    H = hamiltonian(h, v, L, ecore, st)
    if H.shape[0] <= 2000:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    import scipy.sparse.linalg as sla
    return float(sla.eigsh(H, k=1, which="SA", tol=1e-12)[0][0])


def ground_state(h, v, L, nelec, twosz, ecore=0.0):
    """(E0, vector, states) of the (nelec, 2Sz) sector (dense diagonalization; small sectors only)."""
    st = sector_states(L, nelec, twosz)
    H = hamiltonian(h, v, L, ecore, st).toarray()
    w, U = np.linalg.eigh(H)
    return float(w[0]), U[:, 0], st


def expect(h, v, L, vec, states):
    """<vec| H(h, v) |vec> on the given sector states."""
    H = hamiltonian(h, v, L, 0.0, states)
    return float(vec @ (H @ vec))
