"""Quantum-chemistry inputs: FCIDUMP integrals, synthetic integrals and the QC MPO (C ABI ``mpsxx_qc_mpo_*``).

Integral convention (reference src/parsing_integral.cpp:105-150): ``h[p, q]`` one-electron, ``v[p, q, r, s] = (pq|rs)``
in chemist notation with 8-fold symmetry; H = sum h_pq a+_ps a_qs + 1/2 sum (pq|rs) a+_ps a+_rt a_st a_qs + ecore.
"""
import ctypes

import numpy as np

from . import VP, DTensor, _check, lib
from .mpo import mpo_tensor


def read_fcidump(path):
    """(norb, nelec, ms2, h, v, ecore) from a Molpro FCIDUMP (1-based indices, 8-fold symmetric expansion)."""
    with open(path) as f:
        text = f.read()
    head, _, body = text.partition("&END") if "&END" in text else text.partition("/")
    if not body:
        raise ValueError("FCIDUMP without &END or / terminator: " + path)
    head = head.replace(",", " ").replace("=", " = ")
    tok = head.split()

    def field(name, default=0):
        for i, t in enumerate(tok):
            if t.upper() == name and i + 2 < len(tok) and tok[i + 1] == "=":
                return int(tok[i + 2])
        return default

    norb, nelec, ms2 = field("NORB"), field("NELEC"), field("MS2")
    h = np.zeros((norb, norb))
    v = np.zeros((norb, norb, norb, norb))
    ecore = 0.0
    for line in body.splitlines():
        parts = line.split()
        if len(parts) < 5:
            continue
        val = float(parts[0].replace("D", "E").replace("d", "e"))
        i, j, k, l = (int(x) - 1 for x in parts[1:5])
        if i < 0 and j < 0:
            ecore = val
        elif k < 0 and l < 0:
            h[i, j] = h[j, i] = val
        else:
            for (a, b, c, d) in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                                 (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                v[a, b, c, d] = val
    return norb, nelec, ms2, h, v, ecore


def write_fcidump(path, h, v, ecore=0.0, nelec=None, ms2=0):
    """Molpro FCIDUMP with the unique (ij|kl), i >= j, k >= l, ij >= kl, then h_ij (i >= j), then ecore."""
    norb = h.shape[0]
    nelec = norb if nelec is None else nelec
    with open(path, "w") as f:
        f.write(" &FCI NORB=%3d,NELEC=%3d,MS2=%2d,\n  ORBSYM=%s\n  ISYM=1\n &END\n" % (norb, nelec, ms2, "1," * norb))
        for i in range(norb):
            for j in range(i + 1):
                for k in range(i + 1):
                    for l in range(k + 1):
                        if (i * (i + 1) // 2 + j) < (k * (k + 1) // 2 + l):
                            continue
                        x = v[i, j, k, l]
                        if abs(x) > 1e-16:
                            f.write("%28.20e%4d%4d%4d%4d\n" % (x, i + 1, j + 1, k + 1, l + 1))
        for i in range(norb):
            for j in range(i + 1):
                if abs(h[i, j]) > 1e-16:
                    f.write("%28.20e%4d%4d   0   0\n" % (h[i, j], i + 1, j + 1))
        f.write("%28.20e   0   0   0   0\n" % ecore)


def synthetic_integrals(norb, seed, nvec=None, hdiag=(-2.0, -0.5), hsig=0.05, lsig=0.1):
    """Synthetic molecular-like integrals: h symmetric, diagonal linear from hdiag[0] (p = 0) to hdiag[1]
    (p = norb-1), off-diagonal N(0, hsig) symmetrised; v_pqrs = sum_k L^k_pq L^k_rs with nvec (default 2 norb)
    symmetric L^k of N(0, lsig) entries (8-fold symmetric, positive semi-definite).

    Fill order (numpy default_rng(seed)): the upper triangle of h row by row (p < q), then for k = 0..nvec-1 the upper
    triangle including the diagonal of L^k row by row."""
    rng = np.random.default_rng(seed)
    nvec = 2 * norb if nvec is None else nvec
    h = np.diag(np.linspace(hdiag[0], hdiag[1], norb))
    iu = np.triu_indices(norb, 1)
    off = rng.normal(0.0, hsig, size=len(iu[0]))
    h[iu] = off
    h[(iu[1], iu[0])] = off
    iu0 = np.triu_indices(norb, 0)
    Ls = np.zeros((nvec, norb, norb))
    for k in range(nvec):
        x = rng.normal(0.0, lsig, size=len(iu0[0]))
        Ls[k][iu0] = x
        Ls[k][(iu0[1], iu0[0])] = x
    v = np.einsum("kpq,krs->pqrs", Ls, Ls)
    return h, v


class QcMpo:
    """QC MPO built by the library (host side); site tensors as QSTs (host) or DTensors (device)."""

    def __init__(self, h, v, ecore=0.0, swap_sweep=True):
        self.norb = h.shape[0]
        hh = np.ascontiguousarray(h, dtype=np.float64)
        vv = np.ascontiguousarray(v, dtype=np.float64)
        self.h = VP()
        _check(lib().mpsxx_qc_mpo_create(ctypes.c_int32(self.norb), hh.ctypes.data_as(VP), vv.ctypes.data_as(VP),
                                         ctypes.c_double(ecore), ctypes.c_int(int(swap_sweep)), ctypes.byref(self.h)))

    def bond_dims(self):
        d = np.zeros(self.norb + 1, dtype=np.int32)
        n = ctypes.c_int64()
        _check(lib().mpsxx_qc_mpo_info(self.h, d.ctypes.data_as(VP), ctypes.byref(n)))
        return [int(x) for x in d]

    def site_entries(self, k):
        nl, nr, ne = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _check(lib().mpsxx_qc_mpo_site_entries(self.h, ctypes.c_int32(k), ctypes.byref(nl), ctypes.byref(nr),
                                               ctypes.byref(ne), None, None, None, None, None))
        lq = np.zeros(2 * nl.value, dtype=np.int32)
        rq = np.zeros(2 * nr.value, dtype=np.int32)
        wl = np.zeros(max(ne.value, 1), dtype=np.int32)
        wr = np.zeros(max(ne.value, 1), dtype=np.int32)
        mats = np.zeros((max(ne.value, 1), 4, 4))
        _check(lib().mpsxx_qc_mpo_site_entries(self.h, ctypes.c_int32(k), None, None, None, lq.ctypes.data_as(VP),
                                               rq.ctypes.data_as(VP), wl.ctypes.data_as(VP), wr.ctypes.data_as(VP),
                                               mats.ctypes.data_as(VP)))
        lqs = [(int(lq[2 * i]), int(lq[2 * i + 1])) for i in range(nl.value)]
        rqs = [(int(rq[2 * i]), int(rq[2 * i + 1])) for i in range(nr.value)]
        return lqs, rqs, wl[:ne.value], wr[:ne.value], mats[:ne.value]

    def site_qst(self, k):
        lqs, rqs, wl, wr, mats = self.site_entries(k)
        return mpo_tensor(lqs, rqs, {(int(a), int(b)): m for a, b, m in zip(wl, wr, mats)})

    def qsts(self):
        return [self.site_qst(k) for k in range(self.norb)]

    def site_device(self, k):
        t = DTensor()
        _check(lib().mpsxx_qc_mpo_site(self.h, ctypes.c_int32(k), t.h, None))
        return t

    def devices(self):
        return [self.site_device(k) for k in range(self.norb)]

    def gradient(self, grads):
        """(dh, dlambda, energy) from the MPO gradients of a state: dh[p, q] = sum_s <a+_ps a_qs>, dlambda = <V>,
        energy = <H> (see mpsxx_qc_mpo_gradient)."""
        K = self.norb
        arr = (VP * K)(*[g.h.value for g in grads])
        dh = np.zeros((K, K))
        dl = ctypes.c_double()
        e = ctypes.c_double()
        _check(lib().mpsxx_qc_mpo_gradient(self.h, ctypes.c_int32(K), arr, dh.ctypes.data_as(VP), ctypes.byref(dl),
                                           ctypes.byref(e), None))
        return dh, dl.value, e.value

    def free(self):
        if self.h:
            lib().mpsxx_qc_mpo_free(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
