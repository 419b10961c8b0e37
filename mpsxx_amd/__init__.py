"""mpsxx_amd: MI355X (gfx950) DMRG effective-Hamiltonian hot path.

Python access to the C ABI of ``lib/libmpsxx_amd.so`` (declared in include/mpsxx_amd.h).  The HIP library is the
only compute path: importing the package loads it and every operation raises if it is missing.
"""
import ctypes
import os

import numpy as np

from .qst import QST, QstDesc, random_fill  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmpsxx_amd.so")

FP64 = 0
FP32 = 1
BF16X3 = 2

_lib = None


class MpsxxError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpsxxError("libmpsxx_amd.so not built: run `python -m mpsxx_amd.build` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.mpsxx_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MpsxxError("mpsxx error %d: %s" % (rc, lib().mpsxx_last_error().decode()))


VP = ctypes.c_void_p


def plan_cache_stats():
    """Plan cache counters (mpsxx_plan_cache_stats): hits, builds, evictions, build_ms, plans, device_bytes,
    releases, restores."""
    i = [ctypes.c_int64() for _ in range(7)]
    ms = ctypes.c_double()
    _check(lib().mpsxx_plan_cache_stats(ctypes.byref(i[0]), ctypes.byref(i[1]), ctypes.byref(i[2]), ctypes.byref(ms),
                                        ctypes.byref(i[3]), ctypes.byref(i[4]), ctypes.byref(i[5]),
                                        ctypes.byref(i[6])))
    c = ctypes.c_int64()
    _check(lib().mpsxx_plan_cache_collisions(ctypes.byref(c)))
    return {"hits": i[0].value, "builds": i[1].value, "evictions": i[2].value, "build_ms": ms.value,
            "plans": i[3].value, "device_bytes": i[4].value, "releases": i[5].value, "restores": i[6].value,
            "collisions": c.value}


def set_structure_hash_bits(bits):
    """Keep only the low `bits` bits of structure hashes computed from now on (mpsxx_set_structure_hash_bits; 64
    restores full hashes): distinct block structures then share cache keys, which the plan caches must detect."""
    _check(lib().mpsxx_set_structure_hash_bits(ctypes.c_int32(bits)))


def release_caches():
    """Drop every cached plan and return the pool's unused device buffers (mpsxx_release_caches)."""
    _check(lib().mpsxx_release_caches())


def device_memory_stats():
    """Device allocation counters (mpsxx_device_memory_stats): mallocs, frees, malloc_ms, free_ms, reuses,
    pool_syncs, bytes_live, cached_bytes."""
    i = [ctypes.c_int64() for _ in range(6)]
    ma, fr = ctypes.c_double(), ctypes.c_double()
    _check(lib().mpsxx_device_memory_stats(ctypes.byref(i[0]), ctypes.byref(i[1]), ctypes.byref(ma), ctypes.byref(fr),
                                           ctypes.byref(i[2]), ctypes.byref(i[3]), ctypes.byref(i[4]),
                                           ctypes.byref(i[5])))
    return {"mallocs": i[0].value, "frees": i[1].value, "malloc_ms": ma.value, "free_ms": fr.value,
            "reuses": i[2].value, "pool_syncs": i[3].value, "bytes_live": i[4].value, "cached_bytes": i[5].value}


class DTensor:
    """Device-resident block-sparse array (opaque ``mpsxx_tensor`` handle)."""

    def __init__(self, qst=None):
        self.h = VP()
        if qst is None:
            _check(lib().mpsxx_tensor_create_empty(ctypes.byref(self.h)))
        else:
            d, keep = qst.to_desc()
            _check(lib().mpsxx_tensor_create(ctypes.byref(d), ctypes.byref(self.h)))
            data = np.ascontiguousarray(qst.flat(), dtype=np.float64)
            _check(lib().mpsxx_tensor_upload(self.h, data.ctypes.data_as(VP), ctypes.c_int64(data.size), None))

    def info(self):
        rank = ctypes.c_int32()
        qsym = ctypes.c_int32()
        q = (ctypes.c_int32 * 2)()
        ns = ctypes.c_int64()
        nb = ctypes.c_int64()
        ne = ctypes.c_int64()
        _check(lib().mpsxx_tensor_info(self.h, ctypes.byref(rank), ctypes.byref(qsym), q, ctypes.byref(ns),
                                       ctypes.byref(nb), ctypes.byref(ne)))
        return rank.value, qsym.value, (q[0], q[1]), ns.value, nb.value, ne.value

    def to_qst(self):
        rank, qsym, q, ns, nb, ne = self.info()
        if rank == 0:
            return None
        nsec = np.zeros(rank, dtype=np.int32)
        secq = np.zeros(2 * max(ns, 1), dtype=np.int32)
        secd = np.zeros(max(ns, 1), dtype=np.int32)
        bidx = np.zeros((max(nb, 1), rank), dtype=np.int32)
        boff = np.zeros(max(nb, 1), dtype=np.int64)
        _check(lib().mpsxx_tensor_layout(self.h, nsec.ctypes.data_as(VP), secq.ctypes.data_as(VP),
                                         secd.ctypes.data_as(VP), bidx.ctypes.data_as(VP), boff.ctypes.data_as(VP)))
        data = np.zeros(max(ne, 1))
        _check(lib().mpsxx_tensor_download(self.h, data.ctypes.data_as(VP), ctypes.c_int64(ne), None))
        return QST.from_layout(rank, qsym, q, nsec, secq, secd, bidx[:nb], data[:ne])

    def clear(self):
        _check(lib().mpsxx_tensor_clear(self.h))

    def free(self):
        if self.h:
            lib().mpsxx_tensor_free(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _h(t):
    return t.h


def _out(out):
    return DTensor() if out is None else out


def synchronize():
    _check(lib().mpsxx_synchronize(None))


# -- hot path ------------------------------------------------------------------------------------------------------
def compute_sigma_vector_twosite(lmpo, rmpo, lopr, ropr, wfn0, sgv0=None, precision=FP64):
    """sigma += H_eff psi, two-site 4-index form (reference src/compute_sigma_vector.hpp:54-70)."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_twosite(_h(lmpo), _h(rmpo), _h(lopr), _h(ropr), _h(wfn0), _h(sgv0),
                                                    ctypes.c_int(precision), None))
    return sgv0


def compute_sigma_vector_twosite_ex(lmpo, rmpo, lopr, ropr, wfn0, sgv0, precision, convert_mask):
    """Reduced-precision two-site sigma refreshing only the FP32 copies of the inputs whose bit is set."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_twosite_ex(_h(lmpo), _h(rmpo), _h(lopr), _h(ropr), _h(wfn0), _h(sgv0),
                                                       ctypes.c_int(precision), ctypes.c_uint32(convert_mask), None))
    return sgv0


def compute_sigma_vector_merged_ex(lopr, ropr, wfn0, sgv0, precision, convert_mask):
    """Reduced-precision merged sigma refreshing only the FP32 copies of the inputs whose bit is set."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_merged_ex(_h(lopr), _h(ropr), _h(wfn0), _h(sgv0), ctypes.c_int(precision),
                                                      ctypes.c_uint32(convert_mask), None))
    return sgv0


def compute_sigma_vector_onesite(mpo0, lopr, ropr, wfn0, sgv0=None, precision=FP64):
    """One-site form (src/compute_sigma_vector.hpp:10-22)."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_onesite(_h(mpo0), _h(lopr), _h(ropr), _h(wfn0), _h(sgv0),
                                                    ctypes.c_int(precision), None))
    return sgv0


def compute_sigma_vector_merged(lopr, ropr, wfn0, sgv0=None, precision=FP64):
    """Merged 2-index form (src/compute_sigma_vector.hpp:86-96)."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_merged(_h(lopr), _h(ropr), _h(wfn0), _h(sgv0), ctypes.c_int(precision),
                                                   None))
    return sgv0


def compute_sigma_vector_overlap(lsov, rsov, wfn0, sgv0=None):
    """S|psi> forms (src/compute_sigma_vector.hpp:25-35, 73-83, 99-109)."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_overlap(_h(lsov), _h(rsov), _h(wfn0), _h(sgv0), None))
    return sgv0


def compute_sigma_vector_conj(mpo0, lopr, ropr, wfn0, sgv0=None):
    """<psi|H^* form (src/compute_sigma_vector.hpp:38-51)."""
    sgv0 = _out(sgv0)
    _check(lib().mpsxx_compute_sigma_vector_conj(_h(mpo0), _h(lopr), _h(ropr), _h(wfn0), _h(sgv0), None))
    return sgv0


def compute_diagonal_elements_twosite(lmpo, rmpo, lopr, ropr, diag):
    """diag of H_eff, two-site (src/compute_diagonal_elements.hpp:35-61); ``diag`` carries q/qshape."""
    _check(lib().mpsxx_compute_diagonal_elements_twosite(_h(lmpo), _h(rmpo), _h(lopr), _h(ropr), _h(diag), None))
    return diag


def compute_diagonal_elements_onesite(mpo0, lopr, ropr, diag):
    _check(lib().mpsxx_compute_diagonal_elements_onesite(_h(mpo0), _h(lopr), _h(ropr), _h(diag), None))
    return diag


def compute_diagonal_elements_merged(lopr, ropr, diag):
    _check(lib().mpsxx_compute_diagonal_elements_merged(_h(lopr), _h(ropr), _h(diag), None))
    return diag


def renormalize(forward, mpo0, opr0, bra0, ket0, opr1=None, precision=FP64):
    """Environment update (src/renormalize.hpp:10-32)."""
    opr1 = _out(opr1)
    _check(lib().mpsxx_renormalize(ctypes.c_int(1 if forward else 0), _h(mpo0), _h(opr0), _h(bra0), _h(ket0),
                                   _h(opr1), ctypes.c_int(precision), None))
    return opr1


def renormalize_overlap(forward, sov0, bra0, ket0, sov1=None):
    sov1 = _out(sov1)
    _check(lib().mpsxx_renormalize_overlap(ctypes.c_int(1 if forward else 0), _h(sov0), _h(bra0), _h(ket0),
                                           _h(sov1), None))
    return sov1


def renormalize_merged(forward, opr0, bra0, ket0, opr1=None, precision=FP64):
    opr1 = _out(opr1)
    _check(lib().mpsxx_renormalize_merged(ctypes.c_int(1 if forward else 0), _h(opr0), _h(bra0), _h(ket0), _h(opr1),
                                          ctypes.c_int(precision), None))
    return opr1


# -- machinery -----------------------------------------------------------------------------------------------------
def contract(alpha, a, ca, b, cb, beta=1.0, c=None, conj_a=False, conj_b=False):
    """btas::Contract (include/btas/QSPARSE/QSTCONTRACT.h:16-57)."""
    c = _out(c)
    ca = np.array(ca, dtype=np.int32)
    cb = np.array(cb, dtype=np.int32)
    _check(lib().mpsxx_contract(ctypes.c_double(alpha), _h(a), ctypes.c_int(int(conj_a)), ca.ctypes.data_as(VP),
                                _h(b), ctypes.c_int(int(conj_b)), cb.ctypes.data_as(VP), ctypes.c_int32(len(ca)),
                                ctypes.c_double(beta), _h(c), None))
    return c


def tie(x, idx, y=None):
    y = _out(y)
    t = np.array(idx, dtype=np.int32)
    _check(lib().mpsxx_tie(_h(x), t.ctypes.data_as(VP), ctypes.c_int32(len(t)), _h(y), None))
    return y


def dotc(x, y):
    r = ctypes.c_double()
    _check(lib().mpsxx_dotc(_h(x), _h(y), ctypes.byref(r), None))
    return r.value


def nrm2(x):
    r = ctypes.c_double()
    _check(lib().mpsxx_nrm2(_h(x), ctypes.byref(r), None))
    return r.value


def axpy(alpha, x, y):
    _check(lib().mpsxx_axpy(ctypes.c_double(alpha), _h(x), _h(y), None))
    return y


def scal(alpha, x):
    _check(lib().mpsxx_scal(ctypes.c_double(alpha), _h(x), None))
    return x


def copy(x, y=None):
    y = _out(y)
    _check(lib().mpsxx_tensor_copy(_h(x), _h(y), None))
    return y


def compute_merged_operator(forward, mpo0, opr0, out=None):
    """Merged environment operator (src/optimize.hpp:273-368): forward merges (l, n) of Contract(L, W); backward
    merges (n, r) of Contract(W, R)."""
    out = _out(out)
    _check(lib().mpsxx_compute_merged_operator(ctypes.c_int(1 if forward else 0), _h(mpo0), _h(opr0), _h(out), None))
    return out


def davidson_twosite(groups, diag, wfnc, tol, precision=FP64, lower=()):
    """Davidson on the two-site (or, with groups of (None, None, L~, R~), the merged) effective Hamiltonian
    (src/diagonalize.hpp:310-449); wfnc is overwritten by the normalized eigenvector.  lower = [(H|k>, S|k>, E_k)]
    projects out lower roots (excited states, diagonalize.hpp:36-89).  Returns (energy, number of sigma vectors)."""
    n = len(groups)
    merged = groups[0][0] is None
    arr = lambda i: (VP * n)(*[g[i].h.value for g in groups])
    lm = None if merged else arr(0)
    rm = None if merged else arr(1)
    nl = len(lower)
    ls = (VP * max(nl, 1))(*[x[0].h.value for x in lower]) if nl else None
    lo = (VP * max(nl, 1))(*[x[1].h.value for x in lower]) if nl else None
    le = (ctypes.c_double * max(nl, 1))(*[float(x[2]) for x in lower]) if nl else None
    e = ctypes.c_double()
    ns = ctypes.c_int32()
    _check(lib().mpsxx_davidson_twosite_ex(ctypes.c_int(n), lm, rm, arr(2), arr(3), _h(diag), _h(wfnc), ctypes.c_int(nl),
                                           ls, lo, le, ctypes.c_double(tol), ctypes.c_int(precision), ctypes.byref(e),
                                           ctypes.byref(ns), None))
    return e.value, ns.value


TRUNCATION_REFERENCE = 0
TRUNCATION_STABLE = 1


def canonicalize_twosite(forward, wfnc, M, noise=0.0, rule=TRUNCATION_STABLE):
    """canonicalize (src/canonicalize.hpp:27-46): make_noise on wfnc in place, then the blocked SVD truncated to M
    states with the given rule.  Returns (lmps, rmps, discarded weight)."""
    U, V = DTensor(), DTensor()
    d = ctypes.c_double()
    _check(lib().mpsxx_canonicalize_twosite_ex(ctypes.c_int(1 if forward else 0), _h(wfnc), _h(U), _h(V),
                                               ctypes.c_int(M), ctypes.c_double(noise), ctypes.c_int(rule),
                                               ctypes.byref(d), None))
    return U, V, d.value


def make_noise(x, noise):
    """make_noise (src/canonicalize.hpp:11-24), in place."""
    _check(lib().mpsxx_make_noise(_h(x), ctypes.c_double(noise), None))
    return x


def qst_merge(x, nrow, ncol, out=None):
    """QSTmerge of the first nrow and the last ncol indices (include/btas/QSPARSE/QSTmerge.h)."""
    out = _out(out)
    _check(lib().mpsxx_qst_merge(_h(x), ctypes.c_int32(nrow), ctypes.c_int32(ncol), _h(out), None))
    return out
