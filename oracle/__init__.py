"""ORACLE -- test infrastructure only.

ctypes access to ``liboracle_mpsxx.so``: the FP64 CPU restatement of the reference's block-sparse algorithm
(oracle.cpp, every function cited there against naokin/mpsxx).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this package, as the checker or the timed CPU baseline.

Pinning (see tests/test_oracle_*.py): dense numpy brute force of every contraction on random block-sparse inputs,
exact diagonalization of the many-body Hamiltonian (H_eff = P^T H P, energies), closed forms (two-site Hubbard,
free fermions), and the reference's own logged energies for the Hubbard L=12 chain.
"""
import ctypes
import os

import numpy as np

from mpsxx_amd.qst import QST

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_mpsxx.so")
_lib = None
VP = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build(verbose=False)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_create.restype = VP
        L.orc_empty.restype = VP
        L.orc_dotc.restype = ctypes.c_double
        _lib = L
    return _lib


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(int(n)))


_blas = None


def openblas():
    """(library, version string) of numpy's bundled ILP64 OpenBLAS (scipy-openblas), loaded once."""
    global _blas
    if _blas is None:
        import glob
        libs = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
        cand = sorted(glob.glob(os.path.join(libs, "libscipy_openblas64_*.so")))
        if not cand:
            raise RuntimeError("oracle: numpy's bundled OpenBLAS not found under " + libs)
        b = ctypes.CDLL(cand[0])
        b.scipy_openblas_get_config64_.restype = ctypes.c_char_p
        _blas = (b, b.scipy_openblas_get_config64_().decode())
    return _blas


def use_blas(enable=True, threads=1):
    """Route the oracle's block dgemm (C += A B per block pair) through numpy's bundled OpenBLAS cblas_dgemm
    (64-bit integers) with `threads` BLAS threads, as the reference calls CBLAS dgemm per block pair; False restores
    the oracle's own loops.  For CPU timings: the values agree with the loops to rounding (tests/test_oracle_blas.py)."""
    if not enable:
        lib().orc_set_dgemm(None)
        return None
    b, cfg = openblas()
    b.scipy_openblas_set_num_threads64_(ctypes.c_int(int(threads)))
    lib().orc_set_dgemm(ctypes.cast(b.scipy_cblas_dgemm64_, VP))
    return cfg


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


class OTensor:
    def __init__(self, qst=None):
        if qst is None:
            self.h = VP(lib().orc_empty())
        else:
            d, keep = qst.to_desc()
            data = np.ascontiguousarray(qst.flat(), dtype=np.float64)
            self.h = VP(lib().orc_create(ctypes.byref(d), data.ctypes.data_as(VP) if data.size else None))

    def to_qst(self):
        rank = ctypes.c_int32()
        qsym = ctypes.c_int32()
        q = (ctypes.c_int32 * 2)()
        ns = ctypes.c_int64()
        nb = ctypes.c_int64()
        ne = ctypes.c_int64()
        lib().orc_info(self.h, ctypes.byref(rank), ctypes.byref(qsym), q, ctypes.byref(ns), ctypes.byref(nb),
                       ctypes.byref(ne))
        if rank.value == 0:
            return None
        r = rank.value
        nsec = np.zeros(r, dtype=np.int32)
        secq = np.zeros(2 * max(ns.value, 1), dtype=np.int32)
        secd = np.zeros(max(ns.value, 1), dtype=np.int32)
        bidx = np.zeros((max(nb.value, 1), r), dtype=np.int32)
        data = np.zeros(max(ne.value, 1))
        lib().orc_layout(self.h, nsec.ctypes.data_as(VP), secq.ctypes.data_as(VP), secd.ctypes.data_as(VP),
                         bidx.ctypes.data_as(VP), data.ctypes.data_as(VP))
        return QST.from_layout(r, qsym.value, (q[0], q[1]), nsec, secq, secd, bidx[:nb.value], data[:ne.value])

    def __del__(self):
        try:
            if self.h:
                lib().orc_free(self.h)
        except Exception:
            pass


def _o(x):
    return x if isinstance(x, OTensor) else OTensor(x)


def _run(fn, *args):
    _check(fn(*args))


def sigma_twosite(lmpo, rmpo, lopr, ropr, wfn, sgv=None):
    out = OTensor(sgv) if sgv is not None else OTensor()
    ins = [_o(x) for x in (lmpo, rmpo, lopr, ropr, wfn)]
    _run(lib().orc_sigma_twosite, *[x.h for x in ins], out.h)
    return out.to_qst()


def sigma_onesite(mpo, lopr, ropr, wfn, sgv=None):
    out = OTensor(sgv) if sgv is not None else OTensor()
    ins = [_o(x) for x in (mpo, lopr, ropr, wfn)]
    _run(lib().orc_sigma_onesite, *[x.h for x in ins], out.h)
    return out.to_qst()


def sigma_merged(lopr, ropr, wfn, sgv=None):
    out = OTensor(sgv) if sgv is not None else OTensor()
    ins = [_o(x) for x in (lopr, ropr, wfn)]
    _run(lib().orc_sigma_merged, *[x.h for x in ins], out.h)
    return out.to_qst()


def sigma_overlap(lsov, rsov, wfn, sgv=None):
    out = OTensor(sgv) if sgv is not None else OTensor()
    ins = [_o(x) for x in (lsov, rsov, wfn)]
    _run(lib().orc_sigma_overlap, *[x.h for x in ins], out.h)
    return out.to_qst()


def diag_twosite(lmpo, rmpo, lopr, ropr, diag):
    out = OTensor(diag)
    ins = [_o(x) for x in (lmpo, rmpo, lopr, ropr)]
    _run(lib().orc_diag_twosite, *[x.h for x in ins], out.h)
    return out.to_qst()


def diag_onesite(mpo, lopr, ropr, diag):
    out = OTensor(diag)
    ins = [_o(x) for x in (mpo, lopr, ropr)]
    _run(lib().orc_diag_onesite, *[x.h for x in ins], out.h)
    return out.to_qst()


def diag_merged(lopr, ropr, diag):
    out = OTensor(diag)
    ins = [_o(x) for x in (lopr, ropr)]
    _run(lib().orc_diag_merged, *[x.h for x in ins], out.h)
    return out.to_qst()


def renormalize(forward, mpo, opr0, bra, ket, opr1=None):
    out = OTensor(opr1) if opr1 is not None else OTensor()
    ins = [_o(x) for x in (mpo, opr0, bra, ket)]
    _run(lib().orc_renormalize, ctypes.c_int(1 if forward else 0), *[x.h for x in ins], out.h)
    return out.to_qst()


def renormalize_overlap(forward, sov0, bra, ket, sov1=None):
    out = OTensor(sov1) if sov1 is not None else OTensor()
    ins = [_o(x) for x in (sov0, bra, ket)]
    _run(lib().orc_renormalize_overlap, ctypes.c_int(1 if forward else 0), *[x.h for x in ins], out.h)
    return out.to_qst()


def renormalize_merged(forward, opr0, bra, ket, opr1=None):
    out = OTensor(opr1) if opr1 is not None else OTensor()
    ins = [_o(x) for x in (opr0, bra, ket)]
    _run(lib().orc_renormalize_merged, ctypes.c_int(1 if forward else 0), *[x.h for x in ins], out.h)
    return out.to_qst()


def contract(alpha, a, ca, b, cb, beta=1.0, c=None, conj_a=False, conj_b=False):
    out = OTensor(c) if c is not None else OTensor()
    A, B = _o(a), _o(b)
    ca = np.array(ca, dtype=np.int32)
    cb = np.array(cb, dtype=np.int32)
    _run(lib().orc_contract, ctypes.c_double(alpha), A.h, ctypes.c_int(int(conj_a)), ca.ctypes.data_as(VP), B.h,
         ctypes.c_int(int(conj_b)), cb.ctypes.data_as(VP), ctypes.c_int32(len(ca)), ctypes.c_double(beta), out.h)
    return out.to_qst()


def tie(x, idx):
    out = OTensor()
    X = _o(x)
    t = np.array(idx, dtype=np.int32)
    _run(lib().orc_tie, X.h, t.ctypes.data_as(VP), ctypes.c_int32(len(t)), out.h)
    return out.to_qst()


def davidson_twosite(groups, diag, wfn, tol):
    """groups: list of (lmpo, rmpo, lopr, ropr) QSTs.  Returns (energy, wfn, n_sigma)."""
    n = len(groups)
    objs = [[_o(x) for x in g] for g in groups]
    arrs = [(VP * n)(*[objs[g][k].h for g in range(n)]) for k in range(4)]
    D = _o(diag)
    W = OTensor(wfn)
    e = ctypes.c_double()
    ns = ctypes.c_int32()
    _run(lib().orc_davidson_twosite, ctypes.c_int(n), arrs[0], arrs[1], arrs[2], arrs[3], D.h, W.h,
         ctypes.c_double(tol), ctypes.byref(e), ctypes.byref(ns))
    return e.value, W.to_qst(), ns.value


def jacobi_eig(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    _run(lib().orc_jacobi_eig, ctypes.c_int(n), a.ctypes.data_as(VP), w.ctypes.data_as(VP), v.ctypes.data_as(VP))
    return w, v


def canonicalize_twosite(forward, wfn, M):
    W = _o(wfn)
    L, R = OTensor(), OTensor()
    d = ctypes.c_double()
    _run(lib().orc_canonicalize_twosite, ctypes.c_int(1 if forward else 0), W.h, L.h, R.h, ctypes.c_int(M),
         ctypes.byref(d))
    return L.to_qst(), R.to_qst(), d.value


def dmrg(mpos, init_mps, Ms, tol):
    """Two-site DMRG on the CPU (one forward + one backward sweep per entry of Ms).  Returns (energies, n_sigma)."""
    L = len(mpos)
    mo = [_o(x) for x in mpos]
    io = [_o(x) for x in init_mps]
    am = (VP * L)(*[x.h for x in mo])
    ai = (VP * L)(*[x.h for x in io])
    Ms = np.array(Ms, dtype=np.int32)
    e = np.zeros(len(Ms))
    ns = ctypes.c_int32()
    _run(lib().orc_dmrg, ctypes.c_int(L), am, ai, ctypes.c_int(len(Ms)), Ms.ctypes.data_as(VP), ctypes.c_double(tol),
         e.ctypes.data_as(VP), ctypes.byref(ns))
    return e, ns.value


def sigma_conj(mpo, lopr, ropr, wfn, sgv=None):
    """compute_sigma_vector_conj (src/compute_sigma_vector.hpp:38-51), result in sgv's own quantum numbers."""
    out = OTensor(sgv) if sgv is not None else OTensor()
    ins = [_o(x) for x in (mpo, lopr, ropr, wfn)]
    _run(lib().orc_sigma_conj, *[x.h for x in ins], out.h)
    return out.to_qst()


def make_noise(wfn, noise):
    """make_noise (src/canonicalize.hpp:11-24) on a copy of wfn."""
    W = OTensor(wfn)
    _run(lib().orc_make_noise, W.h, ctypes.c_double(noise))
    return W.to_qst()


def canonicalize_twosite_ex(forward, wfn, M, noise=0.0, rule=0):
    """canonicalize (src/canonicalize.hpp:27-46) with noise; rule 0 = the reference's truncation
    (QSTLAPACK.h:209-213), 1 = the stable variant (tie band, 1e-16 floor, numerical rank)."""
    W = _o(wfn)
    L, R = OTensor(), OTensor()
    d = ctypes.c_double()
    _run(lib().orc_canonicalize_twosite_ex, ctypes.c_int(1 if forward else 0), W.h, L.h, R.h, ctypes.c_int(M),
         ctypes.c_double(noise), ctypes.c_int(rule), ctypes.byref(d))
    return L.to_qst(), R.to_qst(), d.value


def truncation_cutoff(values, M, rule=0):
    v = np.ascontiguousarray(values, dtype=np.float64)
    c = ctypes.c_double()
    _run(lib().orc_truncation_cutoff, ctypes.c_int(v.size), v.ctypes.data_as(VP), ctypes.c_int(M), ctypes.c_int(rule),
         ctypes.byref(c))
    return c.value


def merged_operator(forward, mpo, opr):
    """compute_merged_operator for one group (src/optimize.hpp:296-368)."""
    out = OTensor()
    M, O = _o(mpo), _o(opr)
    _run(lib().orc_merged_operator, ctypes.c_int(1 if forward else 0), M.h, O.h, out.h)
    return out.to_qst()


def qst_merge(x, nrow, ncol):
    """QSTmerge of the first nrow and the last ncol indices (include/btas/QSPARSE/QSTmerge.h)."""
    out = OTensor()
    X = _o(x)
    _run(lib().orc_qst_merge, X.h, ctypes.c_int(nrow), ctypes.c_int(ncol), out.h)
    return out.to_qst()


def mt19937_uniform(n):
    """First n draws of std::uniform_real_distribution<double>(-1, 1) on a default-seeded std::mt19937."""
    out = np.zeros(n)
    _run(lib().orc_mt19937_uniform, ctypes.c_int(n), out.ctypes.data_as(VP))
    return out


def merged_overlap(forward, mpo, sov):
    """Merged overlap S~ = merge(S (x) I_site) of compute_merged_operator (src/optimize.hpp:326-328, 363-365)."""
    out = OTensor()
    M, S = _o(mpo), _o(sov)
    _run(lib().orc_merged_overlap, ctypes.c_int(1 if forward else 0), M.h, S.h, out.h)
    return out.to_qst()
