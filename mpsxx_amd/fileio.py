"""On-disk block-sparse arrays and the reference's site-file names (reference src/fileio.h:9-25, src/fileio.cpp:7-13).

The reference archives every QSTArray with boost (``save(obj, getfile(name, prefix, site, i, j))``); this library
writes its own self-describing format (magic "MPSXXQST", version 1, structure, FP64 data in canonical block order,
FNV-1a-64 checksum; layout in include/mpsxx_amd.h).  A damaged, truncated or foreign file raises ``MpsxxError``.

Host arrays (``QST``) are saved and loaded without a device; device tensors (``DTensor``) through the same format.
"""
import ctypes

import numpy as np

from . import DTensor, MpsxxError, VP, _check, lib
from .qst import QST, QstDesc


def _b(path):
    return str(path).encode()


def getfile(name, prefix, site=0, istate=0, jstate=0):
    """'[prefix]/[name].[site].[istate].[jstate].0.tmp' (mpsxx::getfile, src/fileio.cpp:7-13; last field rank 0)."""
    n = ctypes.c_int64()
    cap = 64 + 2 * (len(str(name)) + len(str(prefix)))
    buf = ctypes.create_string_buffer(cap)
    _check(lib().mpsxx_getfile(_b(name), _b(prefix), ctypes.c_int(site), ctypes.c_int(istate), ctypes.c_int(jstate),
                               buf, ctypes.c_int64(cap), ctypes.byref(n)))
    return buf.value.decode()


def save_qst(a, path):
    """Write host array ``a`` (``None`` writes an array with no structure, the reference's size() == 0)."""
    if a is None:
        d = QstDesc()
        d.rank = 0
        _check(lib().mpsxx_qst_save(_b(path), ctypes.byref(d), None, ctypes.c_int64(0)))
        return
    d, keep = a.to_desc()
    data = np.ascontiguousarray(a.flat(), dtype=np.float64)
    _check(lib().mpsxx_qst_save(_b(path), ctypes.byref(d), data.ctypes.data_as(VP), ctypes.c_int64(data.size)))


def file_info(path):
    """(rank, qsym, q_total, number of sectors, number of blocks, number of elements) of a saved array."""
    rank, qsym, ns, nb, ne = (ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64(),
                              ctypes.c_int64())
    q = (ctypes.c_int32 * 2)()
    _check(lib().mpsxx_qst_file_info(_b(path), ctypes.byref(rank), ctypes.byref(qsym), q, ctypes.byref(ns),
                                     ctypes.byref(nb), ctypes.byref(ne)))
    return rank.value, qsym.value, (q[0], q[1]), ns.value, nb.value, ne.value


def load_qst(path):
    """Read a saved array as a host ``QST`` (``None`` for an array with no structure)."""
    rank, qsym, q, ns, nb, ne = file_info(path)
    if rank == 0:
        return None
    nsec = np.zeros(rank, dtype=np.int32)
    secq = np.zeros(2 * max(ns, 1), dtype=np.int32)
    secd = np.zeros(max(ns, 1), dtype=np.int32)
    bidx = np.zeros((max(nb, 1), rank), dtype=np.int32)
    data = np.zeros(max(ne, 1))
    _check(lib().mpsxx_qst_file_read(_b(path), nsec.ctypes.data_as(VP), secq.ctypes.data_as(VP),
                                     secd.ctypes.data_as(VP), bidx.ctypes.data_as(VP), data.ctypes.data_as(VP),
                                     ctypes.c_int64(ne)))
    return QST.from_layout(rank, qsym, q, nsec, secq, secd, bidx[:nb], data[:ne])


def save_tensor(t, path):
    """Write device tensor ``t`` (synchronous)."""
    _check(lib().mpsxx_tensor_save(t.h, _b(path), None))


def load_tensor(path, out=None):
    """Read a saved array into a device tensor (``out``'s structure and data are replaced)."""
    out = DTensor() if out is None else out
    _check(lib().mpsxx_tensor_load(_b(path), out.h, None))
    return out


def save_mpos(mpos, prefix, name="mpo"):
    """One file per site, getfile(name, prefix, site) (src/naive_mpogen.cpp:115, src/gen_qc_operators.cpp:149)."""
    for i, w in enumerate(mpos):
        if isinstance(w, DTensor):
            save_tensor(w, getfile(name, prefix, i))
        else:
            save_qst(w, getfile(name, prefix, i))


def load_mpos(L, prefix, name="mpo", device=True):
    """The L site operators saved by ``save_mpos`` (device tensors, or host arrays with device=False)."""
    out = []
    for i in range(L):
        f = getfile(name, prefix, i)
        out.append(load_tensor(f) if device else load_qst(f))
    if any(x is None for x in out):
        raise MpsxxError("load_mpos: an operator file holds an empty array")
    return out
