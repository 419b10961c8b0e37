"""Host-side block-sparse array with U(1)xU(1) quantum numbers (numpy), mirroring btas::QSTArray.

Reference: include/btas/QSPARSE/QSTArray.h:16-328 (allowed-block rule :34-36, conjugate :293-300),
include/btas/SPARSE/STArray.h:338-348 (row-major block tag).  Used to build inputs, to move arrays across the C ABI
(canonical layout: blocks in ascending tag order, each dense row-major) and to expand arrays densely in tests.
"""
import ctypes

import numpy as np


class QstDesc(ctypes.Structure):
    """Mirror of ``mpsxx_qst_desc`` (include/mpsxx_amd.h)."""

    _fields_ = [
        ("rank", ctypes.c_int32),
        ("qsym", ctypes.c_int32),
        ("q_total", ctypes.c_int32 * 2),
        ("nsec", ctypes.POINTER(ctypes.c_int32)),
        ("sec_q", ctypes.POINTER(ctypes.c_int32)),
        ("sec_dim", ctypes.POINTER(ctypes.c_int32)),
        ("nblocks", ctypes.c_int64),
        ("block_idx", ctypes.POINTER(ctypes.c_int32)),
    ]


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class QST:
    """Block-sparse array.  ``qshape[i]`` is a list of (particle, 2Sz) tuples, ``dshape[i]`` the sector dims."""

    def __init__(self, q, qshape, dshape, qsym=True):
        self.q = (int(q[0]), int(q[1]))
        self.qsym = bool(qsym)
        self.dshape = [list(map(int, d)) for d in dshape]
        if qshape is None:
            qshape = [[(0, 0)] * len(d) for d in self.dshape]
        self.qshape = [[(int(a), int(b)) for (a, b) in qs] for qs in qshape]
        assert len(self.qshape) == len(self.dshape)
        for qs, ds in zip(self.qshape, self.dshape):
            assert len(qs) == len(ds)
        self.blocks = {}

    # -- structure ---------------------------------------------------------------------------------------------------
    @property
    def rank(self):
        return len(self.dshape)

    def nsec(self, i):
        return len(self.dshape[i])

    def tag(self, idx):
        t = 0
        for i, k in enumerate(idx):
            t = t * self.nsec(i) + k
        return t

    def allowed(self, idx):
        if not self.qsym:
            return True
        p = sum(self.qshape[i][k][0] for i, k in enumerate(idx))
        s = sum(self.qshape[i][k][1] for i, k in enumerate(idx))
        return (p, s) == self.q

    def block_shape(self, idx):
        return tuple(self.dshape[i][k] for i, k in enumerate(idx))

    def insert(self, idx, arr):
        idx = tuple(int(x) for x in idx)
        if not self.allowed(idx):
            raise ValueError("block %r not allowed by quantum numbers" % (idx,))
        arr = np.asarray(arr, dtype=np.float64)
        if arr.shape != self.block_shape(idx):
            arr = arr.reshape(self.block_shape(idx))
        self.blocks[idx] = np.array(arr, dtype=np.float64, copy=True)

    def sorted_keys(self):
        return sorted(self.blocks.keys(), key=self.tag)

    def nelem(self):
        return int(sum(b.size for b in self.blocks.values()))

    def copy(self):
        c = QST(self.q, self.qshape, self.dshape, self.qsym)
        c.blocks = {k: v.copy() for k, v in self.blocks.items()}
        return c

    def conjugate(self):
        c = QST((-self.q[0], -self.q[1]), [[(-a, -b) for (a, b) in qs] for qs in self.qshape], self.dshape, self.qsym)
        c.blocks = dict(self.blocks)
        return c

    def empty_like(self):
        return QST(self.q, self.qshape, self.dshape, self.qsym)

    # -- ABI ---------------------------------------------------------------------------------------------------------
    def to_desc(self):
        """Return (desc, keepalive) for the C ABI."""
        nsec = np.array([self.nsec(i) for i in range(self.rank)], dtype=np.int32)
        secq = np.array([x for qs in self.qshape for (a, b) in qs for x in (a, b)], dtype=np.int32)
        if secq.size == 0:
            secq = np.zeros(2, dtype=np.int32)
        secd = np.array([d for ds in self.dshape for d in ds], dtype=np.int32)
        keys = self.sorted_keys()
        bidx = np.array(keys, dtype=np.int32).reshape(len(keys), self.rank) if keys else np.zeros((1, self.rank),
                                                                                                   dtype=np.int32)
        d = QstDesc()
        d.rank = self.rank
        d.qsym = 1 if self.qsym else 0
        d.q_total[0], d.q_total[1] = self.q
        d.nsec = _ptr(nsec, ctypes.c_int32)
        d.sec_q = _ptr(secq, ctypes.c_int32)
        d.sec_dim = _ptr(secd, ctypes.c_int32)
        d.nblocks = len(keys)
        d.block_idx = _ptr(bidx, ctypes.c_int32)
        return d, (nsec, secq, secd, bidx)

    def flat(self):
        keys = self.sorted_keys()
        if not keys:
            return np.zeros(0)
        return np.concatenate([self.blocks[k].ravel() for k in keys])

    @staticmethod
    def from_layout(rank, qsym, q, nsec, secq, secd, bidx, data):
        """Build from arrays returned by the C ABI (canonical layout)."""
        qshape, dshape = [], []
        k = 0
        for i in range(rank):
            qshape.append([(int(secq[2 * (k + j)]), int(secq[2 * (k + j) + 1])) for j in range(nsec[i])])
            dshape.append([int(secd[k + j]) for j in range(nsec[i])])
            k += nsec[i]
        a = QST(q, qshape, dshape, qsym)
        off = 0
        for b in range(len(bidx)):
            idx = tuple(int(x) for x in bidx[b])
            n = int(np.prod(a.block_shape(idx)))
            a.blocks[idx] = np.array(data[off:off + n]).reshape(a.block_shape(idx))
            off += n
        return a

    # -- dense expansion (tests) -------------------------------------------------------------------------------------
    def offsets(self, i):
        return np.concatenate([[0], np.cumsum(self.dshape[i])]).astype(int)

    def dense(self):
        shape = tuple(sum(d) for d in self.dshape)
        out = np.zeros(shape)
        offs = [self.offsets(i) for i in range(self.rank)]
        for idx, blk in self.blocks.items():
            sl = tuple(slice(offs[i][k], offs[i][k] + self.dshape[i][k]) for i, k in enumerate(idx))
            out[sl] = blk
        return out

    def block_set(self):
        return set(self.blocks.keys())

    def max_abs_diff(self, other):
        """Max |a - b| over the union of blocks (absent = zero)."""
        m = 0.0
        for k in set(self.blocks) | set(other.blocks):
            a = self.blocks.get(k)
            b = other.blocks.get(k)
            if a is None:
                a = np.zeros_like(b)
            if b is None:
                b = np.zeros_like(a)
            if a.size:
                m = max(m, float(np.max(np.abs(a - b))))
        return m

    def norm(self):
        return float(np.sqrt(sum(float(np.sum(b * b)) for b in self.blocks.values())))

    def scaled(self, alpha):
        c = self.copy()
        c.blocks = {k: alpha * v for k, v in c.blocks.items()}
        return c

    def diff_norm(self, other):
        """Frobenius norm of (self - other) over the union of blocks (absent = zero)."""
        t = 0.0
        for k in set(self.blocks) | set(other.blocks):
            a = self.blocks.get(k)
            b = other.blocks.get(k)
            if a is None:
                a = np.zeros_like(b)
            if b is None:
                b = np.zeros_like(a)
            t += float(np.sum((a - b) ** 2))
        return float(np.sqrt(t))


def random_fill(a, rng, scale=1.0, all_allowed=True, density=1.0):
    """Fill every allowed block (or a random subset of them) with N(0,1)*scale entries."""
    import itertools
    for idx in itertools.product(*[range(a.nsec(i)) for i in range(a.rank)]):
        if not a.allowed(idx):
            continue
        if not all_allowed and rng.random() > density:
            continue
        shp = a.block_shape(idx)
        a.blocks[tuple(idx)] = rng.standard_normal(shp) * scale
    return a
