"""Two-site DMRG driver on the device (Python side of the mpsxx_dmrg_* C ABI).

Initial MPS (reference src/make_random_mpss.cpp:20-94, 120-171): bond quantum numbers from
``generate_quantum_states`` (all products of site quantum numbers with particle number <= N, pruned to those that
can still reach the target, optionally capped at M sectors around the middle), every sector of dimension 1, total
QN 0 on every site except the last, which carries (N, 2Sz).  Entries: N(0,1) per element of every allowed block
from numpy's default_rng(seed) (the configuration's seeded initialisation).
"""
import ctypes

import numpy as np

from . import DTensor, _check, lib
from .mpo import SITE_Q
from .qst import QST


def generate_quantum_states(L, qt, max_blocks=0):
    """Bond quantum-number lists qb[0..L-1] (qb[i] = bond right of site i, left convention)."""
    site = list(SITE_Q)
    qb = [None] * L
    qb[0] = list(site)
    for i in range(1, L - 1):
        bare = sorted({(a[0] + b[0], a[1] + b[1]) for a in qb[i - 1] for b in site})
        qb[i] = [q for q in bare if q[0] <= qt[0]]
    qb[L - 1] = [tuple(qt)]
    for i in range(L - 1, 0, -1):
        qr = qb[i]
        qb[i - 1] = [q for q in qb[i - 1] if any((q[0] + n[0], q[1] + n[1]) in qr for n in site)]
        assert qb[i - 1]
    for i in range(L - 1):
        ql, qr = qb[i], qb[i + 1]
        if max_blocks > 0 and len(ql) > max_blocks:
            offs = (len(ql) - max_blocks) // 2
            qb[i] = ql[offs:offs + max_blocks]
            qb[i + 1] = [q for q in qr if any((l[0] + n[0], l[1] + n[1]) == q for l in qb[i] for n in site)]
            assert qb[i + 1]
    return qb


def initial_mps(L, qt, M, seed=0):
    rng = np.random.default_rng(seed)
    qb = generate_quantum_states(L, qt, M)
    mps = []
    ql = [(0, 0)]
    for i in range(L):
        qr = qb[i] if i < L - 1 else [(0, 0)]
        total = (0, 0) if i < L - 1 else tuple(qt)
        a = QST(total, [ql, SITE_Q, [(-p, -s) for (p, s) in qr]], [[1] * len(ql), [1] * 4, [1] * len(qr)])
        import itertools
        for idx in itertools.product(range(len(ql)), range(4), range(len(qr))):
            if a.allowed(idx):
                a.blocks[idx] = rng.standard_normal((1, 1, 1))
        nrm = a.norm()
        if nrm > 0:
            for k in a.blocks:
                a.blocks[k] /= nrm
        mps.append(a)
        ql = qr
    return mps


def _mt19937_uniform_pm1(rs):
    """One std::uniform_real_distribution<double>(-1, 1) draw of libstdc++ from the std::mt19937 behind the legacy
    numpy RandomState ``rs`` (seeded with an integer, it initialises the Mersenne Twister exactly as std::mt19937's
    seed constructor does): generate_canonical<double, 53> combines two 32-bit outputs, (x0 + x1 2^32) / 2^64 (a
    value that rounds to 1 becomes the largest double below 1), and the distribution returns a + (b - a) u."""
    x0, x1 = rs.randint(0, 2 ** 32, size=2, dtype=np.uint64)
    u = (float(x0) + float(x1) * 4294967296.0) / 18446744073709551616.0
    if u >= 1.0:
        u = float(np.nextafter(1.0, 0.0))
    return u * 2.0 - 1.0


def reference_initial_mps(L, qt, M):
    """The reference's initial MPS (src/make_random_mpss.cpp:97-171, identical in dmrg/make_random_mpss.cpp): a
    default-seeded std::mt19937 (seed 5489) and uniform_real_distribution(-1, 1) give ONE value per site, in site
    order, and every element of every allowed block of that site is set to it (QSTArray::resize's constant overload
    wins, :153,171), then the site tensor is normalized.  Bond sectors as generate_quantum_states (capped at M), every
    sector of dimension 1, total QN 0 on every site except the last, which carries (N, 2Sz)."""
    rs = np.random.RandomState(5489)
    qb = generate_quantum_states(L, qt, M)
    mps = []
    ql = [(0, 0)]
    import itertools
    for i in range(L):
        qr = qb[i] if i < L - 1 else [(0, 0)]
        total = (0, 0) if i < L - 1 else tuple(qt)
        a = QST(total, [ql, SITE_Q, [(-p, -s) for (p, s) in qr]], [[1] * len(ql), [1] * 4, [1] * len(qr)])
        v = _mt19937_uniform_pm1(rs)
        for idx in itertools.product(range(len(ql)), range(4), range(len(qr))):
            if a.allowed(idx):
                a.blocks[idx] = np.full((1, 1, 1), v)
        nrm = a.norm()
        if nrm > 0:
            for k in a.blocks:
                a.blocks[k] /= nrm
        mps.append(a)
        ql = qr
    return mps


class SweepStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("sigma_ms", ctypes.c_double), ("davidson_s", ctypes.c_double),
                ("svd_s", ctypes.c_double), ("renorm_s", ctypes.c_double), ("diag_s", ctypes.c_double),
                ("nsigma", ctypes.c_int32), ("max_bond", ctypes.c_int32), ("discarded_weight", ctypes.c_double),
                ("nsigma_fp32", ctypes.c_int32), ("pad", ctypes.c_int32)]


SWEEP_FOURINDEX, SWEEP_MERGED, SWEEP_ONESITE = 0, 1, 2


class DeviceDMRG:
    """Device-resident two-site DMRG over one operator group.

    mpos: host QSTs or device tensors (e.g. ``QcMpo.devices()``); form: SWEEP_FOURINDEX (dmrg/make_sweep.cpp) or
    SWEEP_MERGED (src/make_sweep.cpp, optimize_twosite_merged: the reference's quantum-chemistry path)."""

    def __init__(self, mpos, init_mps, M, form=SWEEP_FOURINDEX):
        L = len(mpos)
        self.L = L
        self._mpo = [w if isinstance(w, DTensor) else DTensor(w) for w in mpos]
        init = [DTensor(a) for a in init_mps]
        arr_m = (ctypes.c_void_p * L)(*[t.h.value for t in self._mpo])
        arr_i = (ctypes.c_void_p * L)(*[t.h.value for t in init])
        self.h = ctypes.c_void_p()
        lib().mpsxx_dmrg_last_error.restype = ctypes.c_char_p
        rc = lib().mpsxx_dmrg_create(ctypes.c_int(L), arr_m, arr_i, ctypes.c_int(M), ctypes.byref(self.h), None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_create: " + lib().mpsxx_dmrg_last_error().decode())
        if form != SWEEP_FOURINDEX:
            rc = lib().mpsxx_dmrg_set_form(self.h, ctypes.c_int(form))
            if rc != 0:
                raise RuntimeError("mpsxx_dmrg_set_form: " + lib().mpsxx_dmrg_last_error().decode())

    def sweep(self, M, tol, precision=0):
        e = ctypes.c_double()
        st = SweepStats()
        rc = lib().mpsxx_dmrg_sweep(self.h, ctypes.c_int(M), ctypes.c_double(tol), ctypes.c_int(precision),
                                    ctypes.byref(e), ctypes.byref(st), None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_sweep: " + lib().mpsxx_dmrg_last_error().decode())
        return e.value, {k: getattr(st, k) for k, _ in SweepStats._fields_}

    def set_form(self, form):
        """Sweep form of later sweeps: SWEEP_FOURINDEX, SWEEP_MERGED or SWEEP_ONESITE."""
        rc = lib().mpsxx_dmrg_set_form(self.h, ctypes.c_int(form))
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_set_form: " + lib().mpsxx_dmrg_last_error().decode())

    def sweep_trace(self):
        """Davidson solves of the last sweep: list of (site, forward, energy, [Ritz value per micro iteration],
        [||r||^2 tested after each])."""
        ns, ne = ctypes.c_int32(), ctypes.c_int32()
        rc = lib().mpsxx_dmrg_sweep_trace(self.h, 0, ctypes.byref(ns), None, None, None, None, 0, ctypes.byref(ne),
                                          None, None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_sweep_trace: " + lib().mpsxx_dmrg_last_error().decode())
        n, m = ns.value, ne.value
        site = np.zeros(max(n, 1), dtype=np.int32)
        fwd = np.zeros(max(n, 1), dtype=np.int32)
        en = np.zeros(max(n, 1))
        cnt = np.zeros(max(n, 1), dtype=np.int32)
        ev = np.zeros(max(m, 1))
        rn = np.zeros(max(m, 1))
        vp = ctypes.c_void_p
        rc = lib().mpsxx_dmrg_sweep_trace(self.h, ctypes.c_int32(n), ctypes.byref(ns), site.ctypes.data_as(vp),
                                          fwd.ctypes.data_as(vp), en.ctypes.data_as(vp), cnt.ctypes.data_as(vp),
                                          ctypes.c_int32(m), ctypes.byref(ne), ev.ctypes.data_as(vp),
                                          rn.ctypes.data_as(vp))
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_sweep_trace: " + lib().mpsxx_dmrg_last_error().decode())
        out, k = [], 0
        for i in range(n):
            out.append((int(site[i]), bool(fwd[i]), float(en[i]), [float(x) for x in ev[k:k + cnt[i]]],
                        [float(x) for x in rn[k:k + cnt[i]]]))
            k += cnt[i]
        return out

    def set_truncation(self, noise=0.0, rule=1):
        """Noise (DMRGInput::noise) and truncation rule (1 stable, 0 the reference's) of later sweeps."""
        rc = lib().mpsxx_dmrg_set_truncation(self.h, ctypes.c_double(noise), ctypes.c_int(rule))
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_set_truncation: " + lib().mpsxx_dmrg_last_error().decode())

    def add_lower_root(self, root, energy):
        """Project the converged state of DeviceDMRG `root` (energy E_k) out of later sweeps (excited states)."""
        rc = lib().mpsxx_dmrg_add_lower_root(self.h, root.h, ctypes.c_double(energy), None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_add_lower_root: " + lib().mpsxx_dmrg_last_error().decode())

    def save(self, prefix, iroot=0):
        """Write the state as the reference names it (src/make_sweep.cpp:124-247): "lmps" left of the centre, "wave"
        at it, "rmps" right of it, each at getfile(name, prefix, site, iroot), plus the record
        getfile("dmrg", prefix, 0, iroot) of nsites and the centre (``fileio.getfile``)."""
        rc = lib().mpsxx_dmrg_save(self.h, str(prefix).encode(), ctypes.c_int(iroot), None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_save: " + lib().mpsxx_dmrg_last_error().decode())

    @classmethod
    def load(cls, mpos, prefix, M, iroot=0, form=SWEEP_FOURINDEX):
        """Restart from a state written by ``save``: the saved site tensors, every right environment rebuilt on the
        device.  Settings (noise, lower roots) start at their defaults."""
        self = cls.__new__(cls)
        self.h = ctypes.c_void_p()
        self.L = len(mpos)
        self._mpo = [w if isinstance(w, DTensor) else DTensor(w) for w in mpos]
        arr_m = (ctypes.c_void_p * self.L)(*[t.h.value for t in self._mpo])
        lib().mpsxx_dmrg_last_error.restype = ctypes.c_char_p
        rc = lib().mpsxx_dmrg_load(ctypes.c_int(self.L), arr_m, str(prefix).encode(), ctypes.c_int(iroot),
                                   ctypes.c_int(M), ctypes.byref(self.h), None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_load: " + lib().mpsxx_dmrg_last_error().decode())
        if form != SWEEP_FOURINDEX:
            self.set_form(form)
        return self

    def mpo_gradient(self):
        """[<psi| dH/dW_k |psi>] for every site (DTensors; W_k's block indices with quantum numbers negated)."""
        grads = [DTensor() for _ in range(self.L)]
        arr = (ctypes.c_void_p * self.L)(*[g.h.value for g in grads])
        rc = lib().mpsxx_dmrg_mpo_gradient(self.h, arr, None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_mpo_gradient: " + lib().mpsxx_dmrg_last_error().decode())
        return grads

    def site_device(self, i):
        out = DTensor()
        rc = lib().mpsxx_dmrg_site(self.h, ctypes.c_int(i), out.h, None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_site: " + lib().mpsxx_dmrg_last_error().decode())
        return out

    def site(self, i):
        out = DTensor()
        rc = lib().mpsxx_dmrg_site(self.h, ctypes.c_int(i), out.h, None)
        if rc != 0:
            raise RuntimeError("mpsxx_dmrg_site: " + lib().mpsxx_dmrg_last_error().decode())
        return out.to_qst()

    def __del__(self):
        try:
            if self.h:
                lib().mpsxx_dmrg_free(self.h)
                self.h = ctypes.c_void_p()
        except Exception:
            pass


def expectation(mpos, mps):
    """<psi|O|psi>/<psi|psi> for device MPO and MPS site tensors (DTensors), src/expectation.hpp:55-114."""
    L = len(mps)
    am = (ctypes.c_void_p * L)(*[t.h.value for t in mpos])
    aa = (ctypes.c_void_p * L)(*[t.h.value for t in mps])
    v = ctypes.c_double()
    rc = lib().mpsxx_expectation(ctypes.c_int(L), am, aa, ctypes.byref(v), None)
    if rc != 0:
        raise RuntimeError("mpsxx_expectation: " + lib().mpsxx_dmrg_last_error().decode())
    return v.value
