"""mpsxx_dotc_noalloc against mpsxx_dotc through the C ABI (needs an MI355X).

Bar: bit-identical results (same kernels, same summation order), on two calls in a row with the second on a smaller
input, and no device allocation by the allocation-free calls (the library's allocation counters do not move)."""
import ctypes

import numpy as np
import pytest

import mpsxx_amd as mx
from helpers import random_twosite_problem

pytestmark = pytest.mark.gpu


def _dot(fn, x, y):
    r = ctypes.c_double()
    rc = fn(x.h, y.h, ctypes.byref(r), None)
    assert rc == 0, mx.lib().mpsxx_last_error()
    return r.value


def _alloc_counters():
    st = mx.device_memory_stats()
    return st["mallocs"], st["reuses"], st["bytes_live"]


def _pair(seed, dlo, dhi):
    """Two wavefunction-shaped arrays with partly different block sets (x lacks some of y's blocks)."""
    rng = np.random.default_rng(seed)
    _, _, _, _, psi = random_twosite_problem(12, 5, rng, lambda q: int(rng.integers(dlo, dhi + 1)))
    y = psi.copy()
    for k in y.blocks:
        y.blocks[k] = rng.standard_normal(y.blocks[k].shape)
    x = psi.copy()
    for k in list(x.blocks)[::3]:
        del x.blocks[k]
    return x, y


def test_dotc_noalloc_matches_dotc_twice_second_smaller():
    L = mx.lib()
    for seed, dlo, dhi in ((1, 20, 120), (2, 1, 12)):
        print("dotc_noalloc seed", seed, "dims", dlo, dhi, flush=True)
        xq, yq = _pair(seed, dlo, dhi)
        x, y = mx.DTensor(xq), mx.DTensor(yq)
        ref = _dot(L.mpsxx_dotc, x, y)
        before = _alloc_counters()
        got = _dot(L.mpsxx_dotc_noalloc, x, y)
        got_self = _dot(L.mpsxx_dotc_noalloc, y, y)
        assert _alloc_counters() == before, "the allocation-free dot allocated device memory"
        assert got == ref, (got, ref)
        assert got_self == _dot(L.mpsxx_dotc, y, y)
        # and against the host value (FP64 summation in another order: rounding-level agreement)
        host = sum(float(np.sum(xq.blocks[k] * yq.blocks[k])) for k in xq.blocks)
        assert abs(got - host) <= 1e-12 * max(1.0, float(np.sqrt(xq.norm() ** 2 * yq.norm() ** 2)))


def test_dotc_noalloc_empty_and_null_result():
    L = mx.lib()
    e = mx.DTensor()
    xq, yq = _pair(3, 2, 6)
    y = mx.DTensor(yq)
    assert _dot(L.mpsxx_dotc_noalloc, e, y) == 0.0
    assert L.mpsxx_dotc_noalloc(y.h, y.h, None, None) == 5  # MPSXX_ERR_ARG
