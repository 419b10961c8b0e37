"""The drop-in C++ header include/mpsxx_amd/mpsxx.hpp (the reference's template signatures over a QSTArray-like
container).  tests/adapter/mock_qstarray.hpp stands in for btas::QSTArray (the API subset the header uses);
tests/adapter/sigma_adapter.cpp calls the templates as src/optimize.hpp:60-75 does.

* CPU: the header compiles against the container API and links against the library's exported symbols;
* GPU: sigma (twice, accumulating), diagonal elements into a structure-only container, the two-group diagonal sum,
  a forward renormalization, compute_merged_operator over two groups in both directions (operators and overlaps) and
  canonicalize with noise through the header agree with the oracle (identical block sets, 1e-12 of the output
  scale)."""
import os
import subprocess

import numpy as np
import pytest

import mpsxx_amd as mx
import oracle
from helpers import hubbard_problem
from mpsxx_amd import mps
from mpsxx_amd.qst import QST

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def _build(tmp_path):
    libdir = os.path.join(REPO, "mpsxx_amd", "lib")
    if not os.path.exists(os.path.join(libdir, "libmpsxx_amd.so")):
        pytest.skip("libmpsxx_amd.so not built")
    exe = str(tmp_path / "sigma_adapter")
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(HERE, "adapter"),
           os.path.join(HERE, "adapter", "sigma_adapter.cpp"), "-o", exe, "-L" + libdir, "-lmpsxx_amd",
           "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


def _write(f, a):
    f.write("%d %d %d\n" % (a.rank, a.q[0], a.q[1]))
    for qs, ds in zip(a.qshape, a.dshape):
        f.write("%d %s\n" % (len(ds), " ".join("%d %d %d" % (q[0], q[1], d) for q, d in zip(qs, ds))))
    keys = a.sorted_keys()
    f.write("%d\n" % len(keys))
    for k in keys:
        f.write(" ".join(map(str, k)) + " " + " ".join(repr(float(x)) for x in a.blocks[k].ravel()) + "\n")


def _read(tok):
    rank, qp, qs = int(next(tok)), int(next(tok)), int(next(tok))
    qshape, dshape = [], []
    for _ in range(rank):
        n = int(next(tok))
        sec = [(int(next(tok)), int(next(tok)), int(next(tok))) for _ in range(n)]
        qshape.append([(p, s) for p, s, _ in sec])
        dshape.append([d for _, _, d in sec])
    a = QST((qp, qs), qshape, dshape)
    for _ in range(int(next(tok))):
        idx = tuple(int(next(tok)) for _ in range(rank))
        n = int(np.prod(a.block_shape(idx)))
        a.insert(idx, np.array([float(next(tok)) for _ in range(n)]))
    return a


def test_header_compiles_and_links(tmp_path):
    _build(tmp_path)


def _same(got, ref):
    assert got.block_set() == ref.block_set()
    assert got.q == ref.q and got.qshape == ref.qshape and got.dshape == ref.dshape
    scale = max(1.0, max((float(np.max(np.abs(b))) for b in ref.blocks.values() if b.size), default=1.0))
    assert got.max_abs_diff(ref) <= 1e-12 * scale


@pytest.mark.gpu
def test_header_calls_match_oracle(tmp_path):
    print("adapter header", flush=True)
    exe = _build(tmp_path)
    L, i = 8, 3
    rng = np.random.default_rng(31)
    mpos, ch, lenv, renv = hubbard_problem(L, i, rng, dmax=4)
    psi = ch["psi"]
    nq = mps.left_bond_qns(L, i + 1, (L, L % 2))
    ket = mps.fill_allowed(mps.site_tensor(ch["lq"][i], ch["ld"][i], nq, [2] * len(nq)), rng)
    from helpers import chain_overlaps
    lsov, rsov = chain_overlaps(ch, i, L)
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    with open(inp, "w") as f:
        for a in (mpos[i], mpos[i + 1], lenv, renv, psi, ket, lsov, rsov):
            _write(f, a)
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tok = iter(open(out).read().split())
    sgv, diag, lnext, diag2 = _read(tok), _read(tok), _read(tok), _read(tok)
    mlh0, mlh1, mls, mrh1, mrs, lcan, rcan, wnoisy = [_read(tok) for _ in range(8)]
    # compute_merged_operator: both groups and the overlaps (optimize.hpp:273-368)
    _same(mlh0, oracle.merged_operator(True, mpos[i], lenv))
    _same(mlh1, oracle.merged_operator(True, mpos[i], lenv))
    _same(mls, oracle.merged_overlap(True, mpos[i], lsov))
    _same(mrh1, oracle.merged_operator(False, mpos[i + 1], renv))
    _same(mrs, oracle.merged_overlap(False, mpos[i + 1], rsov))
    # canonicalize with noise and the reference truncation rule; wfnc keeps the noise (canonicalize.hpp:35)
    _same(wnoisy, oracle.make_noise(psi, 1e-3))
    Ur, Vr, _ = oracle.canonicalize_twosite_ex(True, psi, 12, 1e-3, 0)
    assert lcan.dshape == Ur.dshape and rcan.dshape == Vr.dshape and lcan.block_set() == Ur.block_set()
    rec = oracle.contract(1.0, lcan, [2], rcan, [0])
    _same(rec, oracle.contract(1.0, Ur, [2], Vr, [0]))

    ref = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, psi)
    ref2 = oracle.sigma_twosite(mpos[i], mpos[i + 1], lenv, renv, psi, sgv=ref.copy())
    _same(sgv, ref2)
    d0 = QST(psi.q, psi.qshape, psi.dshape)
    dref = oracle.diag_twosite(mpos[i], mpos[i + 1], lenv, renv, d0)
    _same(diag, dref)
    _same(diag2, dref.scaled(2.0))
    _same(lnext, oracle.renormalize(True, mpos[i], lenv, ket, ket))
    # the library's own ctypes path gives the same sigma
    dev = mx.compute_sigma_vector_twosite(*[mx.DTensor(x) for x in (mpos[i], mpos[i + 1], lenv, renv, psi)]).to_qst()
    assert dev.max_abs_diff(ref) <= 1e-12 * max(1.0, ref.norm())
