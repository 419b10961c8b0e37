"""Device tensors and sweep states on disk (reference src/fileio.h:9-25, src/make_sweep.cpp:46-247; needs an MI355X).

Bars: saved and reloaded arrays are bit-exact; a restarted sweep state reproduces the saved site tensors bit-exactly
and its later sweeps the energies of the uninterrupted run to 1e-12."""
import struct

import numpy as np
import pytest

import mpsxx_amd as mx
from mpsxx_amd import fileio
from mpsxx_amd.dmrg import DeviceDMRG, expectation, initial_mps
from mpsxx_amd.mpo import hubbard_mpo

pytestmark = pytest.mark.gpu

FNV_OFFSET, FNV_PRIME = 1469598103934665603, 1099511628211


def _fnv1a(b):
    h = FNV_OFFSET
    for x in b:
        h = ((h ^ x) * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def _same(a, b):
    assert a.q == b.q and a.qshape == b.qshape and a.dshape == b.dshape and a.block_set() == b.block_set()
    for k in a.blocks:
        assert np.array_equal(a.blocks[k], b.blocks[k])


def test_device_tensor_round_trips(tmp_path):
    print("fileio tensors", flush=True)
    rng = np.random.default_rng(4)
    mpos = hubbard_mpo(6, 1.0, 4.0)
    a = initial_mps(6, (6, 0), 16, 3)[2]
    for k in a.blocks:
        a.blocks[k] = rng.standard_normal(a.blocks[k].shape)
    d = mx.DTensor(a)
    fileio.save_tensor(d, tmp_path / "d.bin")
    _same(a, fileio.load_qst(tmp_path / "d.bin"))           # device -> file -> host
    fileio.save_qst(a, tmp_path / "h.bin")
    _same(a, fileio.load_tensor(tmp_path / "h.bin").to_qst())  # host -> file -> device
    out = mx.DTensor(mpos[1])                                # load replaces an existing structure
    fileio.load_tensor(tmp_path / "h.bin", out)
    _same(a, out.to_qst())
    fileio.save_tensor(mx.DTensor(), tmp_path / "e.bin")     # size() == 0 array
    assert fileio.load_tensor(tmp_path / "e.bin").to_qst() is None
    fileio.save_mpos([mx.DTensor(w) for w in mpos], str(tmp_path))
    for w, v in zip(mpos, fileio.load_mpos(6, str(tmp_path))):
        _same(w, v.to_qst())


def test_sweep_restart_reproduces_the_run(tmp_path):
    """Save after two sweeps, restart from the files, sweep both twice more: same site tensors, same energies."""
    print("fileio sweep restart", flush=True)
    L, M = 8, 64
    mpos = hubbard_mpo(L, 1.0, 4.0)
    dm = DeviceDMRG(mpos, initial_mps(L, (L, 0), M, 0), M)
    for _ in range(2):
        dm.sweep(M, 1e-10)
    prefix = str(tmp_path)
    dm.save(prefix)
    names = sorted(x.name for x in tmp_path.iterdir())
    assert names == sorted(["dmrg.0.0.0.0.tmp", "wave.0.0.0.0.tmp"] + ["rmps.%d.0.0.0.tmp" % k for k in range(1, L)])
    dm2 = DeviceDMRG.load(mpos, prefix, M)
    for k in range(L):
        _same(dm.site(k), dm2.site(k))
    dmpo = [mx.DTensor(w) for w in mpos]
    e_a = expectation(dmpo, [dm.site_device(k) for k in range(L)])
    e_b = expectation(dmpo, [dm2.site_device(k) for k in range(L)])
    assert e_a == e_b
    for _ in range(2):
        ea, _ = dm.sweep(M, 1e-10)
        eb, _ = dm2.sweep(M, 1e-10)
        print("E", ea, eb, flush=True)
        assert abs(ea - eb) <= 1e-12
    with pytest.raises(RuntimeError, match="sites"):
        DeviceDMRG.load(hubbard_mpo(L + 2, 1.0, 4.0), prefix, M)
    with pytest.raises(RuntimeError, match="cannot open"):
        DeviceDMRG.load(mpos, str(tmp_path / "nowhere"), M)
    with pytest.raises(RuntimeError, match="cannot open"):
        DeviceDMRG.load(mpos, prefix, M, iroot=1)


def test_restart_with_centre_inside_the_chain(tmp_path):
    """A state record with the centre at site 3 ("lmps" 0..2, "wave" 3, "rmps" 4..7, record written by hand in the
    documented layout) is brought to site 0 as mpsxx_dmrg_create does: the same sweeps as a driver created from
    the same site tensors."""
    print("fileio restart centre 3", flush=True)
    L, M, c = 8, 48, 3
    mpos = hubbard_mpo(L, 1.0, 4.0)
    init = initial_mps(L, (L, 0), M, 1)
    for k, a in enumerate(init):
        name = "lmps" if k < c else ("wave" if k == c else "rmps")
        fileio.save_qst(a, fileio.getfile(name, str(tmp_path), k, 0))
    body = b"MPSXXDMR" + struct.pack("<IIiii", 1, 0, 2, L, c)
    with open(fileio.getfile("dmrg", str(tmp_path), 0, 0), "wb") as f:
        f.write(body + struct.pack("<Q", _fnv1a(body)))
    dref = DeviceDMRG(mpos, init, M)
    dres = DeviceDMRG.load(mpos, str(tmp_path), M)
    for _ in range(2):
        ea, _ = dref.sweep(M, 1e-10)
        eb, _ = dres.sweep(M, 1e-10)
        print("E", ea, eb, flush=True)
        assert abs(ea - eb) <= 1e-12
