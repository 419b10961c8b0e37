"""On-disk arrays (reference src/fileio.h:9-25 save/load, src/fileio.cpp:7-13 getfile) on the CPU: the host-only
entry points of the C ABI, bit-exact round trips, the documented byte layout read back by an independent parser, and
rejection of damaged files."""
import struct

import numpy as np
import pytest

import mpsxx_amd as mx
from mpsxx_amd import fileio
from mpsxx_amd.qst import QST

FNV_OFFSET, FNV_PRIME = 1469598103934665603, 1099511628211


def _fnv1a(b):
    h = FNV_OFFSET
    for x in b:
        h = ((h ^ x) * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def _random_qst(rng, rank, qsym=True, zero_dim=False):
    """Random array: 2-4 sectors per index with quantum numbers in a small range, a random subset of allowed blocks."""
    qshape, dshape = [], []
    for i in range(rank):
        ns = int(rng.integers(2, 5))
        # first sector (0, 0) on every index: the all-first block is always allowed
        qshape.append([(0, 0)] + [(int(rng.integers(-1, 2)), int(rng.integers(-1, 2))) for _ in range(ns - 1)])
        dshape.append([int(rng.integers(0 if zero_dim else 1, 5)) for _ in range(ns)])
    a = QST((0, 0), qshape, dshape, qsym)
    for idx in np.ndindex(*[len(d) for d in dshape]):
        if qsym and tuple(map(sum, zip(*[qshape[i][k] for i, k in enumerate(idx)]))) != (0, 0):
            continue
        if rng.random() < 0.6 or not any(idx):
            a.blocks[idx] = rng.standard_normal(a.block_shape(idx))
    return a


def _same(a, b):
    assert a.rank == b.rank and a.qsym == b.qsym and a.q == b.q
    if a.qsym:
        assert a.qshape == b.qshape
    assert a.dshape == b.dshape
    assert a.block_set() == b.block_set()
    for k in a.blocks:
        assert a.blocks[k].shape == b.blocks[k].shape
        assert np.array_equal(a.blocks[k], b.blocks[k])  # bit-exact


@pytest.mark.parametrize("rank,qsym,zero_dim", [(1, True, False), (2, True, False), (3, True, True),
                                                (4, True, False), (6, True, False), (3, False, False)])
def test_round_trip_bit_exact(tmp_path, rank, qsym, zero_dim):
    rng = np.random.default_rng(100 + rank)
    a = _random_qst(rng, rank, qsym, zero_dim)
    p = tmp_path / "a.bin"
    fileio.save_qst(a, p)
    b = fileio.load_qst(p)
    _same(a, b)
    rank_, qsym_, q, ns, nb, ne = fileio.file_info(p)
    assert (rank_, qsym_, q, nb, ne) == (a.rank, int(qsym), a.q, len(a.blocks), a.nelem())
    assert ns == sum(len(d) for d in a.dshape)


def test_empty_array_round_trip(tmp_path):
    p = tmp_path / "e.bin"
    fileio.save_qst(None, p)
    assert fileio.load_qst(p) is None
    assert fileio.file_info(p)[0] == 0


def test_array_without_blocks(tmp_path):
    a = QST((1, 1), [[(0, 0), (1, 1)], [(0, 0)]], [[2, 3], [4]])
    p = tmp_path / "nb.bin"
    fileio.save_qst(a, p)
    b = fileio.load_qst(p)
    _same(a, b)
    assert b.nelem() == 0


def _parse(raw):
    """Independent reader of the documented layout (little-endian)."""
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from("<" + fmt, raw, o)
        o += struct.calcsize("<" + fmt)
        return v

    assert raw[:8] == b"MPSXXQST"
    o = 8
    version, reserved = take("II")
    rank, qsym, qp, qs = take("iiii")
    nsec = take("%di" % rank)
    S = sum(nsec)
    secq = take("%di" % (2 * S))
    secd = take("%di" % S)
    (nb,) = take("q")
    bidx = np.array(take("%di" % (nb * rank)), dtype=np.int64).reshape(nb, rank)
    (ne,) = take("q")
    data = np.frombuffer(raw, dtype="<f8", count=ne, offset=o)
    o += 8 * ne
    (chk,) = take("Q")
    assert o == len(raw)
    return dict(version=version, reserved=reserved, rank=rank, qsym=qsym, q=(qp, qs), nsec=nsec, secq=secq,
                secd=secd, bidx=bidx, data=data, chk=chk)


def test_documented_layout(tmp_path):
    """Blocks inserted in scrambled order are written by ascending tag (row-major sector index), each row-major,
    after the structure fields, and the trailing word is FNV-1a-64 of every byte before it."""
    rng = np.random.default_rng(7)
    up, dn = [(0, 0), (1, 1), (1, -1)], [(0, 0), (-1, -1), (-1, 1), (-2, 0)]
    a = QST((0, 0), [up, up, dn], [[2, 3, 1], [1, 2, 2], [3, 2, 4, 1]])
    for idx in [(0, 0, 0), (1, 0, 1), (2, 0, 2), (0, 1, 1), (0, 2, 2), (1, 2, 3), (2, 1, 3)]:
        a.blocks[idx] = rng.standard_normal(a.block_shape(idx))
    keys = list(a.blocks)
    rng.shuffle(keys)
    a.blocks = {k: a.blocks[k] for k in keys}
    p = tmp_path / "l.bin"
    fileio.save_qst(a, p)
    raw = p.read_bytes()
    f = _parse(raw)
    assert f["version"] == 1 and f["reserved"] == 0 and f["rank"] == 3 and f["qsym"] == 1 and f["q"] == (0, 0)
    assert list(f["nsec"]) == [len(d) for d in a.dshape]
    assert list(f["secq"]) == [x for qs in a.qshape for q in qs for x in q]
    assert list(f["secd"]) == [d for ds in a.dshape for d in ds]
    tags = [int(np.ravel_multi_index(tuple(b), [len(d) for d in a.dshape])) for b in f["bidx"]]
    assert tags == sorted(tags) and len(set(tags)) == len(tags)
    expect = np.concatenate([a.blocks[tuple(int(x) for x in b)].ravel() for b in f["bidx"]])
    assert np.array_equal(f["data"], expect)
    assert f["chk"] == _fnv1a(raw[:-8])


def _expect_io_error(p, what):
    with pytest.raises(mx.MpsxxError, match=what) as e:
        fileio.load_qst(p)
    assert "error 7" in str(e.value)


def test_damaged_files_are_rejected(tmp_path):
    rng = np.random.default_rng(9)
    a = _random_qst(rng, 3)
    p = tmp_path / "d.bin"
    fileio.save_qst(a, p)
    raw = bytearray(p.read_bytes())
    # one flipped bit in the data
    bad = bytearray(raw)
    bad[-16] ^= 0x01
    (tmp_path / "flip.bin").write_bytes(bytes(bad))
    _expect_io_error(tmp_path / "flip.bin", "checksum")
    # truncated, and an extra trailing byte
    (tmp_path / "short.bin").write_bytes(bytes(raw[:-9]))
    _expect_io_error(tmp_path / "short.bin", "size|truncated")
    (tmp_path / "long.bin").write_bytes(bytes(raw) + b"\0")
    _expect_io_error(tmp_path / "long.bin", "size|unexpected|checksum")
    # foreign magic, unknown version
    bad = bytearray(raw)
    bad[0:8] = b"NOTMPSXX"
    (tmp_path / "magic.bin").write_bytes(bytes(bad))
    _expect_io_error(tmp_path / "magic.bin", "not a MPSXXQST")
    bad = bytearray(raw)
    bad[8:12] = struct.pack("<I", 2)
    (tmp_path / "ver.bin").write_bytes(bytes(bad))
    _expect_io_error(tmp_path / "ver.bin", "version")
    # structure edits caught before the checksum: blocks out of order, a block the quantum numbers forbid
    f = _parse(bytes(raw))
    nb = len(f["bidx"])
    assert nb >= 2
    off_b = 8 + 8 + 16 + 4 * 3 + 12 * sum(f["nsec"]) + 8
    bad = bytearray(raw)
    first, second = bytes(bad[off_b:off_b + 12]), bytes(bad[off_b + 12:off_b + 24])
    bad[off_b:off_b + 12], bad[off_b + 12:off_b + 24] = second, first
    (tmp_path / "order.bin").write_bytes(bytes(bad))
    _expect_io_error(tmp_path / "order.bin", "ascending")
    bad = bytearray(raw)
    bad[24:28] = struct.pack("<i", 3)  # total particle number 3: the stored blocks no longer fit
    (tmp_path / "qn.bin").write_bytes(bytes(bad))
    _expect_io_error(tmp_path / "qn.bin", "not allowed")
    # missing file
    _expect_io_error(tmp_path / "missing.bin", "cannot open")


def test_save_rejects_bad_arguments(tmp_path):
    a = QST((0, 0), [[(0, 0), (1, 0)], [(0, 0), (-1, 0)]], [[2, 2], [3, 3]])
    a.blocks[(0, 1)] = np.zeros((2, 3))  # (0,0)+(-1,0) != (0,0)
    with pytest.raises(mx.MpsxxError, match="not allowed"):
        fileio.save_qst(a, tmp_path / "x.bin")
    assert not (tmp_path / "x.bin").exists() and not (tmp_path / "x.bin.part").exists()
    with pytest.raises(mx.MpsxxError, match="cannot open"):
        fileio.save_qst(QST((0, 0), [[(0, 0)]], [[1]]), tmp_path / "no_such_dir" / "x.bin")


def test_overwrite_replaces_whole_file(tmp_path):
    rng = np.random.default_rng(3)
    big, small = _random_qst(rng, 4), _random_qst(rng, 2)
    p = tmp_path / "o.bin"
    fileio.save_qst(big, p)
    fileio.save_qst(small, p)
    _same(small, fileio.load_qst(p))
    assert not (tmp_path / "o.bin.part").exists()


def test_getfile_names():
    """mpsxx::getfile: prefix/name.site.istate.jstate.rank.tmp (src/fileio.cpp:10-12), rank 0."""
    assert fileio.getfile("wave", "/scr", 3, 1) == "/scr/wave.3.1.0.0.tmp"
    assert fileio.getfile("right-H", "scr/x", 11, 2, 5) == "scr/x/right-H.11.2.5.0.tmp"
    assert fileio.getfile("mpo", ".", 0) == "./mpo.0.0.0.0.tmp"
    long_prefix = "p" * 300
    assert fileio.getfile("lmps", long_prefix, 7) == long_prefix + "/lmps.7.0.0.0.tmp"


def test_mpo_files_on_host(tmp_path):
    from mpsxx_amd.mpo import hubbard_mpo
    mpos = hubbard_mpo(6, 1.0, 4.0)
    fileio.save_mpos(mpos, str(tmp_path))
    back = fileio.load_mpos(6, str(tmp_path), device=False)
    for a, b in zip(mpos, back):
        _same(a, b)
    assert sorted(x.name for x in tmp_path.iterdir()) == ["mpo.%d.0.0.0.tmp" % i for i in range(6)]
