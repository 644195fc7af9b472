"""On-disk formats of the reference's partitioned path: packed-read cache
(src/lib.rs:3394-3450), super-k-mer partition records (3209-3216,
2692-2711, 2756-2769) and counted k-mer records (2413-2421, 2657-2678).
Files are decoded here byte by byte from the layout the reference's readers
expect, and checked against the oracle's parser."""
import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import random_seqs, write_fasta


def _unpack(b, n):
    """n packed-seq codes from bytes, 4 per byte from bit 0."""
    a = np.frombuffer(b, dtype=np.uint8)
    c = np.stack([(a >> (2 * i)) & 3 for i in range(4)], axis=1).reshape(-1)
    return c[:n].astype(np.uint8)


def _codes(words, off):
    n = int(off[-1])
    i = np.arange(n, dtype=np.uint64)
    return ((words[(i >> np.uint64(4)).astype(np.int64)] >> ((i & np.uint64(15)) * np.uint64(2)).astype(np.uint32))
            & 3).astype(np.uint8)


@pytest.mark.parametrize("min_len", [1, 31])
def test_read_cache_layout_and_round_trip(tmp_path, min_len):
    rng = np.random.default_rng(min_len)
    seqs = random_seqs(rng, 200, 1, 700, alphabet="ACGTACGTACGTN")
    fa = tmp_path / "r.fa"
    write_fasta(fa, seqs)
    M.write_read_cache(fa, tmp_path / "r.pread", min_len)
    b = (tmp_path / "r.pread").read_bytes()
    assert b[:8] == b"MCRD0001"
    want_c, want_o = O.parse_fastx(str(fa), min_len)
    at, got_c, lens = 8, [], []
    while at < len(b):
        n = int.from_bytes(b[at:at + 4], "big")
        at += 4
        nb = (n + 3) // 4
        got_c.append(_unpack(b[at:at + nb], n))
        if n % 4:  # unused high bits of the last byte are zero
            assert b[at + nb - 1] >> (2 * (n % 4)) == 0
        at += nb
        lens.append(n)
    assert at == len(b)
    assert np.array_equal(np.cumsum([0] + lens), want_o)
    assert np.array_equal(np.concatenate(got_c) if got_c else np.zeros(0, np.uint8), want_c)
    w, off = M.read_read_cache(tmp_path / "r.pread", min_len)
    pw, poff = M.pack_fastx(fa, min_len)
    assert np.array_equal(off, poff) and np.array_equal(w, pw)
    # a larger min_len on reading skips the short ones, as the reference's reader
    w2, off2 = M.read_read_cache(tmp_path / "r.pread", 100)
    assert list(np.diff(off2)) == [x for x in lens if x >= 100]


def test_read_cache_rejects_bad_files(tmp_path):
    (tmp_path / "a").write_bytes(b"MCRD0002")
    with pytest.raises(M.MczaError, match="invalid packed-read cache header"):
        M.read_read_cache(tmp_path / "a")
    (tmp_path / "b").write_bytes(b"MCRD0001" + (10).to_bytes(4, "big") + b"\x00")
    with pytest.raises(M.MczaError, match="truncated packed-read cache record"):
        M.read_read_cache(tmp_path / "b")


def _skm_record(codes):
    n = len(codes)
    head = bytes([n]) if n < 255 else b"\xff" + n.to_bytes(4, "big")
    pad = np.concatenate([codes, np.zeros((-n) % 4, np.uint8)]).reshape(-1, 4).astype(np.uint32)
    body = (pad[:, 0] | (pad[:, 1] << 2) | (pad[:, 2] << 4) | (pad[:, 3] << 6)).astype(np.uint8).tobytes()
    return head + body


def test_superkmer_records_read_as_fragments(tmp_path):
    rng = np.random.default_rng(5)
    recs = [rng.integers(0, 4, size=n).astype(np.uint8) for n in (31, 254, 255, 256, 1000, 3, 57, 1)]
    (tmp_path / "partition_0000.skm").write_bytes(b"".join(_skm_record(c) for c in recs[:5]))
    (tmp_path / "partition_0001.skm").write_bytes(b"".join(_skm_record(c) for c in recs[5:]))
    paths = [tmp_path / "partition_0000.skm", tmp_path / "partition_0001.skm"]
    for min_len in (1, 31):
        keep = [c for c in recs if len(c) >= min_len]
        w, off = M.read_superkmer_partitions(paths, min_len)
        assert list(np.diff(off)) == [len(c) for c in keep]
        assert np.array_equal(_codes(w, off), np.concatenate(keep))
    (tmp_path / "bad.skm").write_bytes(b"\xff\x00\x00")
    with pytest.raises(M.MczaError, match="truncated super-kmer record"):
        M.read_superkmer_partitions([tmp_path / "bad.skm"])


def test_counted_records(tmp_path):
    rng = np.random.default_rng(2)
    a = np.zeros(500, dtype=M.COUNTED)
    a["lo"] = rng.integers(0, 2**63, size=500, dtype=np.uint64)
    a["hi"] = rng.integers(0, 2**62, size=500, dtype=np.uint64)
    a["count"] = rng.integers(1, 2**40, size=500, dtype=np.uint64)
    M.write_counted(tmp_path / "p.cnt", a)
    b = (tmp_path / "p.cnt").read_bytes()
    assert len(b) == 24 * 500
    for i in (0, 17, 499):
        r = b[24 * i:24 * (i + 1)]
        assert int.from_bytes(r[:16], "big") == (int(a[i]["hi"]) << 64) | int(a[i]["lo"])
        assert int.from_bytes(r[16:], "big") == int(a[i]["count"])
    back = M.read_counted(tmp_path / "p.cnt")
    assert np.array_equal(back, a)
    (tmp_path / "t.cnt").write_bytes(b[:30])
    with pytest.raises(M.MczaError, match="truncated counted k-mer record"):
        M.read_counted(tmp_path / "t.cnt")
