"""Host ingest (mczi_pack_fastx): FASTA/FASTQ parse + 2-bit pack, against
the oracle's parser (for_fastx_packed_batches, src/lib.rs:3264-3308) on
well-formed files, and the many-thread FASTA path against the one-pass path
on anything, with chunk boundaries forced onto every byte position."""
import os

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O

CHUNK_ENV = "MCZI_PARSE_CHUNK_BYTES"


def _codes(words, off):
    n = int(off[-1])
    i = np.arange(n, dtype=np.uint64)
    return ((words[(i >> np.uint64(4)).astype(np.int64)] >> ((i & np.uint64(15)) * np.uint64(2)).astype(np.uint32))
            & 3).astype(np.uint8)


def _pack(path, min_len, chunk=None):
    old = os.environ.pop(CHUNK_ENV, None)
    try:
        if chunk is not None:
            os.environ[CHUNK_ENV] = str(chunk)
        w, off = M.pack_fastx(path, min_len)
    finally:
        os.environ.pop(CHUNK_ENV, None)
        if old is not None:
            os.environ[CHUNK_ENV] = old
    nw = (int(off[-1]) + 15) // 16
    assert w.size == nw
    if off[-1] % 16:
        assert w[-1] >> np.uint32(2 * (int(off[-1]) % 16)) == 0  # no stray bits past the end
    return _codes(w, off), off


def _fasta(rng, n_rec, crlf=False, wrap=60):
    out = []
    nl = "\r\n" if crlf else "\n"
    for r in range(n_rec):
        out.append(f">rec{r} some description{nl}")
        L = int(rng.integers(0, 400))
        alphabet = list("ACGTACGTACGTacgtNNRYK-")
        s = "".join(rng.choice(alphabet, size=L))
        for a in range(0, L, wrap):
            out.append(s[a:a + wrap] + nl)
        if rng.random() < 0.2:
            out.append(nl)
    return "".join(out)


def _same(path, min_len, chunks=(None, 1, 2, 3, 5, 7, 16, 61, 1000)):
    want_c, want_o = O.parse_fastx(str(path), min_len)
    for ch in chunks:
        got_c, got_o = _pack(path, min_len, ch)
        assert np.array_equal(got_o, want_o), ch
        assert np.array_equal(got_c, want_c), ch


@pytest.mark.parametrize("seed", range(6))
def test_fasta_matches_oracle(tmp_path, seed):
    rng = np.random.default_rng(seed)
    p = tmp_path / "a.fa"
    p.write_text(_fasta(rng, 40, crlf=seed % 2 == 1, wrap=[60, 80, 7, 1000][seed % 4]))
    for min_len in (1, 5, 31, 200):
        _same(p, min_len)


def test_fastq_matches_oracle(tmp_path):
    rng = np.random.default_rng(9)
    recs = []
    for r in range(300):
        s = "".join(rng.choice(list("ACGTNacgt"), size=int(rng.integers(1, 200))))
        q = "".join(rng.choice(list("!#+@ABCDEFGHIJ"), size=len(s)))
        recs.append(f"@r{r}\n{s}\n+\n{q}\n")
    p = tmp_path / "a.fq"
    p.write_text("".join(recs))
    for min_len in (1, 31):
        _same(p, min_len, chunks=(None, 3))


def test_many_thread_path_agrees_with_one_pass_on_odd_text(tmp_path):
    # '>' inside sequence lines, a header at end of file without newline,
    # empty records, leading blank lines: both paths must agree exactly.
    rng = np.random.default_rng(4)
    for t in range(40):
        parts = ["\n\n" if t % 3 == 0 else ""]
        for r in range(int(rng.integers(1, 12))):
            parts.append(">h" + "".join(rng.choice(list("x>AC \t"), size=int(rng.integers(0, 9)))) + "\n")
            w = np.array([14, 14, 14, 14, 4, 4, 4, 4, 4, 10, 8, 4, 3, 3], dtype=float)
            body = "".join(rng.choice(list("ACGTacgtN>\n\r.*"), p=w / w.sum(), size=int(rng.integers(0, 120))))
            parts.append(body)
        if t % 4 == 1:
            parts.append(">tail-without-newline")
        p = tmp_path / "o.fa"
        p.write_text("".join(parts))
        for min_len in (1, 4):
            one, one_o = _pack(p, min_len, None)
            for ch in (1, 2, 3, 4, 9, 33):
                got, got_o = _pack(p, min_len, ch)
                assert np.array_equal(got_o, one_o), (t, ch)
                assert np.array_equal(got, one), (t, ch)


def test_large_file_default_chunks(tmp_path):
    # larger than the default chunk, so the default many-thread path runs
    rng = np.random.default_rng(11)
    g = rng.integers(0, 4, size=9_000_000).astype(np.uint8)
    lines = []
    for r in range(0, g.size, 150_000):
        lines.append(f">chunk{r}\n")
        s = np.array(list("ACGT"))[g[r:r + 150_000]]
        s[rng.integers(0, s.size, size=20)] = "N"
        seq = "".join(s)
        lines.extend(seq[a:a + 80] + "\n" for a in range(0, len(seq), 80))
    p = tmp_path / "big.fa"
    p.write_text("".join(lines))
    assert p.stat().st_size > (4 << 20)
    _same(p, 31, chunks=(None,))
