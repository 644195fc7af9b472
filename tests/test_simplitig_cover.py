"""Simplitig writer against the oracle's restatement of the reference greedy
(for_each_simplitig_encoded, src/lib.rs:1478-1533), byte for byte, on sets
built to reach every case of its chain layout: dense small-k graphs (hairpins,
self loops, branches everywhere), even k (palindromic k-mers), circular
sequences (cycles of forced links), reverse-complement repeats, wide k, and
sets large enough for the multi-threaded passes."""
import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import canonical_kmers_np


def _arr(hi, lo):
    a = np.zeros(len(lo), dtype=M.COUNTED)
    a["hi"], a["lo"], a["count"] = hi, lo, 1
    a = np.unique(a)
    a.sort(order=["hi", "lo"])
    return a


def _same(tmp_path, a, k, first=0):
    lm = M.write_fasta_after_index(tmp_path / "m.fa", a, k, first)
    lo = O.write_simplitigs(str(tmp_path / "o.fa"), a, k, first)
    assert lm == lo
    assert (tmp_path / "m.fa").read_bytes() == (tmp_path / "o.fa").read_bytes()


def _all_canonical(k):
    v = np.arange(4 ** k, dtype=np.uint64)
    codes = np.stack([(v >> np.uint64(2 * (k - 1 - j))) & np.uint64(3) for j in range(k)], axis=1)
    r = np.zeros_like(v)
    for j in range(k):
        r = (r << np.uint64(2)) | (np.uint64(3) - codes[:, k - 1 - j])
    return np.unique(np.minimum(v, r))


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7])
def test_dense_small_k(tmp_path, k):
    allc = _all_canonical(k)
    rng = np.random.default_rng(k)
    for density in (0.1, 0.3, 0.5, 0.7, 0.9, 1.0):
        for _ in range(12):
            keep = allc[rng.random(allc.size) < density]
            _same(tmp_path, _arr(np.zeros(keep.size, np.uint64), keep), k, int(rng.integers(0, 5)))


def _genome_set(rng, k, glen, circular=False, rc_repeats=0, repeats=0, errors=0):
    g = rng.integers(0, 4, size=glen).astype(np.uint8)
    parts = [g]
    for _ in range(repeats):
        a = int(rng.integers(0, glen - 200))
        parts.append(g[a:a + int(rng.integers(k, 200))])
    for _ in range(rc_repeats):
        a = int(rng.integers(0, glen - 200))
        parts.append((3 - g[a:a + int(rng.integers(k, 200))])[::-1])
    if circular:
        parts = [np.concatenate([g, g[:k - 1]])]
    hs, ls = [], []
    for p in parts:
        p = p.copy()
        for _ in range(errors):
            p[int(rng.integers(0, p.size))] = rng.integers(0, 4)
        h, l = canonical_kmers_np(p, k)
        hs.append(h)
        ls.append(l)
    return _arr(np.concatenate(hs), np.concatenate(ls))


@pytest.mark.parametrize("k", [9, 12, 15, 21, 31, 32, 33, 40, 47, 63, 64])
def test_genome_like_sets(tmp_path, k):
    rng = np.random.default_rng(100 + k)
    for trial in range(6):
        a = _genome_set(rng, k, 3000, circular=(trial == 0), rc_repeats=trial, repeats=trial % 3,
                        errors=trial)
        _same(tmp_path, a, k, trial)


def test_hairpin_and_cycles(tmp_path):
    rng = np.random.default_rng(7)
    for k in (5, 7, 9, 11, 13):
        for _ in range(20):
            half = rng.integers(0, 4, size=int(rng.integers(k, 4 * k))).astype(np.uint8)
            # a sequence followed by its own reverse complement folds back on itself
            s = np.concatenate([half, (3 - half)[::-1]])
            h1, l1 = canonical_kmers_np(s, k)
            c = rng.integers(0, 4, size=int(rng.integers(k, 5 * k))).astype(np.uint8)
            h2, l2 = canonical_kmers_np(np.concatenate([c, c[:k - 1]]), k)
            _same(tmp_path, _arr(np.concatenate([h1, h2]), np.concatenate([l1, l2])), k)


@pytest.mark.parametrize("k", [21, 31, 41])
def test_large_sets_threaded(tmp_path, k):
    # above the size where the neighbour and chain passes run on many threads
    rng = np.random.default_rng(k)
    a = _genome_set(rng, k, 150_000, rc_repeats=40, repeats=40, errors=300)
    assert a.size > 1 << 17
    _same(tmp_path, a, k)


def test_text_and_empty(tmp_path):
    # header numbering continues after first_index, in both writers
    a = _genome_set(np.random.default_rng(3), 31, 500)
    last = M.write_fasta_after_index(tmp_path / "m.fa", a, 31, 41)
    assert last == O.write_simplitigs(str(tmp_path / "o.fa"), a, 31, 41) > 41
    assert (tmp_path / "m.fa").read_bytes() == (tmp_path / "o.fa").read_bytes()
    assert (tmp_path / "m.fa").read_text().startswith(">MC_simplitig_42 ")
    e = np.zeros(0, dtype=M.COUNTED)
    assert M.write_fasta_after_index(tmp_path / "e.fa", e, 31, 4) == 4
    assert (tmp_path / "e.fa").read_bytes() == b""


def _fasta_seqs(path):
    seqs, cur = [], None
    for line in open(path):
        line = line.rstrip("\n")
        if line.startswith(">"):
            if cur is not None:
                seqs.append("".join(cur))
            cur = []
        else:
            cur.append(line)
    if cur is not None:
        seqs.append("".join(cur))
    return seqs


def _canon_all(seqs, k):
    out = []
    for s in seqs:
        for i in range(len(s) - k + 1):
            out.append(O.encode_canonical(s[i:i + k]))
    return sorted(out)


def test_simplitig_sequences_are_the_fasta_records(tmp_path):
    rng = np.random.default_rng(21)
    for k in (3, 7, 21, 31, 45):
        a = _genome_set(rng, k, 2000, rc_repeats=3, repeats=2, errors=4)
        O.write_simplitigs(str(tmp_path / "o.fa"), a, k)
        assert M.simplitig_sequences(a, k) == _fasta_seqs(tmp_path / "o.fa"), k
    assert M.simplitig_sequences(np.zeros(0, dtype=M.COUNTED), 31) == []


def test_parallel_simplitigs_preserve_unique_kmer_set():
    # the reference's own case (src/lib.rs:5395-5416): k = 3, minimizer 2
    kms = [b"AAA", b"AAC", b"ACG", b"CGT", b"AAC"]
    a = np.zeros(len(kms), dtype=M.COUNTED)
    a["lo"] = [O.encode_canonical(x.decode()) for x in kms]
    a["count"] = 1
    a.sort(order=["hi", "lo"])
    want = sorted({O.encode_canonical(x.decode()) for x in kms})
    for parts in (0, 1, 2, 1024):
        assert _canon_all(M.simplitig_sequences_parallel(a, 3, 2, parts), 3) == want, parts


@pytest.mark.parametrize("k,m", [(5, 3), (21, 11), (31, 19), (45, 23)])
def test_parallel_simplitigs_cover_each_kmer_once(k, m):
    rng = np.random.default_rng(k)
    a = _genome_set(rng, k, 20_000, rc_repeats=5, repeats=5, errors=20)
    want = sorted((int(h) << 64) | int(l) for h, l in zip(a["hi"], a["lo"]))
    for parts in (1, 2, 16, 1024):
        got = M.simplitig_sequences_parallel(a, k, m, parts)
        assert _canon_all(got, k) == want, parts
    # one partition is the plain greedy
    assert M.simplitig_sequences_parallel(a, k, m, 1) == M.simplitig_sequences(a, k)


def test_parallel_simplitigs_validation():
    a = _genome_set(np.random.default_rng(2), 9, 100)
    with pytest.raises(M.MczaError, match="minimizer size must be > 0 and <= k"):
        M.simplitig_sequences_parallel(a, 9, 0)
    with pytest.raises(M.MczaError, match="minimizer size must be > 0 and <= k"):
        M.simplitig_sequences_parallel(a, 9, 10)
    with pytest.raises(M.MczaError, match="k must be <= 64"):
        M.simplitig_sequences(a, 65)
