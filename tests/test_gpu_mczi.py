"""MCZI in one process (src/mczi.rs:238-451) on the device: MC-counted index,
device membership index, ZI subtraction, against the oracle (MI355X only)."""
import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import revcomp, write_fasta

pytestmark = pytest.mark.gpu


def test_scan_stats_reference_case(tmp_path):
    # scan_query_stats_count_filtered_and_absent_kmers (src/mczi.rs:1896-1914):
    # index AAA, queries AAAT and CCCC -> scanned 4, filtered 1, not filtered 3
    (tmp_path / "i.fa").write_text(">i\nAAA\n")
    (tmp_path / "q1.fa").write_text(">q1\nAAAT\n")
    (tmp_path / "q2.fa").write_text(">q2\nCCCC\n")
    got, st = M.mczi_absent([tmp_path / "i.fa"], [tmp_path / "q1.fa", tmp_path / "q2.fa"], M.CounterConfig(3, 1, 0))
    assert (st["query_kmers_scanned"], st["query_kmers_filtered"], st["query_kmers_not_filtered"]) == (4, 1, 3)
    assert st["index_kmers"] == 1
    assert sorted(O.decode(int(r["lo"]), int(r["hi"]), 3) for r in got) == ["AAT", "CCC"]


def _genome(rng, n):
    return "".join(rng.choice(list("ACGT"), size=n))


def _reads(rng, src, n, L=150, err=0.01):
    out = []
    for _ in range(n):
        p = int(rng.integers(0, len(src) - L))
        r = list(src[p:p + L])
        for j in range(L):
            if rng.random() < err:
                r[j] = "ACGT"[("ACGT".index(r[j]) + 1) % 4]
        s = "".join(r)
        out.append(s if rng.random() < 0.5 else revcomp(s))
    return out


def _want(index_kmers, queries, k):
    """Absent set and window statistics by the oracle's set difference: the
    index is given to the oracle as one fragment per selected k-mer."""
    idx_seqs = [O.decode(int(r["lo"]), int(r["hi"]), k) for r in index_kmers]
    ic, io, _ = O.frags_of(idx_seqs if idx_seqs else ["A"])  # one base: no window, an empty index
    qc, qo, _ = O.frags_of(queries)
    absent = O.zi_absent(ic, io, qc, qo, k)
    present = {(int(r["hi"]) << 64) | int(r["lo"]) for r in index_kmers}
    scanned = not_filtered = 0
    for q in queries:
        for i in range(len(q) - k + 1):
            scanned += 1
            not_filtered += O.encode_canonical(q[i:i + k]) not in present
    return absent, scanned, not_filtered


@pytest.mark.parametrize("k,m,x", [(31, 21, 1), (21, 11, 0), (45, 27, 2)])
def test_mczi_matches_oracle(tmp_path, k, m, x):
    rng = np.random.default_rng(k)
    g, other = _genome(rng, 30_000), _genome(rng, 30_000)
    idx_reads = _reads(rng, g, 1500)  # ~7.5x coverage: thresholds matter
    idx_paths = []
    for i in range(2):
        p = tmp_path / f"idx{i}.fa"
        write_fasta(p, idx_reads[i::2])
        idx_paths.append(p)
    queries = _reads(rng, g, 150, err=0.005) + _reads(rng, other, 150) + ["ACGTN" * 40]
    qp = tmp_path / "q.fa"
    write_fasta(qp, queries)
    ic, io, _ = O.frags_of(idx_reads)
    index_kmers = O.count_bruteforce(ic, io, k, x)
    assert index_kmers.size > 1000
    want, scanned, not_filtered = _want(index_kmers, [s for q in queries for s in q.split("N") if s], k)
    cfg = M.CounterConfig(k, m, x, 64)
    got, st = M.mczi_absent(idx_paths, [qp], cfg)
    assert np.array_equal(got["lo"], want["lo"]) and np.array_equal(got["hi"], want["hi"])
    assert st["index_kmers"] == index_kmers.size
    assert (st["query_kmers_scanned"], st["query_kmers_not_filtered"]) == (scanned, not_filtered)
    assert st["query_kmers_filtered"] == scanned - not_filtered
    assert st["query_unique_kmers_not_filtered"] == want.size
    # the one-call form writes the absent set's simplitigs
    out = tmp_path / "o.fa"
    st2 = M.mczi_subtract(idx_paths, [qp], cfg, out)
    assert st2 == st
    O.write_simplitigs(str(tmp_path / "w.fa"), want, k)
    assert out.read_bytes() == (tmp_path / "w.fa").read_bytes()
    # no-output mode: statistics only
    assert M.mczi_subtract(idx_paths, [qp], cfg) == st


def test_mczi_dataset_presence_index(tmp_path):
    rng = np.random.default_rng(3)
    g = _genome(rng, 20_000)
    paths, seqs, ds = [], [], []
    for d in range(5):
        v = list(g)
        for _ in range(30):
            q = int(rng.integers(0, len(v)))
            v[q] = "ACGT"[("ACGT".index(v[q]) + 1) % 4]
        rs = _reads(rng, "".join(v), 400, err=0.0)
        p = tmp_path / f"d{d}.fa"
        write_fasta(p, rs)
        paths.append(p)
        seqs += rs
        ds += [d] * len(rs)
    queries = _reads(rng, g, 100, err=0.01)
    qp = tmp_path / "q.fa"
    write_fasta(qp, queries)
    codes, off, dsa = O.frags_of(seqs, datasets=ds)
    for x in (1, 3, 5):
        index_kmers = O.count_datasets_bruteforce(codes, off, dsa, 5, 31, x)
        want, scanned, not_filtered = _want(index_kmers, queries, 31)
        got, st = M.mczi_absent(paths, [qp], M.CounterConfig(31, 21, x, 64), index_datasets=True)
        assert np.array_equal(got["lo"], want["lo"]), x
        assert st["index_kmers"] == index_kmers.size, x
        assert (st["query_kmers_scanned"], st["query_kmers_not_filtered"]) == (scanned, not_filtered), x
