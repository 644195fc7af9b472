"""Pass 1 on the device (MI355X only): minimizer counts, dataset minimizer
presence, the two-pass filter, and the statistics that come from them.

- unique_minimizers is the reference's unique_hashes() of the pass-1 table
  (src/lib.rs:4630-4632), reported as minimizers_above_threshold by the
  k-mer FASTA path (src/lib.rs:326-351) and as Phase12Stats
  unique_minimizer_hashes (src/lib.rs:202);
- minimizers_retained with fof is the number of minimizers the dataset table
  keeps (src/lib.rs:759-760, 815);
- partition_bytes is the size of the reference's partition files
  (src/lib.rs:575): records of the staged runs (all runs single-pass, the
  runs above the threshold two-pass).
All are checked against the oracle's restatement of the reference's tables.
"""
import ctypes as C

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import pack_words, random_seqs, revcomp, same_counted, write_fasta

pytestmark = pytest.mark.gpu


def _inputs(seed, n=1500, L=150, long=True):
    rng = np.random.default_rng(seed)
    genome = "".join(rng.choice(list("ACGT"), size=30000))
    reads = []
    for _ in range(n):
        p = int(rng.integers(0, len(genome) - L))
        r = list(genome[p:p + L])
        for j in range(L):
            if rng.random() < 0.01:
                r[j] = "ACGT"[("ACGT".index(r[j]) + 1) % 4]
        s = "".join(r)
        reads.append(s if rng.random() < 0.5 else revcomp(s))
    if long:
        # runs that cross the scan's tile edges: homopolymers and repeats
        # longer than a tile, and long reads
        reads += ["A" * 7000, "AC" * 3000, genome[:12000], "G" * 1930 + "T" * 1930]
        reads += random_seqs(rng, 3, 4000, 9000)
    return reads


@pytest.mark.parametrize("k,m,x,P", [(31, 21, 1, 64), (31, 21, 5, 1024), (21, 11, 2, 16), (15, 9, 0, 7),
                                     (45, 27, 1, 64)])
def test_two_pass_stats_match_oracle(k, m, x, P):
    print(f"two-pass k={k} m={m} x={x} P={P}", flush=True)
    reads = _inputs(k * 7 + x)
    codes, off, _ = O.frags_of(reads)
    b = M.Batch.from_packed(pack_words(codes), off)
    want_counts, pb, kept = O.count_two_pass(codes, off, k, m, x, P)
    cfg = M.CounterConfig(k, m, x, P)
    for launch in (M.LAUNCH_BATCHED, M.LAUNCH_PLAIN):
        st = b.count(cfg, M.MODE_TWO_PASS, launch)
        assert same_counted(b.download(), want_counts)
        assert st.unique_minimizers == kept
        assert st.partition_bytes == pb, (st.partition_bytes, pb)
    # single pass: every run is staged (the two-pass bytes at x = 0, where
    # every minimizer passes)
    _, pb0, _ = O.count_two_pass(codes, off, k, m, 0, P)
    st = b.count(cfg, M.MODE_SINGLE_PASS)
    assert same_counted(b.download(), want_counts)
    assert st.partition_bytes == pb0


@pytest.mark.parametrize("n_datasets,x", [(7, 2), (33, 5), (100, 40)])
def test_dataset_two_pass_matches_oracle(n_datasets, x):
    print(f"fof two-pass datasets={n_datasets} x={x}", flush=True)
    r = M.SynthReads(M.SynthSpec(genome_len=20_000, genome_seed=9, read_seed=909, n_reads=40 * n_datasets,
                                 read_len=150, sub_rate=0.005, n_datasets=n_datasets, snp_rate=0.003))
    codes = r.codes()
    b = M.Batch.from_packed(r.words, r.frag_off, r.dataset)
    want, kept = O.count_datasets(codes, r.frag_off, r.dataset, n_datasets, 31, 21, x, 64)
    assert same_counted(want, O.count_datasets_bruteforce(codes, r.frag_off, r.dataset, n_datasets, 31, x))
    st = b.count(M.CounterConfig(31, 21, x, 64), M.MODE_DATASETS_TWO_PASS)
    assert same_counted(b.download(), want)
    assert st.minimizers_retained == kept


def test_kmer_fasta_stats_and_phase12(tmp_path):
    reads = _inputs(3, n=800)
    fa = tmp_path / "r.fa"
    write_fasta(fa, reads)
    codes, off = O.parse_fastx(str(fa), 31)
    _, pb, kept = O.count_two_pass(codes, off, 31, 21, 2, 64)
    st = M._FastaStats()
    M._check(M._lib.mczi_count_inputs_to_kmer_fasta_path(M._paths([fa])[0], 1, C.byref(M.CounterConfig(31, 21, 2, 64)._c()),
                                                          str(tmp_path / "o.fa").encode(), C.byref(st)))
    assert st.minimizers_above_threshold == kept
    assert st.selected_kmers == len(O.count_bruteforce(codes, off, 31, 2))
    ph = M.run_inputs_phase12_direct([fa], M.CounterConfig(31, 21, 2, 64))
    assert ph["unique_minimizer_hashes"] == kept and ph["partition_bytes"] == pb
    assert M.count_inputs_minimizer_phase_direct([fa], M.CounterConfig(31, 21, 2, 64)) == kept
    # the dataset path reports the retained minimizers
    paths = []
    for d in range(4):
        p = tmp_path / f"d{d}.fa"
        write_fasta(p, reads[d * 200:(d + 1) * 200])
        paths.append(p)
    dcodes, doff, dds = O.frags_of(reads[:800], datasets=np.repeat(np.arange(4), 200))
    _, dkept = O.count_datasets(dcodes, doff, dds, 4, 31, 21, 1, 64)
    M._check(M._lib.mczi_count_datasets_to_kmer_fasta_path(M._paths(paths)[0], 4,
                                                            C.byref(M.CounterConfig(31, 21, 1, 64)._c()),
                                                            str(tmp_path / "d.fa").encode(), C.byref(st)))
    assert st.minimizers_above_threshold == dkept
