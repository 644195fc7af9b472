"""HIP path vs the oracle: minimizer scan, counting, KFF, fof (MI355X only).

Every comparison is bit-exact: same seeded inputs, the oracle's restatement
(or brute force) on the CPU, the library's device path through the C ABI.
"""
import os

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import as_pairs, random_seqs, revcomp, same_counted, write_fasta
from kff_reader import KffFile, kmers_u64

pytestmark = pytest.mark.gpu


def _pack(codes: np.ndarray) -> bytes:
    n = codes.size
    b = np.zeros((n + 3) // 4, dtype=np.uint8)
    for j in range(4):
        b[: (n - j + 3) // 4] |= (codes[j::4] << (2 * j)).astype(np.uint8)
    return b.tobytes()


@pytest.mark.parametrize("k,m", [(31, 21), (5, 3), (3, 2), (21, 11), (15, 15), (63, 31), (31, 19), (1, 1)])
def test_run_hashes_match_oracle(k, m):
    print(f"run_hashes k={k} m={m}", flush=True)
    rng = np.random.default_rng(k * 100 + m)
    seqs = random_seqs(rng, 6, k, 700) + ["A" * 300, "AC" * 200, "ACGTTGCA" * 50]
    seqs += ["".join(rng.choice(list("ACGT"), size=5000))]
    for s in seqs:
        if len(s) < k:
            continue
        codes = O.codes_of(s)
        h, st = M.minimizer_run_hashes(_pack(codes), len(s), k, m)
        oh, ost = O.minimizer_runs(codes, k, m)
        assert np.array_equal(h, oh) and np.array_equal(st, ost), (k, m, len(s))


def test_run_hashes_known_positions(golden):
    # the reference's position vector, through the device scan: the selected
    # hash of each window is the hash at the known position
    case = golden["minimizer_positions"][0]
    codes = O.codes_of(case["seq"])
    k = case["k"] + case["w"] - 1
    h, st = M.minimizer_run_hashes(_pack(codes), len(case["seq"]), k, case["k"])
    pos_hashes = [O.nthash32(codes[p:p + case["k"]], case["k"], True) for p in case["positions"]]
    assert list(h) == pos_hashes


def _reads(seed, genome_len=20000, n=2000, L=150, err=0.01):
    rng = np.random.default_rng(seed)
    genome = "".join(rng.choice(list("ACGT"), size=genome_len))
    reads = []
    for _ in range(n):
        p = int(rng.integers(0, genome_len - L))
        r = list(genome[p:p + L])
        for j in range(L):
            if rng.random() < err:
                r[j] = "ACGT"[("ACGT".index(r[j]) + int(rng.integers(1, 4))) % 4]
        s = "".join(r)
        reads.append(s if rng.random() < 0.5 else revcomp(s))
    return reads


@pytest.mark.parametrize("k,m,x,P", [(31, 21, 1, 64), (31, 21, 5, 1024), (21, 11, 2, 7), (31, 19, 0, 3),
                                     (15, 9, 3, 1024), (9, 5, 1, 16),
                                     (33, 17, 0, 1024), (41, 21, 1, 64), (45, 27, 3, 7), (63, 31, 2, 1024)])
def test_count_inputs_matches_oracle(tmp_path, k, m, x, P):
    print(f"count_inputs k={k} m={m} x={x} P={P}", flush=True)
    reads = _reads(k + x)
    fa = tmp_path / "r.fa"
    write_fasta(fa, reads)
    got = M.count_inputs([fa], M.CounterConfig(k, m, x, P))
    codes, off = O.parse_fastx(str(fa), k)
    want = O.count_bruteforce(codes, off, k, x)
    assert same_counted(got, want)
    tp, _, _ = O.count_two_pass(codes, off, k, m, x, P)
    assert same_counted(got, tp)


def test_golden_count_cases(golden, tmp_path):
    for i, case in enumerate(golden["count_inputs"]):
        p = tmp_path / f"g{i}.fa"
        p.write_text(case["fasta"])
        got = as_pairs(M.count_inputs([p], M.CounterConfig(case["k"], case["m"], case["x"])), case["k"])
        if "exact" in case:
            assert got == [tuple(x) for x in case["exact"]], case["src"]
        for x in case.get("contains", []):
            assert tuple(x) in got, case["src"]
    case = golden["kmer_fasta"][0]
    p = tmp_path / "kf.fa"
    p.write_text(case["fasta"])
    n = M.count_inputs_to_kmer_fasta_path([p], M.CounterConfig(case["k"], case["m"], case["x"]), tmp_path / "o.fa")
    assert n == case["selected"]
    assert case["contains"] in (tmp_path / "o.fa").read_text()


def test_edge_inputs(tmp_path):
    cfg = M.CounterConfig(31, 21, 0, 64)
    empty = tmp_path / "e.fa"
    empty.write_text("")
    assert len(M.count_inputs([empty], cfg)) == 0
    short = tmp_path / "s.fa"
    short.write_text(">a\nACGT\n>b\nNNNNNNNN\n>c\nACGTNACGTACGTACGT\n")
    assert len(M.count_inputs([short], cfg)) == 0
    # long fragment crossing many tiles, homopolymer run, lowercase, N splits
    rng = np.random.default_rng(9)
    long = "".join(rng.choice(list("ACGT"), size=100000))
    text = ">l\n" + long + "\n>h\n" + "A" * 5000 + "\n>lc\n" + long[:3000].lower() + "N" + long[5000:9000] + "\n"
    f = tmp_path / "l.fa"
    f.write_text(text)
    for x in (0, 1):
        got = M.count_inputs([f], M.CounterConfig(31, 21, x, 64))
        codes, off = O.parse_fastx(str(f), 31)
        assert same_counted(got, O.count_bruteforce(codes, off, 31, x))
    with pytest.raises(M.MczaError, match="odd"):
        M.count_inputs([f], M.CounterConfig(30, 21, 1))


def test_kff_bytes_match_oracle(tmp_path):
    reads = _reads(77, n=3000)
    fa = tmp_path / "r.fa"
    write_fasta(fa, reads)
    codes, off = O.parse_fastx(str(fa), 31)
    for x, P in ((1, 64), (0, 1024), (3, 5)):
        out = tmp_path / f"o{x}.kff"
        M.count_inputs_to_kff_path([fa], M.CounterConfig(31, 21, x, P), out)
        want, _ = O.kff_u8_kmc(codes, off, 31, 21, x, P)
        assert out.read_bytes() == want, (x, P)


def test_fasta_bytes_match_oracle(tmp_path):
    reads = _reads(78, n=3000)
    fa = tmp_path / "r.fa"
    write_fasta(fa, reads)
    codes, off = O.parse_fastx(str(fa), 31)
    out = tmp_path / "o.fa"
    M.count_inputs_to_fasta_path([fa], M.CounterConfig(31, 21, 1, 64), out)
    kms = O.count_bruteforce(codes, off, 31, 1)
    O.write_simplitigs(str(tmp_path / "w.fa"), kms, 31)
    assert out.read_bytes() == (tmp_path / "w.fa").read_bytes()


@pytest.mark.parametrize("k,m", [(31, 21), (45, 27)])
def test_simplitigs_with_device_links_match_host_writer(tmp_path, k, m):
    # count -> FASTA in one call finds the writer's forced links on the device;
    # writing the downloaded k-mers finds them on the host: same bytes. The
    # genome has 5 kbp repeats, so the k-mer graph branches.
    spec = M.SynthSpec(genome_len=400_000, genome_seed=9, read_seed=109, n_reads=60_000, read_len=150,
                       sub_rate=0.002, repeat_fraction=0.05, repeat_len=5000, repeat_copies=4)
    r = M.SynthReads(spec)
    fa = tmp_path / "r.fa"
    r.write_fasta(fa)
    cfg = M.CounterConfig(k, m, 1, 64)
    dev = tmp_path / "dev.fa"
    M.count_inputs_to_fasta_path([fa], cfg, dev)
    kms = M.count_inputs([fa], cfg)
    assert kms.size > 1 << 17  # above the host writer's threading threshold
    host = tmp_path / "host.fa"
    M.write_fasta_path(host, kms, k)
    assert dev.read_bytes() == host.read_bytes()


@pytest.mark.parametrize("k,m", [(31, 21), (45, 27)])
def test_superkmer_partition_files_round_trip(tmp_path, k, m):
    # reads -> device scan -> partition_NNNN.skm files -> device count of the
    # files == count of the reads; each k-mer lives in one partition only
    spec = M.SynthSpec(genome_len=300_000, genome_seed=4, read_seed=104, n_reads=40_000, read_len=150,
                       sub_rate=0.005, repeat_fraction=0.05, repeat_len=2000, repeat_copies=3)
    r = M.SynthReads(spec)
    fa = tmp_path / "r.fa"
    r.write_fasta(fa)
    cfg = M.CounterConfig(k, m, 0, 16)
    out = tmp_path / "parts"
    out.mkdir()
    M.write_superkmer_partitions([fa], cfg, out)
    files = sorted(out.glob("partition_*.skm"))
    assert [f.name for f in files] == [f"partition_{u:04d}.skm" for u in range(16)]
    for x in (0, 2):
        c = M.CounterConfig(k, m, x, 16)
        assert same_counted(M.count_superkmer_partitions(files, c), M.count_inputs([fa], c)), x
    seen = set()
    for f in files:
        kms = M.count_superkmer_partitions([f], cfg)
        mine = set(zip(kms["hi"].tolist(), kms["lo"].tolist()))
        assert not (mine & seen)
        seen |= mine


def test_output_order_with_long_prefix_runs(tmp_path):
    # The encoded-order sort finishes runs of k-mers that share their top 24
    # bits in place; 300 k-mers that start with twelve A's form one run too
    # long for that, and the full sort must take over.
    rng = np.random.default_rng(31)
    reads = ["A" * 12 + "".join(rng.choice(list("ACG"), size=19)) for _ in range(300)]
    reads += ["".join(rng.choice(list("ACGT"), size=150)) for _ in range(2000)]
    fa = tmp_path / "r.fa"
    write_fasta(fa, reads)
    codes, off = O.parse_fastx(str(fa), 31)
    want = O.count_bruteforce(codes, off, 31, 0)
    got = M.count_inputs([fa], M.CounterConfig(31, 21, 0, 64))
    assert same_counted(got, want)
    assert int((want["lo"] >> np.uint64(38) == 0).sum()) >= 300


def test_datasets_match_oracle(tmp_path, golden):
    for case in golden["datasets"]:
        paths = []
        for i, s in enumerate(case["datasets"]):
            p = tmp_path / f"d{i}.fa"
            p.write_text(f">d{i}\n{s}\n")
            paths.append(p)
        got = as_pairs(M.count_datasets(paths, M.CounterConfig(case["k"], case["m"], case["x"])), case["k"])
        for x in case.get("contains", []):
            assert tuple(x) in got, case["src"]
        for a in case.get("absent", []):
            assert a not in [g[0] for g in got], case["src"]
        if "exact" in case:
            assert got == [tuple(x) for x in case["exact"]], case["src"]
    rng = np.random.default_rng(5)
    base = "".join(rng.choice(list("ACGT"), size=5000))
    paths, seqs, ds = [], [], []
    for d in range(7):
        g = list(base)
        for _ in range(20):
            q = int(rng.integers(0, len(g)))
            g[q] = "ACGT"[("ACGT".index(g[q]) + 1) % 4]
        g = "".join(g)
        rs = [g[p:p + 150] for p in rng.integers(0, len(g) - 150, size=200)]
        p = tmp_path / f"r{d}.fa"
        write_fasta(p, rs)
        paths.append(p)
        seqs += rs
        ds += [d] * len(rs)
    codes, off, dsa = O.frags_of(seqs, datasets=ds)
    for x in (0, 2, 5):
        got = M.count_datasets(paths, M.CounterConfig(31, 21, x, 64))
        want = O.count_datasets_bruteforce(codes, off, dsa, 7, 31, x)
        assert same_counted(got, want), x
    # the dataset simplitig path (keys only to the writer); x >= datasets writes nothing
    for x in (2, 7):
        out = tmp_path / f"ds{x}.fa"
        M.count_datasets_to_fasta_path(paths, M.CounterConfig(31, 21, x, 64), out)
        want = O.count_datasets_bruteforce(codes, off, dsa, 7, 31, x)
        O.write_simplitigs(str(tmp_path / "w.fa"), want, 31)
        assert out.read_bytes() == (tmp_path / "w.fa").read_bytes(), x


def test_batch_plain_equals_batched_and_spill(monkeypatch):
    spec = M.SynthSpec(genome_len=200_000, genome_seed=1, read_seed=101, n_reads=40_000, read_len=150,
                       sub_rate=0.005)
    r = M.SynthReads(spec)
    b = M.Batch.from_packed(r.words, r.frag_off)
    codes = r.codes()
    cfg = M.CounterConfig(31, 21, 1, 64)
    want = O.count_bruteforce(codes, r.frag_off, 31, 1)
    s1 = b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_BATCHED)
    a = b.download()
    assert same_counted(a, want)
    s2 = b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_PLAIN)
    assert same_counted(b.download(), want)
    s3 = b.count(cfg, M.MODE_SINGLE_PASS, M.LAUNCH_BATCHED)
    assert same_counted(b.download(), want)
    assert s1.n_out == s2.n_out == s3.n_out == len(want)
    # repeated graph launches give identical results
    for _ in range(3):
        b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_BATCHED)
        assert same_counted(b.download(), want)
    # few, large partitions: most k-mers spill from LDS to the HBM table
    monkeypatch.setenv("MCZI_PART_TARGET", str(10 ** 9))
    b2 = M.Batch.from_packed(r.words, r.frag_off)
    st = b2.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_BATCHED)
    assert st.n_overflow > 0
    assert same_counted(b2.download(), want)


@pytest.mark.parametrize("k,m", [(31, 21), (45, 27)])
def test_staging_grows_past_its_estimate(monkeypatch, k, m):
    """The scan reserves record slots per tile from one counter; when the
    records exceed the staging buffer's first capacity, the buffer is sized to
    the exact total and the batch is scanned again, with the same result."""
    print(f"staging regrow k={k} m={m}", flush=True)
    spec = M.SynthSpec(genome_len=100_000, genome_seed=3, read_seed=103, n_reads=8_000, read_len=150,
                       sub_rate=0.01)
    r = M.SynthReads(spec)
    codes = r.codes()
    cfg = M.CounterConfig(k, m, 1, 64)
    want = O.count_bruteforce(codes, r.frag_off, k, 1)
    b = M.Batch.from_packed(r.words, r.frag_off)
    st = b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_BATCHED)
    assert same_counted(b.download(), want)
    monkeypatch.setenv("MCZI_STAGE_CAPACITY", "1000")
    b2 = M.Batch.from_packed(r.words, r.frag_off)
    st2 = b2.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_BATCHED)
    assert int(st.n_records) > 1000
    assert (int(st2.n_records), int(st2.n_pieces)) == (int(st.n_records), int(st.n_pieces))
    assert same_counted(b2.download(), want)
    monkeypatch.delenv("MCZI_STAGE_CAPACITY")
    # the grown buffer is kept: later counts of the batch scan once
    for launch in (M.LAUNCH_BATCHED, M.LAUNCH_PLAIN):
        b2.count(cfg, M.MODE_TWO_PASS, launch)
        assert same_counted(b2.download(), want)


def test_many_small_inputs_plain_equals_batched():
    """Inputs marked in one packed batch: the plain mode scans one input
    per launch (here 600 inputs of 20 reads), the batched mode all of them in
    one; both equal brute force."""
    r = M.SynthReads(M.SynthSpec(genome_len=50_000, genome_seed=11, read_seed=111, n_reads=12_000, read_len=150,
                                 sub_rate=0.01))
    marks = np.arange(0, 12_001, 20, dtype=np.uint64)
    want = O.count_bruteforce(r.codes(), r.frag_off, 31, 1)
    cfg = M.CounterConfig(31, 21, 1, 64)
    # the same reads marked with empty inputs between and around them
    gaps = np.sort(np.concatenate([marks, marks[::7], [0, 0, 12_000]])).astype(np.uint64)
    for m in (marks, gaps):
        b = M.Batch.from_packed(r.words, r.frag_off, input_frag=m)
        for launch in (M.LAUNCH_PLAIN, M.LAUNCH_BATCHED, M.LAUNCH_PLAIN):
            b.count(cfg, M.MODE_TWO_PASS, launch)
            assert same_counted(b.download(), want), (launch, m.size)


def _concat(a, b):
    """Two packed read sets in one segmented buffer (word-aligned; the gap
    between them is a short fragment that has no k-mer)."""
    na = int(a.frag_off[-1])
    wa = (na + 15) // 16
    pad = wa * 16
    words = np.concatenate([a.words[:wa], b.words])
    off = np.concatenate([a.frag_off, [pad] if pad != na else [], b.frag_off[1:] + pad]).astype(np.uint64)
    return words, off


def test_long_and_short_reads_one_launch():
    long = M.SynthReads(M.SynthSpec(genome_len=300_000, genome_seed=3, read_seed=103, long_bases=3_000_000,
                                    long_median=8000, long_sigma=0.8, long_min=500, long_max=100_000,
                                    sub_rate=0.04, ins_rate=0.02, del_rate=0.02))
    short = M.SynthReads(M.SynthSpec(genome_len=300_000, genome_seed=3, read_seed=203, n_reads=20_000,
                                     read_len=150, sub_rate=0.005))
    words, off = _concat(long, short)
    b = M.Batch.from_packed(words, off)
    gap = int(off[len(long.frag_off)]) - int(long.frag_off[-1]) if len(off) > len(long.frag_off) + len(
        short.frag_off) - 1 else 0
    codes = np.concatenate([long.codes(), np.zeros(gap, dtype=np.uint8), short.codes()])
    cfg = M.CounterConfig(31, 21, 3, 1024)
    want = O.count_bruteforce(codes, off, 31, 3)
    t = b.autotune(cfg)
    assert t > 0
    b.count(cfg)
    assert same_counted(b.download(), want)
    b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_PLAIN)
    assert same_counted(b.download(), want)


def test_fof_32_datasets_batch():
    r = M.SynthReads(M.SynthSpec(genome_len=60_000, genome_seed=4, read_seed=104, n_reads=1200, read_len=150,
                                 sub_rate=0.005, n_datasets=32, snp_rate=0.001))
    b = M.Batch.from_packed(r.words, r.frag_off, r.dataset)
    assert b.info()[2] == 32
    codes = r.codes()
    for x in (5, 20, 31):
        st = b.count(M.CounterConfig(31, 21, x, 1024), M.MODE_DATASETS)
        got = b.download()
        want = O.count_datasets_bruteforce(codes, r.frag_off, r.dataset, 32, 31, x)
        assert same_counted(got, want), x
        assert st.n_out == len(want)


@pytest.mark.parametrize("n_datasets", [33, 100])
def test_fof_beyond_32_datasets(n_datasets):
    """count_dataset_presence_counts_u64 (src/lib.rs:637-703) takes any number
    of datasets: presence counts above 32 datasets equal brute force, in one
    batch and through the file path."""
    print(f"fof datasets={n_datasets}", flush=True)
    r = M.SynthReads(M.SynthSpec(genome_len=40_000, genome_seed=4, read_seed=404, n_reads=60 * n_datasets,
                                 read_len=150, sub_rate=0.005, n_datasets=n_datasets, snp_rate=0.002))
    b = M.Batch.from_packed(r.words, r.frag_off, r.dataset)
    assert b.info()[2] == n_datasets
    codes = r.codes()
    for x in (0, 5, 31, 32, n_datasets - 2):
        st = b.count(M.CounterConfig(31, 21, x, 64), M.MODE_DATASETS)
        want = O.count_datasets_bruteforce(codes, r.frag_off, r.dataset, n_datasets, 31, x)
        assert same_counted(b.download(), want), x
        assert st.n_out == len(want)


@pytest.mark.parametrize("n_datasets", [8, 40])
def test_fof_per_file_plain_equals_batched(n_datasets):
    """The plain launch over one input per dataset (the per-file path) counts
    each input on its own and merges the per-file presence on the device:
    equal to the batched count and to brute force, in both fof modes; inputs
    that mix datasets keep the single-count plain launch, also equal."""
    print(f"fof per-file datasets={n_datasets}", flush=True)
    r = M.SynthReads(M.SynthSpec(genome_len=40_000, genome_seed=4, read_seed=505, n_reads=60, read_len=150,
                                 sub_rate=0.005, n_datasets=n_datasets, snp_rate=0.002))
    n = len(r.frag_off) - 1
    assert n == 60 * n_datasets and np.array_equal(r.dataset[::60], np.arange(n_datasets))
    marks = np.arange(0, n + 1, 60, dtype=np.uint64)
    b = M.Batch.from_packed(r.words, r.frag_off, r.dataset, input_frag=marks)
    codes = r.codes()
    for x in (0, 3, n_datasets - 2):
        want = O.count_datasets_bruteforce(codes, r.frag_off, r.dataset, n_datasets, 31, x)
        cfg = M.CounterConfig(31, 21, x, 64)
        for mode in (M.MODE_DATASETS, M.MODE_DATASETS_TWO_PASS):
            for launch in (M.LAUNCH_PLAIN, M.LAUNCH_BATCHED):
                st = b.count(cfg, mode, launch)
                assert same_counted(b.download(), want), (x, mode, launch)
                assert st.n_out == len(want)
    mixed = np.array([0, 90, n], dtype=np.uint64)
    b2 = M.Batch.from_packed(r.words, r.frag_off, r.dataset, input_frag=mixed)
    want = O.count_datasets_bruteforce(codes, r.frag_off, r.dataset, n_datasets, 31, 3)
    b2.count(M.CounterConfig(31, 21, 3, 64), M.MODE_DATASETS, M.LAUNCH_PLAIN)
    assert same_counted(b2.download(), want)


def test_counts_per_file_plain_thresholds():
    """The per-file plain path in the counting modes: three inputs (two
    genomes, the second sampled twice), thresholds 0 and 2, single and
    two-pass: the merged per-file counts equal brute force over all inputs."""
    a = M.SynthReads(M.SynthSpec(genome_len=30_000, genome_seed=21, read_seed=121, n_reads=2000, read_len=150,
                                 sub_rate=0.01))
    c = M.SynthReads(M.SynthSpec(genome_len=20_000, genome_seed=22, read_seed=122, n_reads=3000, read_len=150,
                                 sub_rate=0.005))
    words, off = _concat(a, c)
    n = len(off) - 1
    n_a = n - (len(c.frag_off) - 1)
    marks = np.array([0, n_a, n_a + 1500, n], dtype=np.uint64)
    b = M.Batch.from_packed(words, off, input_frag=marks)
    codes = np.concatenate([a.codes(), np.zeros(int(off[n_a]) - int(a.frag_off[-1]), dtype=np.uint8), c.codes()])
    for x in (0, 2):
        want = O.count_bruteforce(codes, off, 31, x)
        cfg = M.CounterConfig(31, 21, x, 64)
        for mode in (M.MODE_SINGLE_PASS, M.MODE_TWO_PASS):
            st = b.count(cfg, mode, M.LAUNCH_PLAIN)
            assert same_counted(b.download(), want), (x, mode)
            assert st.n_out == len(want)


def test_fof_file_path_40_datasets(tmp_path):
    rng = np.random.default_rng(40)
    base = "".join(rng.choice(list("ACGT"), size=3000))
    paths, seqs, ds = [], [], []
    for d in range(40):
        rs = [base[p:p + 120] for p in rng.integers(0, 2880, size=30)]
        p = tmp_path / f"r{d}.fa"
        write_fasta(p, rs)
        paths.append(p)
        seqs += rs
        ds += [d] * len(rs)
    codes, off, dsa = O.frags_of(seqs, datasets=ds)
    for x in (3, 33):
        got = M.count_datasets(paths, M.CounterConfig(31, 21, x, 64))
        assert same_counted(got, O.count_datasets_bruteforce(codes, off, dsa, 40, 31, x)), x


def test_c1_full_config_kff_bytes(tmp_path):
    """MC abundance mode, plain method, small: 1 Mbp genome (seed 1), 200k
    150 bp reads, 0.5% substitutions, k=31, m=21, x=1, 64 partitions, KFF."""
    r = M.SynthReads(M.SynthSpec(genome_len=1_000_000, genome_seed=1, read_seed=101, n_reads=200_000,
                                 read_len=150, sub_rate=0.005))
    fa = tmp_path / "c1.fa"
    r.write_fasta(fa)
    out = tmp_path / "c1.kff"
    M.count_inputs_to_kff_path([fa], M.CounterConfig(31, 21, 1, 64), out)
    want, nrec = O.kff_u8_kmc(r.codes(), r.frag_off, 31, 21, 1, 64)
    data = out.read_bytes()
    assert data == want
    assert nrec == 1053631
    # read back by the independent KFF reader: the closed form
    # {K : c(K) > 1} with count min(c(K), 2), each k-mer once
    f = KffFile(data, decode_records=False)
    assert f.variables["k"] == 31 and f.variables["data_size"] == 1 and f.canonical == 1
    keys, vals = kmers_u64(f, data)
    bf = O.count_bruteforce(r.codes(), r.frag_off, 31, 1)
    order = np.argsort(keys)
    assert np.array_equal(keys[order], bf["lo"].astype(np.uint64))
    assert np.array_equal(vals[order], np.minimum(bf["count"].astype(np.uint64), 2))


def _spot_check(words, off, k, x, got, dataset=None, n_emitted=3000, n_windows=3000, seed=0):
    """Full-size parity by sampling: the exact count (fof: number of datasets)
    of random emitted k-mers, and of the k-mers at random read windows, from
    the oracle's spot counter over the same packed reads. A window k-mer is
    emitted iff its value is > x, with that value."""
    rng = np.random.default_rng(seed)
    lo = got["lo"].astype(np.uint64)
    cnt = got["count"].astype(np.uint64)
    assert np.all(got["hi"] == 0) and np.all(np.diff(lo) > 0)
    idx = rng.integers(0, lo.size, size=min(n_emitted, lo.size))
    lens = np.diff(off.astype(np.int64))
    ok = np.nonzero(lens >= k)[0]
    f = rng.choice(ok, size=n_windows)
    pos = off[f].astype(np.uint64) + (rng.random(n_windows) * (lens[f] - k + 1)).astype(np.uint64)
    wk = O.kmers_at(words, pos, k)
    c, m = O.spot_counts(words, off, k, np.concatenate([lo[idx], wk]), dataset, 16)
    val = c if dataset is None else np.array([bin(int(v)).count("1") for v in m], dtype=np.uint64)
    assert np.array_equal(val[: idx.size], cnt[idx])
    vw = val[idx.size:]
    at = np.minimum(np.searchsorted(lo, wk), lo.size - 1)
    found = lo[at] == wk
    assert np.array_equal(found, vw > x)
    assert np.array_equal(cnt[at][found], vw[found])


def test_c2_full_size_properties():
    """Short reads at scale (100 Mbp + repeats, 20M reads): batched and plain
    launches agree exactly; output sorted, unique, all counts > x; the total
    count of emitted k-mers cannot exceed the windows."""
    r = M.SynthReads(M.SynthSpec(genome_len=100_000_000, genome_seed=2, read_seed=102, n_reads=20_000_000,
                                 read_len=150, sub_rate=0.01, repeat_fraction=0.02, repeat_len=5000,
                                 repeat_copies=10))
    b = M.Batch.from_packed(r.words, r.frag_off)
    cfg = M.CounterConfig(31, 21, 5, 1024)
    st = b.count(cfg)
    a = b.download()
    b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_PLAIN)
    c = b.download()
    assert same_counted(a, c)
    assert len(a) == st.n_out > 90_000_000
    assert np.all(a["hi"] == 0) and np.all(np.diff(a["lo"].astype(np.uint64)) > 0)
    assert np.all(a["count"] > 5)
    assert int(a["count"].sum()) <= st.n_windows
    # single pass (KMC-style) gives the same set and counts
    b.count(cfg, M.MODE_SINGLE_PASS)
    assert same_counted(b.download(), a)
    _spot_check(r.words, r.frag_off, 31, 5, a)


def test_c3_full_size_long_and_short_reads():
    """Long reads (50 Mbp genome, seed 3; lognormal lengths, median 8 kbp,
    sigma 0.8, clip [500, 1e5], 20x; 4/2/2% substitutions/insertions/
    deletions) plus 150 bp reads at 10x (0.5% substitutions) in one launch,
    k=31, m=21, x=3: batched equals plain, and sampled exact counts."""
    long = M.SynthReads(M.SynthSpec(genome_len=50_000_000, genome_seed=3, read_seed=103, long_bases=1_000_000_000,
                                    long_median=8000, long_sigma=0.8, long_min=500, long_max=100_000,
                                    sub_rate=0.04, ins_rate=0.02, del_rate=0.02))
    short = M.SynthReads(M.SynthSpec(genome_len=50_000_000, genome_seed=3, read_seed=203, n_reads=3_333_334,
                                     read_len=150, sub_rate=0.005))
    words, off = _concat(long, short)
    del long, short
    b = M.Batch.from_packed(words, off)
    cfg = M.CounterConfig(31, 21, 3, 1024)
    st = b.count(cfg)
    got = b.download()
    assert st.n_bases > 1_400_000_000 and len(got) == st.n_out
    b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_PLAIN)
    assert same_counted(b.download(), got)
    _spot_check(words, off, 31, 3, got, seed=3)


def test_c4_full_size_fof_32_datasets():
    """32 datasets: variants of a 20 Mbp genome (seed 4) with 0.1% SNPs, each
    sequenced as 150 bp reads at 15x with 0.5% errors (9.6 Gbp in one
    segmented buffer), fof k=31, m=21, x=5: sampled exact dataset counts."""
    r = M.SynthReads(M.SynthSpec(genome_len=20_000_000, genome_seed=4, read_seed=104, n_reads=2_000_000,
                                 read_len=150, sub_rate=0.005, n_datasets=32, snp_rate=0.001))
    b = M.Batch.from_packed(r.words, r.frag_off, r.dataset)
    assert b.info()[2] == 32
    st = b.count(M.CounterConfig(31, 21, 5, 1024), M.MODE_DATASETS)
    got = b.download()
    assert st.n_bases == 9_600_000_000 and len(got) == st.n_out
    assert np.all(got["count"] > 5) and np.all(got["count"] <= 32)
    _spot_check(r.words, r.frag_off, 31, 5, got, dataset=r.dataset, seed=4)


@pytest.mark.parametrize("k,m", [(31, 21), (21, 11), (15, 9)])
def test_pieces_and_records_match_oracle_runs(k, m):
    """The scan's super-k-mer split: pieces = the oracle's minimizer runs of
    each fragment, also cut where a scan tile starts inside a run; each piece
    becomes ceil(windows / (57 - k)) records (src/lib.rs:3483-3560)."""
    print(f"pieces k={k} m={m}", flush=True)
    long = M.SynthReads(M.SynthSpec(genome_len=50_000, genome_seed=7, read_seed=17, long_bases=60_000,
                                    long_median=6000, long_sigma=0.5, long_min=500, long_max=20_000,
                                    sub_rate=0.01))
    short = M.SynthReads(M.SynthSpec(genome_len=50_000, genome_seed=7, read_seed=27, n_reads=600,
                                     read_len=150, sub_rate=0.01))
    words, off = _concat(long, short)
    b = M.Batch.from_packed(words, off)
    st = b.count(M.CounterConfig(k, m, 0, 64))
    tile = int(st.tile_windows)
    assert 0 < tile
    n_bases = int(off[-1])
    codes = np.zeros(n_bases, dtype=np.uint8)
    codes[: int(long.frag_off[-1])] = long.codes()
    codes[n_bases - int(short.frag_off[-1]):] = short.codes()
    wrec = 57 - k
    pieces = records = 0
    for f in range(len(off) - 1):
        s0, s1 = int(off[f]), int(off[f + 1])
        nwin = s1 - s0 - k + 1
        if nwin <= 0:
            continue
        _, starts = O.minimizer_runs(np.ascontiguousarray(codes[s0:s1]), k, m)
        cuts = set(int(x) for x in starts)
        first_tile = (s0 // tile + 1) * tile
        cuts.update(t - s0 for t in range(first_tile, s0 + nwin, tile))
        bounds = sorted(cuts) + [nwin]
        pieces += len(bounds) - 1
        records += sum((bounds[i + 1] - bounds[i] + wrec - 1) // wrec for i in range(len(bounds) - 1))
    assert (int(st.n_pieces), int(st.n_records)) == (pieces, records)



def test_wide_k_writers_and_spill(tmp_path, monkeypatch):
    """33 <= k <= 63 (u128 k-mers, src/lib.rs:110-137): FASTA and KFF (8-byte
    counts) bytes equal the oracle writers' over the brute-force counts; plain
    and batched launches agree; with few, large partitions most k-mers spill
    to the HBM table and the result is unchanged."""
    reads = _reads(91, n=3000)
    fa = tmp_path / "r.fa"
    write_fasta(fa, reads)
    for k, m, x in ((41, 21, 1), (63, 25, 0)):
        codes, off = O.parse_fastx(str(fa), k)
        want = O.count_bruteforce(codes, off, k, x)
        assert np.any(want["hi"] != 0)
        out = tmp_path / f"o{k}.fa"
        M.count_inputs_to_fasta_path([fa], M.CounterConfig(k, m, x, 64), out)
        O.write_simplitigs(str(tmp_path / f"w{k}.fa"), want, k)
        assert out.read_bytes() == (tmp_path / f"w{k}.fa").read_bytes()
        kff = tmp_path / f"o{k}.kff"
        M.count_inputs_to_kff_path([fa], M.CounterConfig(k, m, x, 64), kff)
        O.write_kff(str(tmp_path / f"w{k}.kff"), want, k, m, x, count_size=8, ordered=False)
        assert kff.read_bytes() == (tmp_path / f"w{k}.kff").read_bytes()
    r = M.SynthReads(M.SynthSpec(genome_len=200_000, genome_seed=9, read_seed=109, n_reads=40_000,
                                 read_len=150, sub_rate=0.01))
    want = O.count_bruteforce(r.codes(), r.frag_off, 47, 2)
    b = M.Batch.from_packed(r.words, r.frag_off)
    cfg = M.CounterConfig(47, 23, 2, 1024)
    b.count(cfg)
    assert same_counted(b.download(), want)
    b.count(cfg, M.MODE_TWO_PASS, M.LAUNCH_PLAIN)
    assert same_counted(b.download(), want)
    monkeypatch.setenv("MCZI_PART_TARGET", str(10 ** 9))
    b2 = M.Batch.from_packed(r.words, r.frag_off)
    st = b2.count(cfg)
    assert st.n_overflow > 0
    assert same_counted(b2.download(), want)
    with pytest.raises(M.MczaError, match="k <= 32"):
        b2.count(cfg, M.MODE_SINGLE_PASS, M.LAUNCH_BATCHED, M.ORDER_KFF_U8)
