"""libmczi.so without a GPU: exports, validation, host writers, ingest helpers.

Device entry points are not called here (no GPU in this container); the
writers and checks below are host code and are compared byte for byte with
the oracle's restatement of the reference writers.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import random_seqs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "mczi.h")).read()
    declared = sorted(set(re.findall(r"\b(mczi_[a-z0-9_]+)\s*\(", hdr)))
    assert declared == sorted(M.EXPORTED)
    for name in declared:
        assert hasattr(M._lib, name), name


def test_validation_matrix(golden):
    for case in golden["validation"]:
        cfg = M.CounterConfig(case["k"], case["m"], case["x"], 1024)
        if case["ok"]:
            M.validate_config(cfg, case.get("datasets", False))
        else:
            with pytest.raises(M.MczaError):
                M.validate_config(cfg, case.get("datasets", False))


def test_validation_messages():
    with pytest.raises(M.MczaError, match="threshold must be < 255"):
        M.validate_config(M.CounterConfig(31, 21, 255))
    with pytest.raises(M.MczaError, match="k <= 32"):
        M.validate_config(M.CounterConfig(33, 21, 1), datasets=True)


def _counted(seqs):
    a = np.zeros(len(seqs), dtype=M.COUNTED)
    for i, s in enumerate(seqs):
        a[i]["lo"] = O.encode_canonical(s)
        a[i]["count"] = i + 2
    a.sort(order=["hi", "lo"])
    return a


def test_simplitig_writer_matches_oracle(tmp_path, golden):
    for case in golden["simplitigs"] + golden["fasta_writer"]:
        a = _counted(case["kmers"])
        M.write_fasta_path(tmp_path / "m.fa", a, case["k"])
        O.write_simplitigs(str(tmp_path / "o.fa"), a, case["k"])
        assert (tmp_path / "m.fa").read_bytes() == (tmp_path / "o.fa").read_bytes()
    rng = np.random.default_rng(1)
    for k in (3, 5, 11, 31, 32, 33, 63):
        genome = "".join(rng.choice(list("ACGT"), size=3000))
        kms = sorted({O.encode_canonical(genome[i:i + k]) for i in range(0, 2000, 1)})
        a = np.zeros(len(kms), dtype=M.COUNTED)
        a["lo"] = [v & ((1 << 64) - 1) for v in kms]
        a["hi"] = [v >> 64 for v in kms]
        a["count"] = 1
        a.sort(order=["hi", "lo"])
        M.write_fasta_path(tmp_path / "m.fa", a, k)
        O.write_simplitigs(str(tmp_path / "o.fa"), a, k)
        assert (tmp_path / "m.fa").read_bytes() == (tmp_path / "o.fa").read_bytes(), k


def test_fasta_after_index(tmp_path, golden):
    case = golden["fasta_writer"][1]
    a = _counted(case["kmers"])
    last = M.write_fasta_after_index(tmp_path / "m.fa", a, case["k"], case["first_index"])
    assert last == case["last_index"]
    assert case["contains"][0] in (tmp_path / "m.fa").read_text()


def test_kff_writer_matches_oracle(tmp_path):
    a = _counted(["ACGTACGTAAC", "CCCCCCCCCCC", "GATTACAGATT"])
    cfg = M.CounterConfig(11, 7, 1)
    M.write_kff(tmp_path / "m.kff", a, cfg)
    O.write_kff(str(tmp_path / "o.kff"), a, 11, 7, 1, 8, True)
    b = (tmp_path / "m.kff").read_bytes()
    assert b == (tmp_path / "o.kff").read_bytes()
    assert b.startswith(b"KFF") and b.endswith(b"KFF")


def test_decode_kmer():
    assert M.decode_kmer(O.encode_canonical("ACG"), 0, 3) == "ACG"
    v = (1 << 120) | 5
    assert M.decode_kmer(v & ((1 << 64) - 1), v >> 64, 63) == O.decode(v & ((1 << 64) - 1), v >> 64, 63)


def test_expand_fofns(tmp_path):
    d = tmp_path / "sub"
    d.mkdir()
    (d / "a.fa").write_text(">a\nACGT\n")
    fofn = d / "list.txt"
    fofn.write_text("# comment\n\n a.fa \n/abs/b.fa\n")
    assert M.expand_fofns([fofn]) == [str(d / "a.fa"), "/abs/b.fa"]
    empty = tmp_path / "e.txt"
    empty.write_text("# nothing\n")
    with pytest.raises(M.MczaError, match="no inputs found"):
        M.expand_fofns([empty])


def test_synth_is_deterministic_and_parses_back(tmp_path):
    spec = M.SynthSpec(genome_len=20000, genome_seed=1, read_seed=101, n_reads=300, read_len=150,
                       sub_rate=0.01)
    a = M.SynthReads(spec)
    b = M.SynthReads(spec)
    assert np.array_equal(a.words, b.words) and np.array_equal(a.frag_off, b.frag_off)
    assert a.n_bases == 300 * 150
    a.write_fasta(tmp_path / "r.fa")
    codes, off = O.parse_fastx(str(tmp_path / "r.fa"), 1)
    assert np.array_equal(codes, a.codes()) and np.array_equal(off, a.frag_off)


def test_synth_genome_matches_reference_generator():
    # fill_random_genome (bench/simulate_long_reads.rs:166-180) with SplitMix64
    def splitmix(seed):
        s = seed
        while True:
            s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
            yield z ^ (z >> 31)

    gen = splitmix(1)
    bases = []
    while len(bases) < 200:
        v = next(gen)
        for _ in range(32):
            bases.append("ACGT"[v & 3])
            v >>= 2
    words = M.synth_genome(200, 1)
    code2base = "ACTG"
    got = "".join(code2base[(int(words[i // 16]) >> (2 * (i % 16))) & 3] for i in range(200))
    assert got == "".join(bases[:200])


def test_device_entry_points_fail_loudly_without_gpu(tmp_path):
    import subprocess
    import sys
    # no silent CPU fallback: counting without a HIP device reports an error
    code = ("import mczi_amd as M, sys\n"
            "open(sys.argv[1],'w').write('>r\\nACGTACGT\\n')\n"
            "try:\n M.count_inputs([sys.argv[1]], M.CounterConfig(3,2,1))\n"
            "except M.MczaError as e:\n print('ERR', e.code)\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code, str(tmp_path / "x.fa")], capture_output=True, text=True,
                       cwd=ROOT, env=env, timeout=120)
    assert "ERR 3" in r.stdout, r.stdout + r.stderr


def test_mczi_argument_checks(tmp_path):
    # MCZI's run() checks (src/mczi.rs:239-311, 885-889) come before any device work
    f = tmp_path / "a.fa"
    f.write_text(">a\nACGT\n")
    for (k, m, ds), msg in [((2, 1, False), r"k must be in \[3, 63\]"), ((65, 21, False), r"k must be in \[3, 63\]"),
                            ((30, 21, False), "k must be odd"), ((31, 31, False), "minimizer size must be > 0 and < k"),
                            ((31, 0, False), "minimizer size must be > 0 and < k"),
                            ((33, 21, True), "requires k <= 32")]:
        with pytest.raises(M.MczaError, match=msg):
            M.mczi_subtract([f], [f], M.CounterConfig(k, m, 1), index_datasets=ds)
    with pytest.raises(M.MczaError, match="at least one index input"):
        M.mczi_subtract([], [f], M.CounterConfig(31, 21, 1))
    with pytest.raises(M.MczaError, match="at least one query input"):
        M.mczi_subtract([f], [], M.CounterConfig(31, 21, 1))


def test_packed_input_marks_are_checked_before_device_work():
    import numpy as np
    words = np.zeros(4, dtype=np.uint32)
    off = np.array([0, 16, 32, 48], dtype=np.uint64)
    for marks, msg in [([1, 3], "start at fragment 0"), ([0, 2], "end at n_frag"),
                       ([0, 2, 1, 3], "must not decrease")]:
        with pytest.raises(M.MczaError, match=msg):
            M.Batch.from_packed(words, off, input_frag=np.array(marks, dtype=np.uint64))
    with pytest.raises(M.MczaError, match="fragment offsets must not decrease"):
        M.Batch.from_packed(words, np.array([0, 32, 16, 48], dtype=np.uint64), input_frag=np.array([0, 3]))


def test_noalloc_run_hashes_rejects_oversize_and_short_inputs():
    # argument checks run before any device work (no GPU needed)
    h = np.zeros(4, dtype=np.uint32)
    st = np.zeros(4, dtype=np.uint32)
    r = C.c_size_t(99)
    rc = M._lib.mczi_minimizer_run_hashes_noalloc(None, (1 << 24) + 1, 31, 21, h.ctypes.data, st.ctypes.data,
                                                  C.byref(r))
    assert rc == 1 and b"at most 16777216 bases" in M._lib.mczi_last_error()
    rc = M._lib.mczi_minimizer_run_hashes_noalloc(None, 10, 31, 21, h.ctypes.data, st.ctypes.data, C.byref(r))
    assert rc == 0 and r.value == 0
