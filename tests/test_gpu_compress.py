"""Compressed inputs through the device paths (MI355X only): the reference's
own cases (partitioned_counter_accepts_gzipped_fasta / _xz_fasta,
partitioned_counter_skips_invalid_compressed_fastx_inputs, src/lib.rs:5608-5730)
and larger gzip / xz / zstd inputs counted and subtracted exactly as their
plain text."""
import ctypes as C
import gzip
import lzma

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import random_seqs, same_counted, write_fasta
from test_compress_cpu import zstd_compress

pytestmark = pytest.mark.gpu

_COMP = {"gz": lambda b: gzip.compress(b, 1), "xz": lzma.compress, "zst": zstd_compress}


@pytest.mark.parametrize("ext", ["gz", "xz", "zst"])
def test_counter_accepts_compressed_fasta(tmp_path, ext):
    p = tmp_path / f"reads.fa.{ext}"
    p.write_bytes(_COMP[ext](b">read\nACGTACGT\n"))
    got = M.count_inputs([p], M.CounterConfig(k=3, minimizer=2, threshold=1))
    d = {int(r["lo"]): int(r["count"]) for r in got}
    assert d[O.encode_canonical("ACG")] == 4
    assert d[O.encode_canonical("GTA")] == 2


def test_counter_skips_invalid_compressed_input(tmp_path):
    bad = tmp_path / "not_fastx.txt.gz"
    bad.write_bytes(gzip.compress(b"not a FASTA or FASTQ file\n"))
    good = tmp_path / "reads.fa"
    good.write_text(">read\nACGTACGT\n")
    got = M.count_inputs([bad, good], M.CounterConfig(k=3, minimizer=2, threshold=1))
    d = {int(r["lo"]): int(r["count"]) for r in got}
    assert d[O.encode_canonical("ACG")] == 4 and d[O.encode_canonical("GTA")] == 2


@pytest.mark.parametrize("ext", ["gz", "xz", "zst"])
def test_compressed_inputs_count_like_plain(tmp_path, ext):
    rng = np.random.default_rng(31)
    seqs = random_seqs(rng, 3000, 50, 400)
    seqs += seqs[:1000]
    plain = tmp_path / "r.fa"
    write_fasta(plain, seqs)
    z = tmp_path / f"r.fa.{ext}"
    z.write_bytes(_COMP[ext](plain.read_bytes()))
    cfg = M.CounterConfig(k=31, minimizer=21, threshold=1)
    want = M.count_inputs([plain], cfg)
    got = M.count_inputs([z], cfg)
    assert want.size > 1000 and same_counted(got, want)
    # dataset presence over one compressed and one plain file
    got = M.count_datasets([z, plain], M.CounterConfig(k=31, minimizer=21, threshold=1))
    assert np.all(got["count"] == 2) and got.size == np.unique(
        M.count_inputs([plain], M.CounterConfig(k=31, minimizer=21, threshold=0))["lo"]).size


def test_zi_with_compressed_index_and_query(tmp_path):
    rng = np.random.default_rng(32)
    genome = "".join(rng.choice(list("ACGT"), size=20000))
    other = "".join(rng.choice(list("ACGT"), size=20000))
    reads = [genome[p:p + 150] for p in rng.integers(0, 19850, size=200)]
    reads += [other[p:p + 150] for p in rng.integers(0, 19850, size=200)]
    write_fasta(tmp_path / "idx.fa", [genome])
    write_fasta(tmp_path / "q.fa", reads)
    (tmp_path / "idx.fa.gz").write_bytes(gzip.compress((tmp_path / "idx.fa").read_bytes()))
    (tmp_path / "q.fa.zst").write_bytes(zstd_compress((tmp_path / "q.fa").read_bytes()))
    outs = []
    for idx, q in (("idx.fa", "q.fa"), ("idx.fa.gz", "q.fa.zst")):
        out = tmp_path / f"absent_{idx}.fa"
        M._check(M._lib.mczi_zi_subtract(str(tmp_path / idx).encode(), M._paths([tmp_path / q])[0], 1, 31, 0,
                                         str(out).encode(), 4, 8))
        outs.append(out.read_bytes())
    assert outs[0] == outs[1] and outs[0].count(b">") > 10
