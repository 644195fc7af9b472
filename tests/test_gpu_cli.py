"""The MC and ZI command lines (mczi_amd/bin) on the device: output bytes
equal the oracle writers' over the oracle's counts and set difference, the
stderr log carries the reference's phase lines, and invalid k is refused
(src/main.rs, src/zi.rs command lines). MI355X only."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
from helpers import revcomp, write_fasta

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MC = os.path.join(ROOT, "mczi_amd", "bin", "MC")
ZI = os.path.join(ROOT, "mczi_amd", "bin", "ZI")


def _run(args):
    print("run:", " ".join(str(a) for a in args), flush=True)
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)


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


def test_mc_command_line_fasta_and_kff(tmp_path):
    rng = np.random.default_rng(21)
    fa = tmp_path / "r.fa"
    write_fasta(fa, _reads(rng, _genome(rng, 15000), 2500))
    codes, off = O.parse_fastx(str(fa), 31)
    out = tmp_path / "o.fa"
    r = _run([MC, "-i", fa, "-k", 31, "-m", 21, "-x", 1, "-t", 3, "-o", out])
    assert r.returncode == 0, r.stderr
    assert "MC phase" in r.stderr
    O.write_simplitigs(str(tmp_path / "w.fa"), O.count_bruteforce(codes, off, 31, 1), 31)
    assert out.read_bytes() == (tmp_path / "w.fa").read_bytes()
    kff = tmp_path / "o.kff"
    r = _run([MC, "-i", fa, "-k", 31, "-m", 21, "-x", 2, "--format", "kff", "--partition-count", 64, "-o", kff])
    assert r.returncode == 0, r.stderr
    want, _ = O.kff_u8_kmc(codes, off, 31, 21, 2, 64)
    assert kff.read_bytes() == want
    r = _run([MC, "-i", fa, "-k", 30, "-m", 21, "-o", tmp_path / "e.fa"])
    assert r.returncode != 0 and "odd" in r.stderr


def test_zi_command_line_set_difference(tmp_path):
    rng = np.random.default_rng(22)
    genome, other = _genome(rng, 20000), _genome(rng, 20000)
    write_fasta(tmp_path / "idx.fa", [genome])
    q = tmp_path / "q.fa"
    write_fasta(q, _reads(rng, genome, 200, err=0.005) + _reads(rng, other, 200, err=0.005))
    out = tmp_path / "o.fa"
    r = _run([ZI, "--index", tmp_path / "idx.fa", "-i", q, "-k", 31, "-o", out])
    assert r.returncode == 0, r.stderr
    ic, io = O.parse_fastx(str(tmp_path / "idx.fa"), 1)
    qc, qo = O.parse_fastx(str(q), 31)
    O.write_simplitigs(str(tmp_path / "w.fa"), O.zi_absent(ic, io, qc, qo, 31), 31)
    assert out.read_bytes() == (tmp_path / "w.fa").read_bytes()


def test_mc_command_line_fofn(tmp_path):
    """--fofn: one dataset per listed file, -x is a dataset-presence threshold."""
    rng = np.random.default_rng(23)
    base = _genome(rng, 5000)
    paths, seqs, ds = [], [], []
    for d in range(5):
        g = list(base)
        for _ in range(20):
            q = int(rng.integers(0, len(g)))
            g[q] = "ACGT"[("ACGT".index(g[q]) + 1) % 4]
        g = "".join(g)
        rs = [g[p:p + 150] for p in rng.integers(0, len(g) - 150, size=150)]
        p = tmp_path / f"r{d}.fa"
        write_fasta(p, rs)
        paths.append(p)
        seqs += rs
        ds += [d] * len(rs)
    fofn = tmp_path / "list.txt"
    fofn.write_text("".join(f"{p}\n" for p in paths))
    codes, off, dsa = O.frags_of(seqs, datasets=ds)
    out = tmp_path / "o.fa"
    r = _run([MC, "--fofn", "-i", fofn, "-k", 31, "-m", 21, "-x", 2, "-o", out])
    assert r.returncode == 0, r.stderr
    O.write_simplitigs(str(tmp_path / "w.fa"), O.count_datasets_bruteforce(codes, off, dsa, 5, 31, 2), 31)
    assert out.read_bytes() == (tmp_path / "w.fa").read_bytes()
