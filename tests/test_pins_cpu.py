"""Pins for the oracle functions that no reference literal fixes.

- spread_hash_u32 (src/lib.rs:2984-2990) is the published "lowbias32" integer
  hash and spread_hash_u64 (src/lib.rs:2996-3002) is MurmurHash3's fmix64
  finalizer. Each is a bijection with a published inverse whose multipliers
  are the modular inverses of the forward ones; a round trip through that
  inverse fails for any changed constant or shift.
- The KFF bytes of the writers are read back by an independent KFF v1 reader
  (tests/kff_reader.py) and must give back exactly the k-mers and counts.
"""
import numpy as np

import mczi_amd as M
import pyoracle as O
from helpers import revcomp
from kff_reader import KffFile

M32 = (1 << 32) - 1
M64 = (1 << 64) - 1


def lowbias32_inverse(x: int) -> int:
    # published inverse of lowbias32 (hash-prospector, C. Wellons)
    x ^= x >> 16
    x = (x * 0x43021123) & M32
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & M32
    x ^= x >> 16
    return x


def fmix64_inverse(x: int) -> int:
    # inverse of MurmurHash3's fmix64: the shifts undone, multiplied by the
    # inverses of 0xc4ceb9fe1a85ec53 and 0xff51afd7ed558ccd
    x ^= x >> 33
    x = (x * 0x9CB4B2F8129337DB) & M64
    x ^= x >> 33
    x = (x * 0x4F74430C22A54005) & M64
    x ^= x >> 33
    return x


def test_inverse_multipliers_are_modular_inverses():
    assert (0x846CA68B * 0x43021123) & M32 == 1
    assert (0x7FEB352D * 0x1D69E2A5) & M32 == 1
    assert (0xC4CEB9FE1A85EC53 * 0x9CB4B2F8129337DB) & M64 == 1
    assert (0xFF51AFD7ED558CCD * 0x4F74430C22A54005) & M64 == 1


def test_spread_hash_u32_is_lowbias32():
    rng = np.random.default_rng(1)
    vals = [0, 1, 2, 0xFFFFFFFF, 0x80000000] + [int(v) for v in rng.integers(0, 1 << 32, size=5000, dtype=np.uint64)]
    for v in vals:
        h = O.spread_hash_u32(v)
        assert 0 <= h <= M32 and lowbias32_inverse(h) == v, hex(v)
    assert O.spread_hash_u32(0) == 0


def test_spread_hash_u64_is_fmix64():
    rng = np.random.default_rng(2)
    vals = [0, 1, M64, 1 << 63] + [int(v) for v in rng.integers(0, 1 << 63, size=5000, dtype=np.uint64)]
    for v in vals:
        h = O.spread_hash_u64(v)
        assert 0 <= h <= M64 and fmix64_inverse(h) == v, hex(v)
    assert O.spread_hash_u64(0) == 0


def _canon(s):
    r = revcomp(s)
    return min(s, r)


def test_kff_reader_reads_back_the_writers(tmp_path):
    rng = np.random.default_rng(3)
    k = 31
    seqs = sorted({_canon("".join(rng.choice(list("ACGT"), size=k))) for _ in range(3000)})
    arr = np.zeros(len(seqs), dtype=M.COUNTED)
    for i, s in enumerate(seqs):
        arr["lo"][i] = O.encode_canonical(s)
    arr.sort(order="lo")
    arr["count"] = rng.integers(2, 1 << 40, size=arr.size)
    cfg = M.CounterConfig(k, 21, 1, 64)
    M.write_kff(tmp_path / "m.kff", arr, cfg)
    O.write_kff(str(tmp_path / "o.kff"), arr, k, 21, 1, 8, True)
    for p in ("m.kff", "o.kff"):
        f = KffFile((tmp_path / p).read_bytes())
        assert f.unique == 1 and f.canonical == 1
        assert f.variables == {"k": k, "max": 1, "data_size": 8, "ordered": 1, "m": 21, "threshold": 1}
        assert f.free_block == b"producer=MC\nk=31\nm=21\nthreshold=1\nfilter=count>threshold\ncanonical=1\n"
        want = [(O.decode(int(r["lo"]), 0, k), int(r["count"])) for r in arr]
        assert f.records == want
        assert all(s == _canon(s) for s, _ in f.records)


def test_kff_reader_on_the_kmc_style_path():
    # the oracle's KMC-style KFF (u8 counts saturating at x + 1) read back
    # gives the closed form: {K : c(K) > x}, count min(c(K), x + 1)
    rng = np.random.default_rng(4)
    genome = "".join(rng.choice(list("ACGT"), size=3000))
    reads = [genome[p:p + 150] for p in rng.integers(0, 2850, size=400)]
    reads = [r if rng.random() < 0.5 else revcomp(r) for r in reads]
    codes, off, _ = O.frags_of(reads)
    for x in (0, 1, 3):
        data, nrec = O.kff_u8_kmc(codes, off, 31, 21, x, 16)
        f = KffFile(data)
        assert f.variables["data_size"] == 1 and f.variables["max"] == 1 and f.variables["ordered"] == 0
        assert len(f.records) == nrec
        bf = O.count_bruteforce(codes, off, 31, x)
        want = sorted((O.decode(int(r["lo"]), 0, 31), min(int(r["count"]), x + 1)) for r in bf)
        assert sorted(f.records) == want


def test_kff_vectorised_reader_agrees():
    from kff_reader import kmers_u64
    rng = np.random.default_rng(5)
    genome = "".join(rng.choice(list("ACGT"), size=2000))
    codes, off, _ = O.frags_of([genome[p:p + 100] for p in rng.integers(0, 1900, size=300)])
    data, _ = O.kff_u8_kmc(codes, off, 31, 21, 0, 8)
    f = KffFile(data)
    keys, vals = kmers_u64(KffFile(data, decode_records=False), data)
    assert [(O.decode(int(a), 0, 31), int(b)) for a, b in zip(keys, vals)] == f.records
