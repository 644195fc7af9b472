"""The CPU oracle against the reference's own known answers and closed forms.

Every check here runs on the CPU. The oracle is the checker for the HIP path,
so it is pinned first: reference test literals (tests/golden), brute force,
and the hash sweep that fixes the un-vendored ntHash variant.
"""
import itertools
import os

import numpy as np
import pytest

import pyoracle as O
from helpers import as_pairs, random_seqs, revcomp, write_fasta


# ---------------------------------------------------------------------------
# Scan: minimizer positions (vendor/simd-minimizers-2.3.1/src/lib.rs:86-137)
def test_minimizer_positions_known_answers(golden):
    for case in golden["minimizer_positions"]:
        seq = case.get("seq") or revcomp(case["seq_revcomp_of"])
        got = O.minimizer_positions(O.codes_of(seq), case["k"], case["w"], case["canonical"])
        assert got == case["positions"], case["src"]


def test_minimizer_values_known_answer(golden):
    case = golden["minimizer_positions"][0]
    codes = O.codes_of(case["seq"])
    for pos, val in zip(case["positions"], case["values_packed"]):
        # canonical k-mer value in packed-seq encoding, first base least significant
        fw = sum(int(c) << (2 * i) for i, c in enumerate(codes[pos:pos + case["k"]]))
        rcodes = [c ^ 2 for c in codes[pos:pos + case["k"]][::-1]]
        rc = sum(int(c) << (2 * i) for i, c in enumerate(rcodes))
        assert int(val, 2) in (fw, rc), case["src"]


def _py_hash(codes, m, seeds, rot, comb):
    M = 0xFFFFFFFF

    def rol(x, r):
        r %= 32
        return ((x << r) | (x >> (32 - r))) & M if r else x

    fw = rc = 0
    for j in range(m):
        fw ^= rol(seeds[codes[j]], rot * (m - 1 - j))
        rc ^= rol(seeds[codes[j] ^ 2], rot * j)
    if comb is None:
        return fw
    return {"add": (fw + rc) & M, "xor": fw ^ rc, "min": min(fw, rc)}[comb]


def _py_positions(seq, kh, w, canonical, seeds, rot, comb):
    codes = [int(c) for c in O.codes_of(seq)]
    h = [_py_hash(codes[i:], kh, seeds, rot, comb if canonical else None) for i in range(len(codes) - kh + 1)]
    l = kh + w - 1
    out = []
    for j in range(len(codes) - l + 1):
        keys = [h[j + i] >> 16 for i in range(w)]
        mn = min(keys)
        left = j + keys.index(mn)
        right = j + w - 1 - keys[::-1].index(mn)
        p = left
        if canonical and not (2 * sum(1 for c in codes[j:j + l] if c & 2) > l):
            p = right
        if not out or out[-1] != p:
            out.append(p)
    return out


def test_hash_variant_is_the_only_one_matching_known_answers(golden):
    """The three position vectors fix the ntHash variant: of 1008 plausible
    variants (seed order, seed half, rotation, combiner) exactly one matches
    all three, and it is the oracle's (a wrong rotation or combiner fails)."""
    seeds64 = [0x3C8BFBB395C60474, 0x3193C18562A02B4C, 0x20323ED082572324, 0x295549F54BE24456]
    cases = golden["minimizer_positions"]
    seqs = [c.get("seq") or revcomp(c["seq_revcomp_of"]) for c in cases]
    matches = []
    for half in ("lo", "hi"):
        base = [(s & 0xFFFFFFFF) if half == "lo" else (s >> 32) for s in seeds64]
        for perm in itertools.permutations(range(4)):
            seeds = [base[perm[x]] for x in range(4)]
            for rot in (1, 3, 5, 7, 9, 11, 13):
                for comb in ("add", "xor", "min"):
                    if all(_py_positions(s, c["k"], c["w"], c["canonical"], seeds, rot, comb) == c["positions"]
                           for s, c in zip(seqs, cases)):
                        matches.append((half, perm, rot, comb))
    assert matches == [("lo", (0, 1, 2, 3), 7, "add")]
    rng = np.random.default_rng(7)
    for _ in range(20):
        codes = rng.integers(0, 4, size=21).astype(np.uint8)
        want = _py_hash([int(c) for c in codes], 21, [s & 0xFFFFFFFF for s in seeds64], 7, "add")
        assert O.nthash32(codes, 21, True) == want


def test_dedup_with_index(golden):
    # collect_and_dedup_with_index (sm/collect.rs:38-75) through the runs API:
    # a sequence whose window hashes repeat in pairs.
    case = golden["dedup_with_index"][0]
    vals = case["input"]
    starts = [0] + [i for i in range(1, len(vals)) if vals[i] != vals[i - 1]]
    assert [vals[s] for s in starts] == case["values"]
    assert starts == case["starts"]


def test_runs_tile_windows(golden):
    for case in golden["runs_cover_windows"]:
        codes = O.codes_of(case["seq"])
        h, s = O.minimizer_runs(codes, case["k"], case["m"])
        nw = len(case["seq"]) - case["k"] + 1
        assert s[0] == 0 and all(np.diff(s) > 0) and s[-1] < nw, case["src"]
        wh = O.window_hashes(codes, case["k"], case["m"])
        assert len(wh) == nw
        # runs are maximal blocks of equal selected hash
        assert all(wh[s[i]] == h[i] for i in range(len(s)))
        assert all(wh[j] != wh[j - 1] for j in s[1:])


def test_window_hash_reverse_complement_symmetry():
    # sm/test.rs:112-152: the selected hash of a window equals that of its
    # reverse complement (window j <-> n-k-j).
    rng = np.random.default_rng(3)
    for k, m in ((31, 21), (5, 3), (21, 11), (63, 31), (7, 7)):
        for s in random_seqs(rng, 20, k, 200):
            a = O.window_hashes(O.codes_of(s), k, m)
            b = O.window_hashes(O.codes_of(revcomp(s)), k, m)
            assert np.array_equal(a, b[::-1]), (k, m, s)


def test_prefix_xor_form_of_nthash():
    """The scan kernel computes every m-mer hash from two exclusive prefix
    XORs of per-base terms rotated by their position (ntHash is XOR-linear):
    fw(i) = rotl(SA[i+m] ^ SA[i], 7(i+m-1)), rc(i) = rotr(SB[i+m] ^ SB[i], 7i).
    Any origin of the positions gives the same hashes as the oracle's rolling
    NtHasher<true>."""
    seeds = [0x95C60474, 0x62A02B4C, 0x82572324, 0x4BE24456]
    M = 0xFFFFFFFF

    def rotl(x, r):
        r &= 31
        return ((x << r) | (x >> (32 - r))) & M if r else x

    rng = np.random.default_rng(11)
    for m in (1, 2, 3, 11, 19, 21, 31):
        codes = rng.integers(0, 4, size=160).astype(np.uint8)
        for origin in (0, 5, 37):
            sa, sb = [0], [0]
            for j, c in enumerate(codes):
                p = origin + j
                sa.append(sa[-1] ^ rotl(seeds[c], -7 * p))
                sb.append(sb[-1] ^ rotl(seeds[c ^ 2], 7 * p))
            for i in range(len(codes) - m + 1):
                p = origin + i
                fw = rotl(sa[i + m] ^ sa[i], 7 * (p + m - 1))
                rc = rotl(sb[i + m] ^ sb[i], -7 * p)
                want = O.nthash32(np.ascontiguousarray(codes[i:i + m]), m, True)
                assert (fw + rc) & M == want, (m, origin, i)


# ---------------------------------------------------------------------------
# Counting
def test_exact_counts(golden):
    for case in golden["exact_counts"]:
        codes, off, _ = O.frags_of([case["seq"]])
        got = as_pairs(O.count_bruteforce(codes, off, case["k"], 0), case["k"])
        assert got == sorted((a, b) for a, b in case["counts"]), case["src"]


def test_count_inputs_cases(golden, tmp_path):
    for i, case in enumerate(golden["count_inputs"]):
        p = tmp_path / f"in{i}.fa"
        p.write_text(case["fasta"])
        codes, off = O.parse_fastx(str(p), case["k"])
        got, _, _ = O.count_two_pass(codes, off, case["k"], case["m"], case["x"], 1024)
        pairs = as_pairs(got, case["k"])
        if "exact" in case:
            assert pairs == [tuple(x) for x in case["exact"]], case["src"]
        for x in case.get("contains", []):
            assert tuple(x) in pairs, case["src"]


def test_two_pass_equals_bruteforce_random():
    rng = np.random.default_rng(11)
    genome = "".join(rng.choice(list("ACGT"), size=3000))
    reads = []
    for _ in range(400):
        p = int(rng.integers(0, len(genome) - 150))
        r = list(genome[p:p + 150])
        for j in range(len(r)):
            if rng.random() < 0.01:
                r[j] = "ACGT"[(("ACGT".index(r[j])) + 1) % 4]
        s = "".join(r)
        reads.append(s if rng.random() < 0.5 else revcomp(s))
    codes, off, _ = O.frags_of(reads)
    for k, m, x, P in ((31, 21, 1, 64), (31, 21, 5, 1024), (21, 11, 2, 7), (15, 15, 0, 3), (9, 5, 3, 16)):
        bf = O.count_bruteforce(codes, off, k, x)
        tp, _, _ = O.count_two_pass(codes, off, k, m, x, P)
        assert np.array_equal(bf, tp), (k, m, x, P)
        # the multithreaded form (the CPU baseline) gives the same array at
        # any thread count, including more threads than fragments
        for t in (1, 3, 8, 450):
            assert np.array_equal(O.count_two_pass_mt(codes, off, k, m, x, P, t), tp), (k, m, x, P, t)


def test_planted_mistake_is_caught():
    """A dropped window at each fragment end (a plausible off-by-one) must
    break the two-pass == brute-force check the tests rely on."""
    rng = np.random.default_rng(5)
    reads = random_seqs(rng, 50, 40, 80)
    codes, off, _ = O.frags_of(reads)
    bf = O.count_bruteforce(codes, off, 11, 0)
    trimmed = [r[:-1] for r in reads]
    c2, o2, _ = O.frags_of(trimmed)
    bad, _, _ = O.count_two_pass(c2, o2, 11, 7, 0, 8)
    assert not np.array_equal(bf, bad)


def test_kff_kmc_counts_match_closed_form():
    rng = np.random.default_rng(12)
    reads = random_seqs(rng, 300, 31, 120, alphabet="ACGT")
    reads += reads[:100]  # duplicates so some counts exceed the threshold
    codes, off, _ = O.frags_of(reads)
    for x in (0, 1, 3):
        data, nrec = O.kff_u8_kmc(codes, off, 31, 21, x, 16)
        assert data[:3] == b"KFF" and data[-3:] == b"KFF"
        bf = O.count_bruteforce(codes, off, 31, x)
        assert nrec == len(bf)
        # records: 8 bytes k-mer (BE) + 1 byte count, after the header
        body = data[len(data) - 3 - 9 * nrec:len(data) - 3]
        recs = sorted((int.from_bytes(body[i * 9:i * 9 + 8], "big"), body[i * 9 + 8]) for i in range(nrec))
        want = sorted((int(r["lo"]), min(int(r["count"]), min(x + 1, 255))) for r in bf)
        assert recs == want


def test_datasets(golden, tmp_path):
    for case in golden["datasets"]:
        codes, off, ds = O.frags_of(case["datasets"], datasets=list(range(len(case["datasets"]))))
        got, _ = O.count_datasets(codes, off, ds, len(case["datasets"]), case["k"], case["m"], case["x"], 1024)
        pairs = as_pairs(got, case["k"])
        for x in case.get("contains", []):
            assert tuple(x) in pairs, case["src"]
        for a in case.get("absent", []):
            assert a not in [p[0] for p in pairs], case["src"]
        if "exact" in case:
            assert pairs == [tuple(x) for x in case["exact"]], case["src"]
        if "kmer_fasta_selected" in case:
            assert len(pairs) == case["kmer_fasta_selected"], case["src"]
            out = tmp_path / "sel.fa"
            O.write_kmer_fasta(str(out), got, case["k"])
            text = out.read_text()
            assert case["text_contains"] in ("\n" + text), case["src"]
            if "text_lacks" in case:
                assert case["text_lacks"] not in ("\n" + text), case["src"]


def test_datasets_equal_bruteforce_random():
    rng = np.random.default_rng(21)
    base = "".join(rng.choice(list("ACGT"), size=2000))
    seqs, ds = [], []
    for d in range(6):
        g = list(base)
        for _ in range(10):
            p = int(rng.integers(0, len(g)))
            g[p] = "ACGT"[("ACGT".index(g[p]) + 1) % 4]
        g = "".join(g)
        for _ in range(60):
            p = int(rng.integers(0, len(g) - 100))
            seqs.append(g[p:p + 100])
            ds.append(d)
    codes, off, dsa = O.frags_of(seqs, datasets=ds)
    for x in (0, 2, 4, 6):
        bf = O.count_datasets_bruteforce(codes, off, dsa, 6, 21, x)
        got, kept = O.count_datasets(codes, off, dsa, 6, 21, 11, x, 32)
        assert np.array_equal(bf, got), x
        # the multithreaded form (the fof CPU baseline): same k-mers, counts
        # and retained minimizers at any thread count
        for t in (1, 3, 16, 500):
            got_mt, kept_mt = O.count_datasets_mt(codes, off, dsa, 6, 21, 11, x, 32, t)
            assert np.array_equal(got_mt, got) and kept_mt == kept, (x, t)


# ---------------------------------------------------------------------------
# Writers
def test_simplitigs(golden, tmp_path):
    for case in golden["simplitigs"]:
        k = case["k"]
        kms = np.zeros(len(case["kmers"]), dtype=O.COUNTED)
        for i, s in enumerate(case["kmers"]):
            kms[i]["lo"] = O.encode_canonical(s)
            kms[i]["count"] = 1
        kms.sort(order=["hi", "lo"])
        p = tmp_path / "s.fa"
        O.write_simplitigs(str(p), kms, k)
        seqs = [l for l in p.read_text().split("\n") if l and not l.startswith(">")]
        exp = case["expected"]
        assert exp in seqs or revcomp(exp) in seqs, case["src"]
        assert sum(len(s) - k + 1 for s in seqs) == len(set(kms["lo"].tolist()))


def test_fasta_writer(golden, tmp_path):
    for case in golden["fasta_writer"]:
        kms = np.zeros(len(case["kmers"]), dtype=O.COUNTED)
        for i, s in enumerate(case["kmers"]):
            kms[i]["lo"] = O.encode_canonical(s)
            kms[i]["count"] = 1
        p = tmp_path / "f.fa"
        last = O.write_simplitigs(str(p), kms, case["k"], case["first_index"])
        text = p.read_text()
        for c in case["contains"]:
            assert c in text, case["src"]
        if "last_index" in case:
            assert last == case["last_index"]


def test_kff_writer(golden, tmp_path):
    case = golden["kff_writer"][0]
    kms = np.zeros(1, dtype=O.COUNTED)
    kms[0]["lo"] = O.encode_canonical(case["kmers"][0][0])
    kms[0]["count"] = case["kmers"][0][1]
    p = tmp_path / "o.kff"
    O.write_kff(str(p), kms, case["k"], case["m"], case["x"], 8, True)
    b = p.read_bytes()
    assert b.startswith(b"KFF") and b.endswith(b"KFF")


# ---------------------------------------------------------------------------
# ZI and input
def test_zi(golden, tmp_path):
    for case in golden["zi"]:
        k = case["k"]
        if "query" in case:
            (tmp_path / "i.fa").write_text(case["index"])
            (tmp_path / "q.fa").write_text(case["query"])
            ic, io = O.parse_fastx(str(tmp_path / "i.fa"), 1)
            qc, qo = O.parse_fastx(str(tmp_path / "q.fa"), k)
            ab = O.zi_absent(ic, io, qc, qo, k)
            if "absent" in case:
                assert [p[0] for p in as_pairs(ab, k)] == case["absent"], case["src"]
            if "output_contains_any" in case:
                O.write_simplitigs(str(tmp_path / "o.fa"), ab, k)
                text = (tmp_path / "o.fa").read_text()
                assert any(s in text for s in case["output_contains_any"]), case["src"]
        elif "queries" in case:
            ic, io, _ = O.frags_of(case["index_simplitigs"])
            scanned = filtered = 0
            for qi, q in enumerate(case["queries"]):
                (tmp_path / f"q{qi}.fa").write_text(q)
                qc, qo = O.parse_fastx(str(tmp_path / f"q{qi}.fa"), k)
                n = int(qo[-1]) - (len(qo) - 1) * (k - 1)
                scanned += n
                # windows absent = windows whose canonical k-mer is not indexed
                idx = set(as_pairs(O.count_bruteforce(ic, io, k, 0), k))
                idxk = {a for a, _ in idx}
                for f in range(len(qo) - 1):
                    fr = "".join("ACTG"[c] for c in qc[qo[f]:qo[f + 1]])
                    for j in range(len(fr) - k + 1):
                        w = fr[j:j + k]
                        cw = min(w, revcomp(w))
                        if cw in idxk:
                            filtered += 1
            assert scanned == case["scanned"] and filtered == case["filtered"], case["src"]
            assert scanned - filtered == case["absent_windows"], case["src"]
        else:
            p = tmp_path / "e.fa"
            O.write_simplitigs(str(p), np.zeros(0, dtype=O.COUNTED), k)
            assert p.read_text() == case["output"]


def test_fastx_reader(golden, tmp_path):
    for case in golden["fastx_reader"]:
        p = tmp_path / "r.fa"
        p.write_text(case["fasta"])
        codes, off = O.parse_fastx(str(p), case["min_len"])
        frags = ["".join("ACTG"[c] for c in codes[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
        assert frags == case["fragments"], case["src"]


def test_fastq_and_lowercase(tmp_path):
    p = tmp_path / "r.fq"
    p.write_text("@a\nACGTNacgt\n+\nIIIIIIIII\n@b\nGGGG\n+\nIIII\n")
    codes, off = O.parse_fastx(str(p), 3)
    frags = ["".join("ACTG"[c] for c in codes[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
    assert frags == ["ACGT", "ACGT", "GGGG"]


def test_spot_counts_match_bruteforce():
    """The full-size spot counter (threads over fragment ranges, packed
    words) against the brute-force closed form, including k-mers that do not
    occur and dataset-presence masks."""
    rng = np.random.default_rng(21)
    seqs = random_seqs(rng, 300, 20, 400)
    codes, off, _ = O.frags_of(seqs)
    words = np.zeros((codes.size + 15) // 16 + 4, dtype=np.uint32)
    for j in range(16):
        part = codes[j::16].astype(np.uint32) << (2 * j)
        words[: part.size] |= part
    ds = (np.arange(off.size - 1) % 5).astype(np.uint32)
    for k in (31, 15, 5):
        want = O.count_bruteforce(codes, off, k, 0)
        wmask = O.count_datasets_bruteforce(codes, off, ds, 5, k, 0)
        present = dict(zip(want["lo"].tolist(), want["count"].tolist()))
        npres = dict(zip(wmask["lo"].tolist(), wmask["count"].tolist()))
        q = np.concatenate([want["lo"][::7], rng.integers(0, 1 << (2 * k), size=200, dtype=np.uint64)])
        for threads in (1, 6):
            c, m = O.spot_counts(words, off, k, q, ds, threads)
            for key, cnt, mask in zip(q.tolist(), c.tolist(), m.tolist()):
                assert cnt == present.get(key, 0), (k, key)
                assert bin(mask).count("1") == npres.get(key, 0), (k, key)
        # kmers_at gives the canonical k-mer of a window
        f = 3
        p0 = int(off[f])
        assert O.kmers_at(words, [p0], k)[0] == O.encode_canonical("".join("ACTG"[c] for c in codes[p0:p0 + k]))


def test_spot_counts128_match_bruteforce():
    """The 128-bit spot counter (33 <= k <= 64) and kmers_at128 against the
    brute-force counter, with a repeated region so counts exceed one and
    random keys that do not occur."""
    rng = np.random.default_rng(22)
    seqs = random_seqs(rng, 200, 70, 500)
    seqs += seqs[:15]
    codes, off, _ = O.frags_of(seqs)
    words = np.zeros((codes.size + 15) // 16 + 4, dtype=np.uint32)
    for j in range(16):
        part = codes[j::16].astype(np.uint32) << (2 * j)
        words[: part.size] |= part
    for k in (33, 47, 63, 64):
        want = O.count_bruteforce(codes, off, k, 0)
        present = {(lo, hi): c for lo, hi, c in zip(want["lo"].tolist(), want["hi"].tolist(),
                                                    want["count"].tolist())}
        assert max(present.values()) > 1
        qlo = np.concatenate([want["lo"][::5], rng.integers(0, 1 << 63, size=100, dtype=np.uint64)])
        qhi = np.concatenate([want["hi"][::5], rng.integers(0, 1 << (2 * k - 64), size=100, dtype=np.uint64)])
        for threads in (1, 5):
            c = O.spot_counts128(words, off, k, qlo, qhi, threads)
            for lo, hi, cnt in zip(qlo.tolist(), qhi.tolist(), c.tolist()):
                assert cnt == present.get((lo, hi), 0), (k, lo, hi)
        f = 4
        p0 = int(off[f])
        lo, hi = O.kmers_at128(words, [p0], k)
        v = O.encode_canonical("".join("ACTG"[c] for c in codes[p0:p0 + k]))
        assert (int(hi[0]) << 64 | int(lo[0])) == v


def test_zi_lookup_loop_matches_set_difference():
    rng = np.random.default_rng(77)
    g = "".join(rng.choice(list("ACGT"), size=4000))
    other = "".join(rng.choice(list("ACGT"), size=4000))
    qs = [g[p:p + 150] for p in rng.integers(0, 3850, size=60)] + [other[p:p + 150] for p in rng.integers(0, 3850, size=60)]
    ic, io, _ = O.frags_of([g])
    qc, qo, _ = O.frags_of(qs)
    for k in (21, 31, 45):
        nl, na, _, _ = O.zi_lookup_timed(ic, io, qc, qo, k)
        assert nl == 120 * (150 - k + 1)
        assert na == len(O.zi_absent(ic, io, qc, qo, k))
