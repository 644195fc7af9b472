"""ZI on the device vs the oracle's exact set difference (MI355X only)."""
import ctypes as C
import threading

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import random_seqs, revcomp, same_counted, write_fasta

pytestmark = pytest.mark.gpu


def zi_absent(index_path, query_paths, k):
    arr, n = M._paths(query_paths)
    p = C.POINTER(M._Counted)()
    nout = C.c_size_t()
    M._check(M._lib.mczi_zi_absent(str(index_path).encode(), arr, n, k, C.byref(p), C.byref(nout)))
    return M._take(p, nout.value)


def _genome(rng, n):
    return "".join(rng.choice(list("ACGT"), size=n))


def _queries(rng, genome, other, n, L=150, err=0.005):
    out = []
    for i in range(n):
        src = genome if i % 2 == 0 else other
        p = int(rng.integers(0, len(src) - L))
        r = list(src[p:p + L])
        for j in range(L):
            if rng.random() < err:
                r[j] = "ACGT"[("ACGT".index(r[j]) + 1) % 4]
        s = "".join(r)
        out.append(s if rng.random() < 0.5 else revcomp(s))
    return out


@pytest.mark.parametrize("k", [31, 63, 3, 21, 33, 45])
def test_zi_absent_matches_oracle(tmp_path, k):
    print(f"zi k={k}", flush=True)
    rng = np.random.default_rng(k)
    genome = _genome(rng, 30000)
    other = _genome(rng, 30000)
    write_fasta(tmp_path / "idx.fa", [genome[:15000], genome[15000:]])
    q1 = tmp_path / "q1.fa"
    q2 = tmp_path / "q2.fa"
    write_fasta(q1, _queries(rng, genome, other, 300))
    write_fasta(q2, _queries(rng, genome, other, 200) + ["ACGTN" * 40, "A" * 200])
    got = zi_absent(tmp_path / "idx.fa", [q1, q2], k)
    ic, io = O.parse_fastx(str(tmp_path / "idx.fa"), 1)
    c1, o1 = O.parse_fastx(str(q1), k)
    c2, o2 = O.parse_fastx(str(q2), k)
    qc = np.concatenate([c1, c2])
    qo = np.concatenate([o1, o2[1:] + o1[-1]])
    want = O.zi_absent(ic, io, qc, qo, k)
    assert same_counted(got, want), (len(got), len(want))


@pytest.mark.parametrize("k", [31, 63])
@pytest.mark.parametrize("repeat", [0, 100])
def test_zi_large_request_sort_paths(tmp_path, repeat, k):
    """A request with more than 65536 absent windows is ordered by radix
    passes over the top 24 key bits (k = 63: of the high word, carrying the
    low word's position) plus the finishing pass (equal k-mers included);
    with one read repeated 100 times, runs of equal prefixes exceed 64 and
    the request takes the full sort. Both equal the oracle's set."""
    print(f"zi large request repeat={repeat} k={k}", flush=True)
    rng = np.random.default_rng(7 + repeat)
    genome = _genome(rng, 20000)
    other = _genome(rng, 40000)
    write_fasta(tmp_path / "idx.fa", [genome])
    reads = _queries(rng, genome, other, 2400)
    reads += [reads[1]] * repeat
    q = tmp_path / "q.fa"
    write_fasta(q, reads)
    got = zi_absent(tmp_path / "idx.fa", [q], k)
    ic, io = O.parse_fastx(str(tmp_path / "idx.fa"), 1)
    qc, qo = O.parse_fastx(str(q), k)
    want = O.zi_absent(ic, io, qc, qo, k)
    assert len(want) > 20000
    assert same_counted(got, want), (len(got), len(want))


def test_zi_golden(golden, tmp_path):
    for i, case in enumerate(golden["zi"]):
        if "query" not in case:
            continue
        (tmp_path / f"i{i}.fa").write_text(case["index"])
        (tmp_path / f"q{i}.fa").write_text(case["query"])
        out = tmp_path / f"o{i}.fa"
        M._check(M._lib.mczi_zi_subtract(str(tmp_path / f"i{i}.fa").encode(), M._paths([tmp_path / f"q{i}.fa"])[0],
                                         1, case["k"], case["m"], str(out).encode(), 1, 1))
        text = out.read_text()
        if "output_contains_any" in case:
            assert any(s in text for s in case["output_contains_any"]), case["src"]
        if "absent" in case:
            got = zi_absent(tmp_path / f"i{i}.fa", [tmp_path / f"q{i}.fa"], case["k"])
            assert [O.decode(int(r["lo"]), int(r["hi"]), case["k"]) for r in got] == case["absent"]


def test_zi_fasta_bytes_match_oracle(tmp_path):
    rng = np.random.default_rng(4)
    genome = _genome(rng, 20000)
    other = _genome(rng, 20000)
    write_fasta(tmp_path / "idx.fa", [genome])
    write_fasta(tmp_path / "q.fa", _queries(rng, genome, other, 400))
    out = tmp_path / "o.fa"
    M._check(M._lib.mczi_zi_subtract(str(tmp_path / "idx.fa").encode(), M._paths([tmp_path / "q.fa"])[0], 1, 31, 0,
                                     str(out).encode(), 0, 8))
    ic, io = O.parse_fastx(str(tmp_path / "idx.fa"), 1)
    qc, qo = O.parse_fastx(str(tmp_path / "q.fa"), 31)
    want = O.zi_absent(ic, io, qc, qo, 31)
    O.write_simplitigs(str(tmp_path / "w.fa"), want, 31)
    assert out.read_bytes() == (tmp_path / "w.fa").read_bytes()


def _server(index_text, k, budget=1 << 30, inflight=2):
    h = C.c_void_p()
    b = index_text.encode()
    M._check(M._lib.mczi_zi_server_create_from_text(b, len(b), k, 0, budget, inflight, C.byref(h)))
    return h


def _submit(h, text, fasta=False):
    rid = C.c_uint64()
    b = text.encode()
    M._check(M._lib.mczi_zi_server_submit(h, b, len(b), int(fasta), C.byref(rid)))
    return rid.value


def _poll(h, rid):
    p = C.POINTER(M._ZiResult)()
    M._check(M._lib.mczi_zi_server_poll(h, rid, 1, C.byref(p)))
    r = p.contents
    arr = np.zeros(r.n_absent, dtype=M.COUNTED)
    if r.n_absent:
        arr["lo"] = np.ctypeslib.as_array(r.absent_lo, shape=(r.n_absent,))
        if r.absent_hi:
            arr["hi"] = np.ctypeslib.as_array(r.absent_hi, shape=(r.n_absent,))
        arr["count"] = 1
    fasta = C.string_at(r.fasta, r.fasta_len).decode()
    M._lib.mczi_zi_result_free(p)
    return arr, fasta


@pytest.mark.parametrize("k", [31, 63])
def test_server_results_same_alone_and_in_mix(k):
    """Continuous batching: each request's result equals its result alone
    and the oracle's, whatever else shares its batch."""
    rng = np.random.default_rng(100 + k)
    genome = _genome(rng, 40000)
    other = _genome(rng, 40000)
    idx_text = f">g\n{genome}\n"
    reqs = []
    for i in range(24):
        n = int(rng.integers(1, 120))
        reqs.append("".join(f">q{j}\n{s}\n" for j, s in enumerate(_queries(rng, genome, other, n))))
    ic, io, _ = O.frags_of([genome])
    wants = []
    for t in reqs:
        seqs = [l for l in t.split("\n") if l and not l.startswith(">")]
        qc, qo, _ = O.frags_of(seqs)
        wants.append(O.zi_absent(ic, io, qc, qo, k))
    h = _server(idx_text, k, budget=64 << 20, inflight=3)
    try:
        alone = []
        for t in reqs[:4]:
            alone.append(_poll(h, _submit(h, t, fasta=True)))
        # mixed: submit everything from several threads, then collect
        ids = [None] * len(reqs)

        def sub(lo, hi):
            for i in range(lo, hi):
                ids[i] = _submit(h, reqs[i], fasta=True)

        ths = [threading.Thread(target=sub, args=(i * 6, i * 6 + 6)) for i in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        mixed = [_poll(h, i) for i in ids]
    finally:
        M._lib.mczi_zi_server_destroy(h)
    for i in range(len(reqs)):
        assert same_counted(mixed[i][0], wants[i]), i
    for i in range(4):
        assert same_counted(alone[i][0], mixed[i][0]) and alone[i][1] == mixed[i][1]


def _codes_of(words, n):
    """Packed-seq codes (A0 C1 T2 G3) of the first n bases of packed words."""
    b = np.unpackbits(np.ascontiguousarray(words).view(np.uint8)[:, None], axis=1, bitorder="little")
    b = b.reshape(-1, 4, 2)
    return (b[:, :, 0] | (b[:, :, 1] << 1)).reshape(-1)[:n].astype(np.uint8)


@pytest.mark.parametrize("k,m", [(31, 19), (63, 23)])
def test_server_request_stream_spot_check(k, m):
    """The request stream's shape at reduced scale: 48 requests whose read
    counts are log-uniform in [10, 1e5], half from the index genome (0.5%
    substitutions) and half unrelated, submitted from three threads with
    exponential gaps against a 4 Mbp index. Every result has the request's
    window count and a strictly increasing absent set; the eight smallest
    equal the oracle's set difference; for every request, 200 sampled query
    windows are absent exactly when the oracle spot counter finds them 0
    times in the index genome, and 200 sampled absent k-mers occur 0 times."""
    import time
    G = 4_000_000
    ga = M.synth_genome(G, 15)
    gb = M.synth_genome(G, 1015)
    gcodes = _codes_of(ga, G)
    text = b">index\n" + np.frombuffer(b"ACTG", dtype=np.uint8)[gcodes].tobytes() + b"\n"
    rng = np.random.default_rng(40 + k)
    n_req = 48
    counts = np.exp(rng.uniform(np.log(10), np.log(1e5), size=n_req)).astype(np.int64)
    reqs = [M.synth_reads_from([(ga, G), (gb, G)], int(c), 150, 0.005, 500 + i) for i, c in enumerate(counts)]
    gaps = rng.exponential(0.002, size=n_req)
    h = C.c_void_p()
    M._check(M._lib.mczi_zi_server_create_from_text(text, len(text), k, m, 8 << 30, 4, C.byref(h)))
    ids = [None] * n_req
    errs = []

    def sub(which):
        try:
            for i in which:
                time.sleep(gaps[i])
                off = reqs[i].frag_off.astype(np.uint64)
                rid = C.c_uint64()
                M._check(M._lib.mczi_zi_server_submit_packed(h, reqs[i].words.ctypes.data, off.ctypes.data,
                                                              off.size - 1, 0, C.byref(rid)))
                ids[i] = rid.value
        except Exception as e:  # reported after join
            errs.append(e)

    results = []
    try:
        ths = [threading.Thread(target=sub, args=(range(t, n_req, 3),)) for t in range(3)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert not errs, errs
        for i in range(n_req):
            p = C.POINTER(M._ZiResult)()
            M._check(M._lib.mczi_zi_server_poll(h, ids[i], 1, C.byref(p)))
            r = p.contents
            lo = np.ctypeslib.as_array(r.absent_lo, shape=(r.n_absent,)).copy() if r.n_absent else np.zeros(0, np.uint64)
            hi = (np.ctypeslib.as_array(r.absent_hi, shape=(r.n_absent,)).copy() if k > 32 and r.n_absent
                  else np.zeros_like(lo))
            results.append((r.n_windows, lo, hi))
            M._lib.mczi_zi_result_free(p)
    finally:
        M._lib.mczi_zi_server_destroy(h)
    gwords = ga.view(np.uint32)
    goff = np.array([0, G], dtype=np.uint64)
    ic, io = gcodes, np.array([0, G], dtype=np.uint64)
    smallest = set(np.argsort(counts)[:8].tolist())
    for i, (n_windows, lo, hi) in enumerate(results):
        rq = reqs[i]
        off = rq.frag_off.astype(np.uint64)
        assert n_windows == int(counts[i]) * (150 - k + 1), i
        assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))), i
        if i in smallest:
            qc = _codes_of(rq.words, int(off[-1]))
            want = O.zi_absent(ic, io, qc, off, k)
            got = np.zeros(lo.size, dtype=M.COUNTED)
            got["lo"], got["hi"], got["count"] = lo, hi, 1
            assert same_counted(got, want), i
        f = rng.integers(0, off.size - 1, size=200)
        pos = off[f] + rng.integers(0, 150 - k + 1, size=200).astype(np.uint64)
        sa = rng.integers(0, max(lo.size, 1), size=200 if lo.size else 0)
        if k <= 32:
            wk = O.kmers_at(rq.words, pos, k)
            c, _ = O.spot_counts(gwords, goff, k, wk)
            at = np.minimum(np.searchsorted(lo, wk), max(lo.size - 1, 0))
            is_absent = (lo[at] == wk) if lo.size else np.zeros(wk.size, bool)
            assert np.array_equal(is_absent, c == 0), i
            c, _ = O.spot_counts(gwords, goff, k, lo[sa])
        else:
            wlo, whi = O.kmers_at128(rq.words, pos, k)
            c = O.spot_counts128(gwords, goff, k, wlo, whi)
            left, right = np.searchsorted(hi, whi, "left"), np.searchsorted(hi, whi, "right")
            is_absent = []
            for a, b, x in zip(left, right, wlo):
                j = a + np.searchsorted(lo[a:b], x)
                is_absent.append(bool(j < b and lo[j] == x))
            assert np.array_equal(np.array(is_absent), c == 0), i
            c = O.spot_counts128(gwords, goff, k, lo[sa], hi[sa])
        assert np.all(c == 0), i


def test_server_rejects_over_budget():
    rng = np.random.default_rng(1)
    genome = _genome(rng, 5000)
    h = _server(f">g\n{genome}\n", 31, budget=1 << 20, inflight=1)
    try:
        big = "".join(f">q{j}\n{genome}\n" for j in range(40))
        rid = _submit(h, big)
        with pytest.raises(M.MczaError, match="HBM budget"):
            _poll(h, rid)
        # the server keeps serving after a rejection
        arr, _ = _poll(h, _submit(h, f">q\n{genome[:200]}\n"))
        assert len(arr) == 0
    finally:
        M._lib.mczi_zi_server_destroy(h)


@pytest.mark.parametrize("k,m", [(31, 0), (63, 23)])
def test_zi_full_size_index_spot_check(k, m):
    """Serving at full index size: index = canonical k-mers of a 100 Mbp
    genome (seed 5); one request of 200k 150 bp reads, half from the index
    genome (0.5% substitutions) and half from an unrelated genome. Sampled
    absent k-mers never occur in the index; sampled query windows are absent
    exactly when they do not occur in it (oracle spot counter over the index
    genome). k=63 with m=23 is the 128-bit request-stream configuration."""
    G = 100_000_000
    ga = M.synth_genome(G, 5)
    gb = M.synth_genome(G, 1005)
    codes = np.unpackbits(ga.view(np.uint8)[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
    codes = (codes[:, :, 0] | (codes[:, :, 1] << 1)).reshape(-1)[:G]
    text = b">index\n" + np.frombuffer(b"ACTG", dtype=np.uint8)[codes].tobytes() + b"\n"
    del codes
    reads = M.synth_reads_from([(ga, G), (gb, G)], 200_000, 150, 0.005, 6)
    h = C.c_void_p()
    M._check(M._lib.mczi_zi_server_create_from_text(text, len(text), k, m, 64 << 30, 2, C.byref(h)))
    del text
    try:
        off = reads.frag_off.astype(np.uint64)
        rid = C.c_uint64()
        M._check(M._lib.mczi_zi_server_submit_packed(h, reads.words.ctypes.data, off.ctypes.data, off.size - 1, 0,
                                                      C.byref(rid)))
        p = C.POINTER(M._ZiResult)()
        M._check(M._lib.mczi_zi_server_poll(h, rid.value, 1, C.byref(p)))
        r = p.contents
        lo = np.ctypeslib.as_array(r.absent_lo, shape=(r.n_absent,)).copy()
        hi = (np.ctypeslib.as_array(r.absent_hi, shape=(r.n_absent,)).copy() if k > 32
              else np.zeros_like(lo))
        n_windows = r.n_windows
        M._lib.mczi_zi_result_free(p)
    finally:
        M._lib.mczi_zi_server_destroy(h)
    assert n_windows == 200_000 * (150 - k + 1)
    # strictly increasing as 128-bit values
    assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1])))
    assert 0.3 * n_windows < lo.size < n_windows
    rng = np.random.default_rng(5)
    gwords = ga.view(np.uint32)
    goff = np.array([0, G], dtype=np.uint64)
    sa = rng.integers(0, lo.size, size=3000)
    f = rng.integers(0, off.size - 1, size=3000)
    pos = off[f] + rng.integers(0, 150 - k + 1, size=3000).astype(np.uint64)
    if k <= 32:
        # sampled absent k-mers: none occurs in the index genome
        c, _ = O.spot_counts(gwords, goff, k, lo[sa])
        assert np.all(c == 0)
        # sampled query windows: absent iff not in the index genome
        wk = O.kmers_at(reads.words, pos, k)
        c, _ = O.spot_counts(gwords, goff, k, wk)
        at = np.minimum(np.searchsorted(lo, wk), lo.size - 1)
        assert np.array_equal(lo[at] == wk, c == 0)
    else:
        c = O.spot_counts128(gwords, goff, k, lo[sa], hi[sa])
        assert np.all(c == 0)
        wlo, whi = O.kmers_at128(reads.words, pos, k)
        c = O.spot_counts128(gwords, goff, k, wlo, whi)
        left, right = np.searchsorted(hi, whi, "left"), np.searchsorted(hi, whi, "right")
        is_absent = np.array([a < b and lo[a + np.searchsorted(lo[a:b], x)] == x if a + np.searchsorted(lo[a:b], x) < b
                              else False for a, b, x in zip(left, right, wlo)])
        assert np.array_equal(is_absent, c == 0)
        # both outcomes occur among the sampled windows
        assert 0 < is_absent.sum() < is_absent.size


def test_zi_query_passes_equal_one_pass(tmp_path, monkeypatch):
    """ram_limit_gib splits the query fragments into several lookup passes
    (forced small here through MCZI_ZI_PASS_BYTES); the merged absent set and
    the FASTA bytes equal one pass and the oracle."""
    rng = np.random.default_rng(12)
    genome = _genome(rng, 30000)
    other = _genome(rng, 30000)
    write_fasta(tmp_path / "idx.fa", [genome])
    write_fasta(tmp_path / "q.fa", _queries(rng, genome, other, 900))
    ic, io = O.parse_fastx(str(tmp_path / "idx.fa"), 1)
    qc, qo = O.parse_fastx(str(tmp_path / "q.fa"), 31)
    want = O.zi_absent(ic, io, qc, qo, 31)
    outs = []
    for pass_bytes in (None, "200000", "5000"):
        if pass_bytes:
            monkeypatch.setenv("MCZI_ZI_PASS_BYTES", pass_bytes)
        out = tmp_path / f"o{pass_bytes}.fa"
        M._check(M._lib.mczi_zi_subtract(str(tmp_path / "idx.fa").encode(), M._paths([tmp_path / "q.fa"])[0], 1, 31,
                                         0, str(out).encode(), 4, 8))
        outs.append(out.read_bytes())
        assert same_counted(zi_absent(tmp_path / "idx.fa", [tmp_path / "q.fa"], 31), want)
    assert outs[0] == outs[1] == outs[2]
