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


@pytest.mark.parametrize("repeat", [0, 100])
def test_zi_large_request_sort_paths(tmp_path, repeat):
    """A request with more than 65536 absent windows is ordered by radix
    passes over the top 24 key bits plus the finishing pass (equal k-mers
    included); with one read repeated 100 times, runs of equal prefixes exceed
    64 and the request takes the full sort. Both equal the oracle's set."""
    print(f"zi large request repeat={repeat}", flush=True)
    rng = np.random.default_rng(7 + repeat)
    genome = _genome(rng, 20000)
    other = _genome(rng, 40000)
    write_fasta(tmp_path / "idx.fa", [genome])
    reads = _queries(rng, genome, other, 2400)
    reads += [reads[1]] * repeat
    q = tmp_path / "q.fa"
    write_fasta(q, reads)
    got = zi_absent(tmp_path / "idx.fa", [q], 31)
    ic, io = O.parse_fastx(str(tmp_path / "idx.fa"), 1)
    qc, qo = O.parse_fastx(str(q), 31)
    want = O.zi_absent(ic, io, qc, qo, 31)
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


def test_zi_full_size_index_spot_check():
    """Serving at full index size: index = canonical 31-mers of a 100 Mbp
    genome (seed 5); one request of 200k 150 bp reads, half from the index
    genome (0.5% substitutions) and half from an unrelated genome. Sampled
    absent k-mers never occur in the index; sampled query windows are absent
    exactly when they do not occur in it (oracle spot counter over the index
    genome)."""
    G = 100_000_000
    ga = M.synth_genome(G, 5)
    gb = M.synth_genome(G, 1005)
    codes = np.unpackbits(ga.view(np.uint8)[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
    codes = (codes[:, :, 0] | (codes[:, :, 1] << 1)).reshape(-1)[:G]
    text = b">index\n" + np.frombuffer(b"ACTG", dtype=np.uint8)[codes].tobytes() + b"\n"
    del codes
    reads = M.synth_reads_from([(ga, G), (gb, G)], 200_000, 150, 0.005, 6)
    h = C.c_void_p()
    M._check(M._lib.mczi_zi_server_create_from_text(text, len(text), 31, 0, 64 << 30, 2, C.byref(h)))
    try:
        off = reads.frag_off.astype(np.uint64)
        rid = C.c_uint64()
        M._check(M._lib.mczi_zi_server_submit_packed(h, reads.words.ctypes.data, off.ctypes.data, off.size - 1, 0,
                                                      C.byref(rid)))
        p = C.POINTER(M._ZiResult)()
        M._check(M._lib.mczi_zi_server_poll(h, rid.value, 1, C.byref(p)))
        r = p.contents
        absent = np.ctypeslib.as_array(r.absent_lo, shape=(r.n_absent,)).copy()
        n_windows = r.n_windows
        M._lib.mczi_zi_result_free(p)
    finally:
        M._lib.mczi_zi_server_destroy(h)
    assert n_windows == 200_000 * 120
    assert np.all(np.diff(absent) > 0) and 0.3 * n_windows < absent.size < n_windows
    rng = np.random.default_rng(5)
    gwords = ga.view(np.uint32)
    goff = np.array([0, G], dtype=np.uint64)
    # sampled absent k-mers: none occurs in the index genome
    sa = absent[rng.integers(0, absent.size, size=3000)]
    c, _ = O.spot_counts(gwords, goff, 31, sa)
    assert np.all(c == 0)
    # sampled query windows: absent iff not in the index genome
    f = rng.integers(0, off.size - 1, size=3000)
    pos = off[f] + rng.integers(0, 150 - 31 + 1, size=3000).astype(np.uint64)
    wk = O.kmers_at(reads.words, pos, 31)
    c, _ = O.spot_counts(gwords, goff, 31, wk)
    at = np.minimum(np.searchsorted(absent, wk), absent.size - 1)
    assert np.array_equal(absent[at] == wk, c == 0)


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
