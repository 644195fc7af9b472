"""Packed inputs with empty and short fragments (MI355X only).

The packed entry points accept any non-decreasing fragment offsets. Runs of
empty fragments (several fragment starts at one base position) and
fragments shorter than k must leave the counts, dataset presence and ZI
absent sets exactly those of the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O
from helpers import pack_words, random_seqs, same_counted, with_empty_fragments

pytestmark = pytest.mark.gpu


def _seqs(rng, n):
    # short fragments (below k), reads, and a few long ones crossing tiles
    return (random_seqs(rng, n, 1, 30) + random_seqs(rng, n, 31, 200) + random_seqs(rng, 3, 3000, 9000))


@pytest.mark.parametrize("k,m,x", [(31, 21, 0), (21, 11, 1), (45, 27, 0)])
def test_count_with_empty_fragments(k, m, x):
    print(f"empty fragments: count k={k} m={m} x={x}", flush=True)
    rng = np.random.default_rng(k)
    seqs = _seqs(rng, 300)
    rng.shuffle(seqs)
    codes, off, src = with_empty_fragments(rng, seqs, runs=8, run_len=50)
    want = O.count_bruteforce(codes, off, k, x)
    words = pack_words(codes)
    # input marks at every 37th fragment, some of them on empty runs
    marks = np.unique(np.concatenate([np.arange(0, off.size - 1, 37), [off.size - 1]])).astype(np.uint64)
    for b in (M.Batch.from_packed(words, off), M.Batch.from_packed(words, off, input_frag=marks)):
        assert b.info()[1] == int((src >= 0).sum())
        for launch in (M.LAUNCH_BATCHED, M.LAUNCH_PLAIN):
            b.count(M.CounterConfig(k, m, x, 64), M.MODE_TWO_PASS, launch)
            assert same_counted(b.download(), want), launch


def test_datasets_with_empty_fragments():
    print("empty fragments: dataset presence", flush=True)
    rng = np.random.default_rng(7)
    base = "".join(rng.choice(list("ACGT"), size=3000))
    seqs = []
    for d in range(6):
        for _ in range(60):
            p = int(rng.integers(0, 3000 - 120))
            seqs.append(base[p:p + int(rng.integers(10, 120))])
    ds_of_seq = np.repeat(np.arange(6), 60)
    codes, off, src = with_empty_fragments(rng, seqs, runs=10, run_len=33)
    ds = np.where(src >= 0, ds_of_seq[np.maximum(src, 0)], 0).astype(np.uint32)
    b = M.Batch.from_packed(pack_words(codes), off, ds)
    for x in (0, 2, 4):
        b.count(M.CounterConfig(31, 21, x, 64), M.MODE_DATASETS)
        want = O.count_datasets_bruteforce(codes, off, ds, 6, 31, x)
        assert same_counted(b.download(), want), x


def test_zi_packed_request_with_empty_fragments():
    print("empty fragments: ZI packed request", flush=True)
    rng = np.random.default_rng(8)
    genome = "".join(rng.choice(list("ACGT"), size=20000))
    seqs = [genome[p:p + int(rng.integers(5, 200))] for p in rng.integers(0, 19000, size=300)]
    seqs += random_seqs(rng, 200, 5, 200)
    codes, off, _ = with_empty_fragments(rng, seqs, runs=8, run_len=40)
    ic, io, _ = O.frags_of([genome])
    want = O.zi_absent(ic, io, codes, off, 31)
    h = C.c_void_p()
    t = f">g\n{genome}\n".encode()
    M._check(M._lib.mczi_zi_server_create_from_text(t, len(t), 31, 0, 1 << 30, 2, C.byref(h)))
    try:
        words = pack_words(codes)
        rid = C.c_uint64()
        M._check(M._lib.mczi_zi_server_submit_packed(h, words.ctypes.data, off.ctypes.data, off.size - 1, 0,
                                                      C.byref(rid)))
        p = C.POINTER(M._ZiResult)()
        M._check(M._lib.mczi_zi_server_poll(h, rid.value, 1, C.byref(p)))
        r = p.contents
        got = np.ctypeslib.as_array(r.absent_lo, shape=(r.n_absent,)).copy() if r.n_absent else np.zeros(0, np.uint64)
        M._lib.mczi_zi_result_free(p)
    finally:
        M._lib.mczi_zi_server_destroy(h)
    assert np.array_equal(got, want["lo"])
