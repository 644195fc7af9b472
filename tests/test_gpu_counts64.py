"""Counts of 2^32 and more (MI355X only). The reference's counts are u64
(add_count saturates a u64, src/lib.rs:3003-3010); the device tables hold
32-bit counts and carry what passes 2^32. A poly-A fragment of 2^32 + 99
bases holds one canonical k-mer n - k + 1 times; a second fragment adds
ordinary k-mers around it."""
import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu


def _batch(n):
    # fragment 0: n bases of A (code 0); fragment 1: 300 random bases
    rng = np.random.default_rng(64)
    tail = rng.integers(0, 4, size=300).astype(np.uint32)
    total = n + 300
    words = np.zeros((total + 15) // 16 + 32, dtype=np.uint32)
    pos = n + np.arange(300, dtype=np.uint64)
    np.bitwise_or.at(words, (pos >> np.uint64(4)).astype(np.int64),
                     tail << ((pos & np.uint64(15)) * np.uint64(2)).astype(np.uint32))
    off = np.array([0, n, total], dtype=np.uint64)
    return words, off, tail


@pytest.mark.parametrize("k,m,mode", [(31, 21, M.MODE_SINGLE_PASS), (31, 21, M.MODE_TWO_PASS),
                                      (63, 23, M.MODE_SINGLE_PASS)])
def test_count_past_two_to_the_32(k, m, mode):
    n = (1 << 32) + 99
    words, off, tail = _batch(n)
    print(f"k={k} m={m} mode={mode}: poly-A fragment of {n} bases", flush=True)
    b = M.Batch.from_packed(words, off)
    try:
        # threshold 200: the wrapped low 32 bits (n - k + 1 - 2^32 <= 69)
        # alone would drop the k-mer
        b.count(M.CounterConfig(k=k, minimizer=m, threshold=200), mode=mode)
        got = b.download()
        b.count(M.CounterConfig(k=k, minimizer=m, threshold=0), mode=mode)
        got0 = b.download()
    finally:
        b.close()
    poly_a = O.encode_canonical("A" * k)
    want = n - k + 1
    assert want - (1 << 32) < 200
    keys = (got["hi"].astype(object) << 64) | got["lo"].astype(object)
    assert list(keys) == [poly_a] and int(got["count"][0]) == want
    # threshold 0: the poly-A k-mer with its full count, plus the random
    # fragment's k-mers exactly as the oracle counts them
    codes = np.concatenate([np.zeros(k + 5, dtype=np.uint8), tail.astype(np.uint8)])
    # the oracle sees the poly-A fragment shortened (same k-mer, smaller count)
    ref = O.count_bruteforce(codes, np.array([0, k + 5, k + 5 + 300], dtype=np.uint64), k, 0)
    ref_keys = (ref["hi"].astype(object) << 64) | ref["lo"].astype(object)
    keys0 = (got0["hi"].astype(object) << 64) | got0["lo"].astype(object)
    assert list(keys0) == list(ref_keys)
    for key, c, rk, rc in zip(keys0, got0["count"], ref_keys, ref["count"]):
        if key == poly_a:
            assert int(c) == want + (int(rc) - 6)  # the oracle's 6 poly-A windows are the big fragment's
        else:
            assert int(c) == int(rc)
