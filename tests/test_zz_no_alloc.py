"""The allocation-free minimizer scan (mczi_minimizer_run_hashes_noalloc)
against the allocating one (mczi_minimizer_run_hashes), bit for bit, on two
calls in a row, the second with a smaller input (MI355X only)."""
import ctypes as C

import numpy as np
import pytest

import mczi_amd as M
import pyoracle as O

pytestmark = pytest.mark.gpu


def _pack(codes: np.ndarray) -> np.ndarray:
    n = codes.size
    b = np.zeros((n + 3) // 4, dtype=np.uint8)
    for j in range(4):
        b[: (n - j + 3) // 4] |= (codes[j::4] << (2 * j)).astype(np.uint8)
    return b


def _runs(fn, packed, n, k, m):
    nw = max(n - k + 1, 1)
    h = np.zeros(nw, dtype=np.uint32)
    st = np.zeros(nw, dtype=np.uint32)
    r = C.c_size_t()
    M._check(fn(packed.ctypes.data, n, k, m, h.ctypes.data, st.ctypes.data, C.byref(r)))
    return h[: r.value].copy(), st[: r.value].copy()


@pytest.mark.parametrize("k,m", [(31, 21), (63, 31), (21, 11), (5, 3)])
def test_noalloc_equals_allocating(k, m):
    print(f"noalloc run hashes k={k} m={m}", flush=True)
    rng = np.random.default_rng(k)
    for n in (300_000, 7_000):  # the second call smaller than the first
        codes = rng.integers(0, 4, size=n, dtype=np.uint8)
        codes[1000:4000] = 0  # a homopolymer run
        packed = _pack(codes)
        a = _runs(M._lib.mczi_minimizer_run_hashes, packed, n, k, m)
        b = _runs(M._lib.mczi_minimizer_run_hashes_noalloc, packed, n, k, m)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), n
        if n < 10_000:
            oh, ost = O.minimizer_runs(codes, k, m)
            assert np.array_equal(b[0], oh) and np.array_equal(b[1], ost)
