"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, never by the product package.

Fragments are passed as (codes, offsets): ``codes`` is a uint8 numpy array of
packed-seq base codes (A0 C1 T2 G3, one per byte), ``offsets`` a uint64 array
of length nfrag+1. Results are returned as numpy structured arrays with fields
``lo``, ``hi``, ``count`` (uint64), sorted by (hi, lo) where the oracle sorts.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

COUNTED = np.dtype([("lo", "<u8"), ("hi", "<u8"), ("count", "<u8")])


class _Counted(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64), ("count", C.c_uint64)]


class _Frags(C.Structure):
    _fields_ = [
        ("codes", C.c_void_p),
        ("off", C.c_void_p),
        ("nfrag", C.c_size_t),
        ("dataset", C.c_void_p),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        L = _lib
        L.or_nthash32.restype = C.c_uint32
        L.or_nthash32.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_spread_hash_u32.restype = C.c_uint32
        L.or_spread_hash_u32.argtypes = [C.c_uint32]
        L.or_spread_hash_u64.restype = C.c_uint64
        L.or_spread_hash_u64.argtypes = [C.c_uint64]
        L.or_window_hashes.restype = C.c_size_t
        L.or_window_hashes.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p]
        L.or_minimizer_runs.restype = C.c_size_t
        L.or_minimizer_runs.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.or_minimizer_positions.restype = C.c_size_t
        L.or_minimizer_positions.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.or_count_bruteforce.argtypes = [C.POINTER(_Frags), C.c_int, C.c_uint64, C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)]
        L.or_count_two_pass.argtypes = [C.POINTER(_Frags), C.c_int, C.c_int, C.c_uint64, C.c_size_t,
                                        C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_count_two_pass_mt.argtypes = [C.POINTER(_Frags), C.c_int, C.c_int, C.c_uint64, C.c_size_t, C.c_int,
                                           C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)]
        L.or_kff_u8_kmc.argtypes = [C.POINTER(_Frags), C.c_int, C.c_int, C.c_uint64, C.c_size_t,
                                    C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t), C.POINTER(C.c_uint64)]
        L.or_count_datasets.argtypes = [C.POINTER(_Frags), C.c_size_t, C.c_int, C.c_int, C.c_uint64, C.c_size_t,
                                        C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t), C.POINTER(C.c_uint64)]
        L.or_count_datasets_mt.argtypes = [C.POINTER(_Frags), C.c_size_t, C.c_int, C.c_int, C.c_uint64,
                                           C.c_size_t, C.c_int, C.POINTER(C.POINTER(_Counted)),
                                           C.POINTER(C.c_size_t), C.POINTER(C.c_uint64)]
        L.or_count_datasets_bruteforce.argtypes = [C.POINTER(_Frags), C.c_size_t, C.c_int, C.c_uint64,
                                                   C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)]
        L.or_zi_absent.argtypes = [C.POINTER(_Frags), C.POINTER(_Frags), C.c_int,
                                   C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)]
        L.or_spot_counts128.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_size_t, C.c_void_p, C.c_int]
        L.or_zi_lookup_timed.argtypes = [C.POINTER(_Frags), C.POINTER(_Frags), C.c_int, C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_write_simplitigs.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_int, C.c_size_t, C.POINTER(C.c_size_t)]
        L.or_write_kmer_fasta.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_int]
        L.or_write_kff.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int]
        L.or_parse_fastx.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.POINTER(C.c_uint8)),
                                     C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.c_size_t)]
        L.or_free.argtypes = [C.c_void_p]
        L.or_spot_counts.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_size_t, C.c_void_p, C.c_void_p, C.c_int]
    return _lib


# ---------------------------------------------------------------------------
_ASCII2CODE = np.full(256, 255, dtype=np.uint8)
for _ch, _c in (("A", 0), ("C", 1), ("T", 2), ("G", 3)):
    _ASCII2CODE[ord(_ch)] = _c
    _ASCII2CODE[ord(_ch.lower())] = _c


def codes_of(seq: str | bytes) -> np.ndarray:
    b = np.frombuffer(seq.encode() if isinstance(seq, str) else seq, dtype=np.uint8)
    c = _ASCII2CODE[b]
    assert (c != 255).all(), "only ACGT"
    return c.copy()


def frags_of(seqs, datasets=None):
    """(codes, offsets, datasets) from a list of ACGT strings."""
    cs = [codes_of(s) for s in seqs]
    off = np.zeros(len(cs) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(c) for c in cs]) if cs else []
    codes = np.concatenate(cs) if cs else np.zeros(0, dtype=np.uint8)
    ds = None if datasets is None else np.asarray(datasets, dtype=np.uint32)
    return codes, off, ds


def _frags(codes, off, ds=None) -> tuple[_Frags, tuple]:
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    keep = [codes, off]
    f = _Frags()
    f.codes = codes.ctypes.data if codes.size else None
    f.off = off.ctypes.data
    f.nfrag = off.size - 1
    if ds is not None:
        ds = np.ascontiguousarray(ds, dtype=np.uint32)
        keep.append(ds)
        f.dataset = ds.ctypes.data
    else:
        f.dataset = None
    return f, tuple(keep)


def _take_counted(ptr, n) -> np.ndarray:
    out = np.zeros(n, dtype=COUNTED)
    if n:
        C.memmove(out.ctypes.data, ptr, n * COUNTED.itemsize)
    lib().or_free(C.cast(ptr, C.c_void_p))
    return out


def nthash32(codes: np.ndarray, m: int, canonical: bool = True) -> int:
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    return lib().or_nthash32(codes.ctypes.data, m, int(canonical))


def spread_hash_u32(v: int) -> int:
    return lib().or_spread_hash_u32(v)


def spread_hash_u64(v: int) -> int:
    return lib().or_spread_hash_u64(v)


def window_hashes(codes, k, m) -> np.ndarray:
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    n = codes.size
    out = np.zeros(max(n - k + 1, 0), dtype=np.uint32)
    nw = lib().or_window_hashes(codes.ctypes.data, n, k, m, out.ctypes.data)
    return out[:nw]


def minimizer_runs(codes, k, m):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    n = codes.size
    h = np.zeros(max(n, 1), dtype=np.uint32)
    s = np.zeros(max(n, 1), dtype=np.uint32)
    r = lib().or_minimizer_runs(codes.ctypes.data, n, k, m, h.ctypes.data, s.ctypes.data)
    return h[:r], s[:r]


def minimizer_positions(codes, k_hash, w, canonical=True) -> list[int]:
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros(max(codes.size, 1), dtype=np.uint32)
    r = lib().or_minimizer_positions(codes.ctypes.data, codes.size, k_hash, w, int(canonical), out.ctypes.data)
    return out[:r].tolist()


def count_bruteforce(codes, off, k, x) -> np.ndarray:
    f, keep = _frags(codes, off)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    rc = lib().or_count_bruteforce(C.byref(f), k, x, C.byref(p), C.byref(n))
    assert rc == 0
    return _take_counted(p, n.value)


def count_two_pass(codes, off, k, m, x, partitions):
    f, keep = _frags(codes, off)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    pb = C.c_uint64()
    kept = C.c_uint64()
    rc = lib().or_count_two_pass(C.byref(f), k, m, x, partitions, C.byref(p), C.byref(n), C.byref(pb), C.byref(kept))
    if rc != 0:
        raise ValueError("oracle rejected configuration")
    return _take_counted(p, n.value), pb.value, kept.value


def count_two_pass_mt(codes, off, k, m, x, partitions, threads):
    f, keep = _frags(codes, off)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    rc = lib().or_count_two_pass_mt(C.byref(f), k, m, x, partitions, threads, C.byref(p), C.byref(n))
    if rc != 0:
        raise ValueError("oracle rejected configuration")
    return _take_counted(p, n.value)


def kff_u8_kmc(codes, off, k, m, x, partitions) -> tuple[bytes, int]:
    f, keep = _frags(codes, off)
    p = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    rec = C.c_uint64()
    rc = lib().or_kff_u8_kmc(C.byref(f), k, m, x, partitions, C.byref(p), C.byref(n), C.byref(rec))
    if rc != 0:
        raise ValueError("oracle rejected configuration")
    data = C.string_at(p, n.value)
    lib().or_free(C.cast(p, C.c_void_p))
    return data, rec.value


def count_datasets(codes, off, ds, n_datasets, k, m, x, partitions):
    f, keep = _frags(codes, off, ds)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    kept = C.c_uint64()
    rc = lib().or_count_datasets(C.byref(f), n_datasets, k, m, x, partitions, C.byref(p), C.byref(n), C.byref(kept))
    if rc != 0:
        raise ValueError("oracle rejected configuration")
    return _take_counted(p, n.value), kept.value


def count_datasets_mt(codes, off, ds, n_datasets, k, m, x, partitions, threads):
    f, keep = _frags(codes, off, ds)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    kept = C.c_uint64()
    rc = lib().or_count_datasets_mt(C.byref(f), n_datasets, k, m, x, partitions, threads, C.byref(p), C.byref(n),
                                    C.byref(kept))
    if rc != 0:
        raise ValueError("oracle rejected configuration")
    return _take_counted(p, n.value), kept.value


def count_datasets_bruteforce(codes, off, ds, n_datasets, k, x) -> np.ndarray:
    f, keep = _frags(codes, off, ds)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    rc = lib().or_count_datasets_bruteforce(C.byref(f), n_datasets, k, x, C.byref(p), C.byref(n))
    assert rc == 0
    return _take_counted(p, n.value)


def zi_absent(icodes, ioff, qcodes, qoff, k) -> np.ndarray:
    fi, k1 = _frags(icodes, ioff)
    fq, k2 = _frags(qcodes, qoff)
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    rc = lib().or_zi_absent(C.byref(fi), C.byref(fq), k, C.byref(p), C.byref(n))
    assert rc == 0
    return _take_counted(p, n.value)


def zi_lookup_timed(icodes, ioff, qcodes, qoff, k):
    """(lookups, unique absent, index build s, query s) of ZI's one-thread
    query loop."""
    fi, k1 = _frags(icodes, ioff)
    fq, k2 = _frags(qcodes, qoff)
    nl, na = C.c_uint64(), C.c_uint64()
    b, q = C.c_double(), C.c_double()
    assert lib().or_zi_lookup_timed(C.byref(fi), C.byref(fq), k, C.byref(nl), C.byref(na), C.byref(b), C.byref(q)) == 0
    return nl.value, na.value, b.value, q.value


def write_simplitigs(path, kmers: np.ndarray, k, first_index=0) -> int:
    kmers = np.ascontiguousarray(kmers, dtype=COUNTED)
    last = C.c_size_t()
    rc = lib().or_write_simplitigs(path.encode(), kmers.ctypes.data if kmers.size else None, kmers.size, k,
                                   first_index, C.byref(last))
    assert rc == 0
    return last.value


def write_kmer_fasta(path, kmers: np.ndarray, k) -> None:
    kmers = np.ascontiguousarray(kmers, dtype=COUNTED)
    assert lib().or_write_kmer_fasta(path.encode(), kmers.ctypes.data if kmers.size else None, kmers.size, k) == 0


def write_kff(path, kmers: np.ndarray, k, m, x, count_size=8, ordered=True) -> None:
    kmers = np.ascontiguousarray(kmers, dtype=COUNTED)
    assert lib().or_write_kff(path.encode(), kmers.ctypes.data if kmers.size else None, kmers.size, k, m, x,
                              count_size, int(ordered)) == 0


def parse_fastx(path, min_len):
    cp = C.POINTER(C.c_uint8)()
    op = C.POINTER(C.c_uint64)()
    n = C.c_size_t()
    rc = lib().or_parse_fastx(path.encode(), min_len, C.byref(cp), C.byref(op), C.byref(n))
    if rc != 0:
        raise OSError(f"cannot read {path}")
    off = np.zeros(n.value + 1, dtype=np.uint64)
    C.memmove(off.ctypes.data, op, off.nbytes)
    codes = np.zeros(int(off[-1]), dtype=np.uint8)
    if codes.size:
        C.memmove(codes.ctypes.data, cp, codes.size)
    lib().or_free(C.cast(cp, C.c_void_p))
    lib().or_free(C.cast(op, C.c_void_p))
    return codes, off


# ---------------------------------------------------------------------------
_MC = "ACGT"


def decode(lo: int, hi: int, k: int) -> str:
    v = (int(hi) << 64) | int(lo)
    return "".join(_MC[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))


def encode_canonical(s: str) -> int:
    k = len(s)
    f = 0
    r = 0
    for ch in s:
        b = "ACGT".index(ch)
        f = (f << 2) | b
    for ch in reversed(s):
        b = 3 - "ACGT".index(ch)
        r = (r << 2) | b
    return min(f, r)


# ---------------------------------------------------------------------------
# Full-size spot checks (reads packed 16 bases per u32 word)
def spot_counts(words, frag_off, k, query, dataset=None, threads=16):
    """Exact count and dataset-presence mask of each query k-mer (canonical
    MC encoding, k <= 32) over packed reads (oracle.c or_spot_counts)."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    off = np.ascontiguousarray(frag_off, dtype=np.uint64)
    q = np.ascontiguousarray(query, dtype=np.uint64)
    ds = None if dataset is None else np.ascontiguousarray(dataset, dtype=np.uint32)
    counts = np.zeros(q.size, dtype=np.uint64)
    masks = np.zeros(q.size, dtype=np.uint32)
    rc = lib().or_spot_counts(words.ctypes.data, off.ctypes.data, off.size - 1,
                              ds.ctypes.data if ds is not None else None, k, q.ctypes.data, q.size,
                              counts.ctypes.data, masks.ctypes.data, threads)
    if rc:
        raise ValueError("or_spot_counts failed")
    return counts, masks


def spot_counts128(words, frag_off, k, qlo, qhi, threads=16):
    """Exact count of each 128-bit canonical query k-mer (lo, hi words;
    33 <= k <= 64) over packed reads."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    off = np.ascontiguousarray(frag_off, dtype=np.uint64)
    lo = np.ascontiguousarray(qlo, dtype=np.uint64)
    hi = np.ascontiguousarray(qhi, dtype=np.uint64)
    counts = np.zeros(lo.size, dtype=np.uint64)
    rc = lib().or_spot_counts128(words.ctypes.data, off.ctypes.data, off.size - 1, k, lo.ctypes.data, hi.ctypes.data,
                                 lo.size, counts.ctypes.data, threads)
    if rc:
        raise ValueError("or_spot_counts128 failed")
    return counts


def kmers_at128(words, pos, k):
    """Canonical MC encoding of the k-mer (k <= 64) at each base position, as
    (lo, hi) uint64 arrays."""
    words = np.asarray(words, dtype=np.uint32)
    pos = np.asarray(pos, dtype=np.uint64)
    n = pos.size
    fh, fl, rh, rl = (np.zeros(n, dtype=np.uint64) for _ in range(4))
    for i in range(k):
        p = pos + np.uint64(i)
        c = (words[(p >> np.uint64(4)).astype(np.int64)] >> ((p & np.uint64(15)) * np.uint64(2)).astype(np.uint32)) & 3
        b = (c ^ (c >> 1)).astype(np.uint64)
        # forward: shift the 128-bit value left by 2 and add b
        fh = (fh << np.uint64(2)) | (fl >> np.uint64(62))
        fl = (fl << np.uint64(2)) | b
        # reverse complement: base i's complement at bit position 2i
        sh = 2 * i
        cb = b ^ np.uint64(3)
        if sh < 64:
            rl = rl | (cb << np.uint64(sh))
        else:
            rh = rh | (cb << np.uint64(sh - 64))
    if 2 * k < 128:
        fh = fh & np.uint64((1 << (2 * k - 64)) - 1) if k > 32 else np.zeros(n, dtype=np.uint64)
    take_r = (rh < fh) | ((rh == fh) & (rl < fl))
    return np.where(take_r, rl, fl), np.where(take_r, rh, fh)


def kmers_at(words, pos, k):
    """Canonical MC encoding of the k-mer starting at each base position
    (k <= 32) of packed reads."""
    words = np.asarray(words, dtype=np.uint32)
    pos = np.asarray(pos, dtype=np.uint64)
    fwd = np.zeros(pos.size, dtype=np.uint64)
    rev = np.zeros(pos.size, dtype=np.uint64)
    for i in range(k):
        p = pos + np.uint64(i)
        c = (words[(p >> np.uint64(4)).astype(np.int64)] >> ((p & np.uint64(15)) * np.uint64(2)).astype(np.uint32)) & 3
        b = (c ^ (c >> 1)).astype(np.uint64)
        fwd = (fwd << np.uint64(2)) | b
        rev = rev | ((b ^ np.uint64(3)) << np.uint64(2 * i))
    return np.minimum(fwd, rev)
