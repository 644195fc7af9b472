"""Shared helpers for the parity tests (inputs, comparisons)."""
import numpy as np

import pyoracle as O

REV = {"A": "T", "C": "G", "G": "C", "T": "A"}


def revcomp(s: str) -> str:
    return "".join(REV[c] for c in reversed(s))


def enc(s: str) -> int:
    """Forward MC encoding (A0 C1 G2 T3, first base most significant)."""
    v = 0
    for ch in s:
        v = (v << 2) | "ACGT".index(ch)
    return v


def as_pairs(arr, k):
    return [(O.decode(int(r["lo"]), int(r["hi"]), k), int(r["count"])) for r in arr]


def random_seqs(rng, n, lo, hi, alphabet="ACGT"):
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi + 1))
        out.append("".join(rng.choice(list(alphabet), size=L)))
    return out


def write_fasta(path, seqs, name="r"):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(f">{name}{i}\n{s}\n")


def same_counted(a, b) -> bool:
    return a.shape == b.shape and bool(np.all(a["lo"] == b["lo"]) and np.all(a["hi"] == b["hi"])
                                       and np.all(a["count"] == b["count"]))


def canonical_kmers_np(codes, k):
    """Canonical k-mers (A0 C1 G2 T3) of a code array as (hi, lo) uint64
    arrays: the value is hi << 64 | lo. Vectorised over positions."""
    codes = np.asarray(codes, dtype=np.uint64)
    n = codes.size - k + 1
    if n <= 0:
        z = np.zeros(0, dtype=np.uint64)
        return z, z

    def val(get, j0, j1):
        v = np.zeros(n, dtype=np.uint64)
        for j in range(j0, j1):
            v = (v << np.uint64(2)) | get(j)
        return v

    fwd = lambda j: codes[j:j + n]
    rev = lambda j: np.uint64(3) - codes[k - 1 - j:k - 1 - j + n]
    split = max(0, k - 32)
    fh, fl = val(fwd, 0, split), val(fwd, split, k)
    rh, rl = val(rev, 0, split), val(rev, split, k)
    take_r = (rh < fh) | ((rh == fh) & (rl < fl))
    return np.where(take_r, rh, fh), np.where(take_r, rl, fl)


def pack_words(codes) -> np.ndarray:
    """Packed-seq codes (A0 C1 T2 G3) as 16 bases per little-endian u32 word,
    with 32 zero words of slack."""
    codes = np.asarray(codes, dtype=np.uint32)
    n = codes.size
    w = np.zeros((n + 15) // 16 + 32, dtype=np.uint32)
    for j in range(16):
        part = codes[j::16]
        w[: part.size] |= part << np.uint32(2 * j)
    return w


def with_empty_fragments(rng, seqs, runs=6, run_len=40):
    """Fragment offsets of seqs (concatenated) with runs of empty fragments
    inserted at random fragment boundaries; returns (codes, off, src) where
    src[i] is the index in seqs of fragment i (-1 for an empty one)."""
    codes = np.concatenate([O.codes_of(s) for s in seqs]).astype(np.uint8)
    lens = [len(s) for s in seqs]
    at = set(int(x) for x in rng.integers(0, len(seqs) + 1, size=runs))
    off, src, pos = [0], [], 0
    for i in range(len(seqs) + 1):
        if i in at:
            for _ in range(run_len):
                off.append(pos)
                src.append(-1)
        if i < len(seqs):
            pos += lens[i]
            off.append(pos)
            src.append(i)
    return codes, np.array(off, dtype=np.uint64), np.array(src, dtype=np.int64)
