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
