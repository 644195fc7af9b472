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
