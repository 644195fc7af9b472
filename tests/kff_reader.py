"""Independent reader of KFF v1 files, written from the published format
(Dufresne et al., "The K-mer File Format: a standardized and compact disk
representation of sets of k-mers", Bioinformatics 2022, and its
specification): it is test infrastructure that checks the bytes the writers
produce without sharing any code with them.

Layout read here:
- header: "KFF", version major, minor, encoding byte (2 bits per nucleotide
  for A, C, G, T, most significant pair first), uniqueness and canonicity
  flags, then a free block (u32 size + bytes);
- sections, each opened by one byte: 'v' = variables (u64 count, then
  NUL-terminated name + u64 value per variable), 'r' = raw k-mer blocks
  (u64 block count; per block: the k-mer count on the bytes needed for
  `max`, omitted when max = 1, the ceil((n + k - 1) / 4) sequence bytes with
  the padding in the high bits of the first byte, then n x data_size data
  bytes);
- the closing "KFF" signature.
All integers are big-endian.
"""


def _be(b: bytes) -> int:
    return int.from_bytes(b, "big")


class KffFile:
    def __init__(self, data: bytes, decode_records: bool = True):
        if data[:3] != b"KFF" or data[-3:] != b"KFF":
            raise ValueError("missing KFF signature")
        self.version = (data[3], data[4])
        if self.version != (1, 0):
            raise ValueError(f"unsupported version {self.version}")
        enc = data[5]
        # code -> nucleotide: the pair at bits 7-6 is A's code, ... 1-0 T's
        self.decode = {(enc >> (6 - 2 * i)) & 3: "ACGT"[i] for i in range(4)}
        if len(self.decode) != 4:
            raise ValueError("encoding byte does not give four distinct codes")
        self.unique, self.canonical = data[6], data[7]
        n_free = _be(data[8:12])
        self.free_block = data[12:12 + n_free]
        pos = 12 + n_free
        self.variables = {}
        self.records = []  # (k-mer string, data value)
        self.raw = []      # (offset, blocks) of each raw section
        self._decode_records = decode_records
        end = len(data) - 3
        while pos < end:
            kind = data[pos:pos + 1]
            pos += 1
            if kind == b"v":
                n = _be(data[pos:pos + 8])
                pos += 8
                for _ in range(n):
                    z = data.index(b"\0", pos)
                    name = data[pos:z].decode()
                    self.variables[name] = _be(data[z + 1:z + 9])
                    pos = z + 9
            elif kind == b"r":
                pos = self._raw(data, pos)
            else:
                raise ValueError(f"unknown section {kind!r} at byte {pos - 1}")
        if pos != end:
            raise ValueError("sections overrun the closing signature")

    def _raw(self, data: bytes, pos: int) -> int:
        k = self.variables["k"]
        mx = self.variables["max"]
        dsz = self.variables["data_size"]
        nbytes = 0 if mx == 1 else (mx.bit_length() + 7) // 8
        blocks = _be(data[pos:pos + 8])
        pos += 8
        self.raw.append((pos, blocks))
        if not self._decode_records and mx == 1:
            return pos + blocks * ((k + 3) // 4 + dsz)
        for _ in range(blocks):
            n = 1
            if nbytes:
                n = _be(data[pos:pos + nbytes])
                pos += nbytes
            L = n + k - 1
            sb = (L + 3) // 4
            v = _be(data[pos:pos + sb])
            pos += sb
            if v >> (2 * L):
                raise ValueError("nonzero padding bits in a sequence")
            seq = "".join(self.decode[(v >> (2 * (L - 1 - i))) & 3] for i in range(L))
            for j in range(n):
                self.records.append((seq[j:j + k], _be(data[pos:pos + dsz])))
                pos += dsz
        return pos


def kmers_u64(f: KffFile, data: bytes):
    """Vectorised read of the raw sections of a file with max = 1 and k <= 32:
    (k-mers as 2-bit values with A0 C1 G2 T3, first base most significant;
    data values), in file order."""
    import numpy as np

    k, dsz = f.variables["k"], f.variables["data_size"]
    assert f.variables["max"] == 1 and k <= 32
    sb = (k + 3) // 4
    lut = np.zeros(4, dtype=np.uint64)
    for code, nt in f.decode.items():
        lut[code] = "ACGT".index(nt)
    keys, vals = [], []
    for pos, blocks in f.raw:
        rec = np.frombuffer(data, dtype=np.uint8, count=blocks * (sb + dsz), offset=pos).reshape(blocks, sb + dsz)
        v = np.zeros(blocks, dtype=np.uint64)
        for j in range(sb):
            v = (v << np.uint64(8)) | rec[:, j].astype(np.uint64)
        if np.any(v >> np.uint64(2 * k)):
            raise ValueError("nonzero padding bits in a sequence")
        out = np.zeros(blocks, dtype=np.uint64)
        for i in range(k):
            out |= lut[(v >> np.uint64(2 * i)) & np.uint64(3)] << np.uint64(2 * i)
        c = np.zeros(blocks, dtype=np.uint64)
        for j in range(dsz):
            c = (c << np.uint64(8)) | rec[:, sb + j].astype(np.uint64)
        keys.append(out)
        vals.append(c)
    return np.concatenate(keys), np.concatenate(vals)
