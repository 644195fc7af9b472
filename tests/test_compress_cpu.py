"""Compressed FASTX inputs (gzip, xz, zstd; by magic bytes) and compressed
outputs (.gz/.gzip, .xz, .zst/.zstd by extension, src/lib.rs:1383-1395) on
the host: a compressed input packs to the same fragments as its plain text,
and a compressed output decompresses to the plain writer's bytes."""
import ctypes as C
import gzip
import lzma

import numpy as np
import pytest

import mczi_amd as M

_zstd = C.CDLL("libzstd.so.1")
_zstd.ZSTD_compressBound.restype = C.c_size_t
_zstd.ZSTD_compressBound.argtypes = [C.c_size_t]
_zstd.ZSTD_compress.restype = C.c_size_t
_zstd.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
_zstd.ZSTD_isError.restype = C.c_uint
_zstd.ZSTD_isError.argtypes = [C.c_size_t]


def zstd_compress(b: bytes, level=3) -> bytes:
    cap = _zstd.ZSTD_compressBound(len(b))
    out = C.create_string_buffer(cap)
    n = _zstd.ZSTD_compress(out, cap, b, len(b), level)
    assert not _zstd.ZSTD_isError(n)
    return out.raw[:n]


class _ZBuf(C.Structure):
    _fields_ = [("p", C.c_void_p), ("size", C.c_size_t), ("pos", C.c_size_t)]


_zstd.ZSTD_createDStream.restype = C.c_void_p
_zstd.ZSTD_freeDStream.argtypes = [C.c_void_p]
_zstd.ZSTD_decompressStream.restype = C.c_size_t
_zstd.ZSTD_decompressStream.argtypes = [C.c_void_p, C.POINTER(_ZBuf), C.POINTER(_ZBuf)]


def zstd_decompress_frames(b: bytes) -> bytes:
    """Streaming decompression of one or more frames (content size unknown)."""
    ds = _zstd.ZSTD_createDStream()
    src = C.create_string_buffer(b, len(b))
    ib = _ZBuf(C.cast(src, C.c_void_p), len(b), 0)
    chunk = C.create_string_buffer(1 << 20)
    out = []
    try:
        while True:
            ob = _ZBuf(C.cast(chunk, C.c_void_p), len(chunk), 0)
            rc = _zstd.ZSTD_decompressStream(ds, C.byref(ob), C.byref(ib))
            assert not _zstd.ZSTD_isError(rc)
            out.append(chunk.raw[:ob.pos])
            if ib.pos >= ib.size and ob.pos < ob.size:
                break
    finally:
        _zstd.ZSTD_freeDStream(ds)
    return b"".join(out)


def _fastx(rng, n, fastq=False):
    out = []
    for r in range(n):
        L = int(rng.integers(1, 300))
        s = "".join(rng.choice(list("ACGTACGTNacgt"), size=L))
        if fastq:
            out.append(f"@r{r}\n{s}\n+\n{'I' * L}\n")
        else:
            out.append(f">r{r} d\n" + "\n".join(s[i:i + 60] for i in range(0, L, 60)) + "\n")
    return "".join(out).encode()


def _compressors():
    return {
        "gz": lambda b: gzip.compress(b, 1),
        "gz_multi": lambda b: gzip.compress(b[: len(b) // 2], 6) + gzip.compress(b[len(b) // 2:], 1),
        "xz": lambda b: lzma.compress(b),
        "xz_multi": lambda b: lzma.compress(b[: len(b) // 3]) + lzma.compress(b[len(b) // 3:]),
        "zst": lambda b: zstd_compress(b),
        "zst_multi": lambda b: zstd_compress(b[: len(b) // 2]) + zstd_compress(b[len(b) // 2:], 1),
    }


@pytest.mark.parametrize("kind", sorted(_compressors()))
@pytest.mark.parametrize("fastq", [False, True])
def test_compressed_input_packs_like_plain(tmp_path, kind, fastq):
    rng = np.random.default_rng(sorted(_compressors()).index(kind) * 2 + int(fastq))
    text = _fastx(rng, 400, fastq)
    plain = tmp_path / "r.fx"
    plain.write_bytes(text)
    # no extension: the format comes from the magic bytes
    z = tmp_path / "r.bin"
    z.write_bytes(_compressors()[kind](text))
    w0, o0 = M.pack_fastx(plain)
    w1, o1 = M.pack_fastx(z)
    assert np.array_equal(o0, o1) and np.array_equal(w0, w1)
    assert o0[-1] > 10000


def test_compressed_invalid_fastx_is_skipped(tmp_path):
    """A valid gzip stream of non-FASTX text is skipped like the plain file
    (partitioned_counter_skips_invalid_compressed_fastx_inputs, src/lib.rs)."""
    p = tmp_path / "not_fastx.txt.gz"
    p.write_bytes(gzip.compress(b"not a FASTA or FASTQ file\n"))
    w, off = M.pack_fastx(p)
    assert off[-1] == 0


@pytest.mark.parametrize("kind", ["gz", "xz", "zst"])
def test_corrupt_or_truncated_input_is_an_error(tmp_path, kind):
    text = _fastx(np.random.default_rng(3), 200)
    z = _compressors()[kind](text)
    p = tmp_path / "t.bin"
    p.write_bytes(z[: len(z) // 2])
    with pytest.raises(M.MczaError):
        M.pack_fastx(p)


def _kmers(rng, n, k):
    lo = np.unique(rng.integers(0, 1 << (2 * k), size=n, dtype=np.uint64))
    a = np.zeros(lo.size, dtype=M.COUNTED)
    a["lo"] = lo
    a["count"] = rng.integers(1, 1000, size=lo.size)
    return a


@pytest.mark.parametrize("ext,dec", [("gz", gzip.decompress), ("GZIP", gzip.decompress), ("xz", lzma.decompress),
                                     ("zst", zstd_decompress_frames), ("zstd", zstd_decompress_frames)])
def test_compressed_outputs_decompress_to_plain_bytes(tmp_path, ext, dec):
    rng = np.random.default_rng(9)
    k = 21
    a = _kmers(rng, 30000, k)
    M.write_fasta_path(tmp_path / "o.fa", a, k)
    M.write_fasta_path(tmp_path / f"o.fa.{ext}", a, k)
    plain = (tmp_path / "o.fa").read_bytes()
    z = (tmp_path / f"o.fa.{ext}").read_bytes()
    assert z != plain and len(z) < len(plain)
    assert dec(z) == plain
    cfg = M.CounterConfig(k=k, minimizer=11, threshold=1)
    M._check(M._lib.mczi_write_kff_path(str(tmp_path / "o.kff").encode(), a.ctypes.data, a.size,
                                        C.byref(cfg._c())))
    M._check(M._lib.mczi_write_kff_path(str(tmp_path / f"o.kff.{ext}").encode(), a.ctypes.data, a.size,
                                        C.byref(cfg._c())))
    assert dec((tmp_path / f"o.kff.{ext}").read_bytes()) == (tmp_path / "o.kff").read_bytes()


def test_empty_compressed_output(tmp_path):
    a = np.zeros(0, dtype=M.COUNTED)
    M.write_fasta_path(tmp_path / "e.fa.gz", a, 5)
    assert gzip.decompress((tmp_path / "e.fa.gz").read_bytes()) == b""
    M.write_fasta_path(tmp_path / "e.fa.xz", a, 5)
    assert lzma.decompress((tmp_path / "e.fa.xz").read_bytes()) == b""
