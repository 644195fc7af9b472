"""MI355X-native MCZI k-mer counting and subtraction path (Python host API).

A thin ctypes layer over ``libmczi.so`` (C ABI in ``include/mczi.h``). Names
and argument meanings follow the reference's ``lib.rs`` / ``zi.rs`` entry
points (``count_inputs``, ``count_datasets``, ``count_inputs_to_fasta_path``,
``count_inputs_to_kff_path``, ``write_fasta_path``, ``expand_fofns``,
``decode_kmer``, ...). Errors raise :class:`MczaError` carrying the library's
context chain. The library must be built (``make -C mczi_amd``); importing
this module without it raises ``ImportError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmczi.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run `make -C {_HERE}` (HIP extension required)")

_lib = C.CDLL(LIB_PATH)

DEFAULT_PARTITION_COUNT = 1024
MODE_TWO_PASS, MODE_SINGLE_PASS, MODE_DATASETS, MODE_DATASETS_TWO_PASS = 0, 1, 2, 3
LAUNCH_BATCHED, LAUNCH_PLAIN = 0, 1
ORDER_ENCODED, ORDER_KFF_U8, ORDER_NONE = 0, 1, 2

COUNTED = np.dtype([("lo", "<u8"), ("hi", "<u8"), ("count", "<u8")])


class MczaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class _Config(C.Structure):
    _fields_ = [("k", C.c_uint32), ("minimizer", C.c_uint32), ("threshold", C.c_uint64),
                ("partition_count", C.c_uint32)]


class _Counted(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64), ("count", C.c_uint64)]


class _FastaStats(C.Structure):
    _fields_ = [("selected_kmers", C.c_uint64), ("minimizers_above_threshold", C.c_uint64),
                ("total_seconds", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("n_bases", "n_windows", "n_pieces", "n_records", "n_distinct",
                                         "n_overflow", "n_out", "tile_windows", "partition_bytes")] + [
        ("internal_partitions", C.c_uint32), ("lds_slots", C.c_uint32)] + [
        (n, C.c_double) for n in ("ms_scan1", "ms_scan2", "ms_count", "ms_sort", "ms_total")] + [
        ("unique_minimizers", C.c_uint64), ("minimizers_retained", C.c_uint64), ("ms_pass1", C.c_double),
        ("partition_groups", C.c_uint64), ("scan_blocks_per_cu", C.c_uint64)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class _SynthSpec(C.Structure):
    _fields_ = [("genome_len", C.c_uint64), ("genome_seed", C.c_uint64), ("read_seed", C.c_uint64),
                ("n_reads", C.c_uint64), ("read_len", C.c_uint32), ("sub_rate", C.c_double),
                ("ins_rate", C.c_double), ("del_rate", C.c_double), ("long_bases", C.c_uint64),
                ("long_median", C.c_double), ("long_sigma", C.c_double), ("long_min", C.c_uint32),
                ("long_max", C.c_uint32), ("repeat_fraction", C.c_double), ("repeat_len", C.c_uint32),
                ("repeat_copies", C.c_uint32), ("n_datasets", C.c_uint32), ("snp_rate", C.c_double)]


class _SynthReads(C.Structure):
    _fields_ = [("words", C.POINTER(C.c_uint32)), ("n_words", C.c_uint64),
                ("frag_off", C.POINTER(C.c_uint64)), ("n_frag", C.c_uint64),
                ("dataset", C.POINTER(C.c_uint32)), ("n_bases", C.c_uint64)]


class McziStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("index_kmers", "query_kmers_scanned", "query_kmers_filtered",
                                         "query_kmers_not_filtered", "query_unique_kmers_not_filtered")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class _ZiResult(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("n_windows", C.c_uint64), ("n_absent", C.c_uint64),
                ("absent_lo", C.POINTER(C.c_uint64)), ("absent_hi", C.POINTER(C.c_uint64)),
                ("fasta", C.c_char_p), ("fasta_len", C.c_size_t),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


_P = C.c_void_p
_CP = C.POINTER(C.c_char_p)


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("mczi_last_error", C.c_char_p, [])
_sig("mczi_version", C.c_char_p, [])
_sig("mczi_set_threads", C.c_int, [C.c_uint32])
_sig("mczi_free", None, [_P])
_sig("mczi_validate_config", C.c_int, [C.POINTER(_Config), C.c_int])
_sig("mczi_expand_fofns", C.c_int, [_CP, C.c_size_t, C.POINTER(C.POINTER(C.c_char_p)), C.POINTER(C.c_size_t)])
_sig("mczi_free_strings", None, [C.POINTER(C.c_char_p), C.c_size_t])
for _n in ("mczi_count_inputs", "mczi_count_datasets"):
    _sig(_n, C.c_int, [_CP, C.c_size_t, C.POINTER(_Config), C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)])
for _n in ("mczi_count_inputs_to_fasta_path", "mczi_count_datasets_to_fasta_path",
           "mczi_count_inputs_to_kff_path", "mczi_count_datasets_to_kff_path"):
    _sig(_n, C.c_int, [_CP, C.c_size_t, C.POINTER(_Config), C.c_char_p])
for _n in ("mczi_count_inputs_to_kmer_fasta_path", "mczi_count_datasets_to_kmer_fasta_path"):
    _sig(_n, C.c_int, [_CP, C.c_size_t, C.POINTER(_Config), C.c_char_p, C.POINTER(_FastaStats)])
class Phase12Stats(C.Structure):
    _fields_ = [("phase1", C.c_double), ("phase2", C.c_double), ("total", C.c_double),
                ("unique_minimizer_hashes", C.c_uint64), ("partition_bytes", C.c_uint64)]


_sig("mczi_run_inputs_phase12_direct", C.c_int, [_CP, C.c_size_t, C.POINTER(_Config), C.POINTER(Phase12Stats)])
_sig("mczi_count_inputs_minimizer_phase_direct", C.c_int, [_CP, C.c_size_t, C.POINTER(_Config),
                                                           C.POINTER(C.c_uint64)])
_sig("mczi_write_fasta_path", C.c_int, [C.c_char_p, _P, C.c_size_t, C.c_uint32])
_sig("mczi_write_fasta_after_index", C.c_int, [C.c_char_p, _P, C.c_size_t, C.c_uint32, C.c_size_t,
                                               C.POINTER(C.c_size_t)])
_sig("mczi_write_kff_path", C.c_int, [C.c_char_p, _P, C.c_size_t, C.POINTER(_Config)])
_sig("mczi_simplitig_sequences", C.c_int, [_P, C.c_size_t, C.c_uint32, C.POINTER(C.POINTER(C.c_char_p)),
                                           C.POINTER(C.c_size_t)])
_sig("mczi_simplitig_sequences_parallel", C.c_int, [_P, C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32,
                                                    C.POINTER(C.POINTER(C.c_char_p)), C.POINTER(C.c_size_t)])
_sig("mczi_decode_kmer", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_char_p])
_sig("mczi_minimizer_run_hashes", C.c_int, [_P, C.c_uint64, C.c_uint32, C.c_uint32, _P, _P, C.POINTER(C.c_size_t)])
_sig("mczi_minimizer_run_hashes_noalloc", C.c_int, [_P, C.c_uint64, C.c_uint32, C.c_uint32, _P, _P,
                                                    C.POINTER(C.c_size_t)])
_sig("mczi_batch_from_files", C.c_int, [_CP, C.c_size_t, C.c_uint32, C.c_int, C.POINTER(_P)])
_sig("mczi_pack_fastx", C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.POINTER(C.c_uint32)),
                                  C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.c_size_t)])
_sig("mczi_batch_from_packed", C.c_int, [_P, _P, C.c_size_t, _P, C.POINTER(_P)])
_sig("mczi_batch_from_packed_inputs", C.c_int, [_P, _P, C.c_size_t, _P, _P, C.c_size_t, C.POINTER(_P)])
_WORDS_OUT = [C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.c_size_t)]
_sig("mczi_write_read_cache", C.c_int, [C.c_char_p, C.c_uint32, C.c_char_p])
_sig("mczi_read_read_cache", C.c_int, [C.c_char_p, C.c_uint32] + _WORDS_OUT)
_sig("mczi_write_superkmer_partitions", C.c_int, [_CP, C.c_size_t, C.POINTER(_Config), C.c_char_p])
_sig("mczi_read_superkmer_partitions", C.c_int, [_CP, C.c_size_t, C.c_uint32] + _WORDS_OUT)
_sig("mczi_count_superkmer_partitions", C.c_int, [_CP, C.c_size_t, C.POINTER(_Config), C.POINTER(C.POINTER(_Counted)),
                                                  C.POINTER(C.c_size_t)])
_sig("mczi_write_counted", C.c_int, [C.c_char_p, _P, C.c_size_t])
_sig("mczi_read_counted", C.c_int, [C.c_char_p, C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)])
_sig("mczi_batch_info", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)])
_sig("mczi_batch_free", None, [_P])
_sig("mczi_batch_count", C.c_int, [_P, C.POINTER(_Config), C.c_int, C.c_int, C.c_int, C.POINTER(RunStats)])
_sig("mczi_batch_autotune", C.c_int, [_P, C.POINTER(_Config), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)])
_sig("mczi_batch_download", C.c_int, [_P, C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t)])
_sig("mczi_copy_probe", C.c_int, [C.c_uint64, C.c_int, C.POINTER(C.c_double)])
_sig("mczi_device_info", C.c_int, [C.c_char_p, C.c_size_t])
_sig("mczi_zi_subtract", C.c_int, [C.c_char_p, _CP, C.c_size_t, C.c_uint32, C.c_uint32, C.c_char_p,
                                   C.c_uint32, C.c_uint32])
_sig("mczi_zi_absent", C.c_int, [C.c_char_p, _CP, C.c_size_t, C.c_uint32, C.POINTER(C.POINTER(_Counted)),
                                 C.POINTER(C.c_size_t)])
_sig("mczi_mczi_subtract", C.c_int, [_CP, C.c_size_t, C.c_int, _CP, C.c_size_t, C.POINTER(_Config), C.c_char_p,
                                     C.POINTER(McziStats)])
_sig("mczi_mczi_absent", C.c_int, [_CP, C.c_size_t, C.c_int, _CP, C.c_size_t, C.POINTER(_Config),
                                   C.POINTER(C.POINTER(_Counted)), C.POINTER(C.c_size_t), C.POINTER(McziStats)])
_sig("mczi_zi_server_create", C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(_P)])
_sig("mczi_zi_server_create_from_text", C.c_int, [C.c_char_p, C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint64,
                                                  C.c_uint32, C.POINTER(_P)])
_sig("mczi_zi_server_submit", C.c_int, [_P, C.c_char_p, C.c_size_t, C.c_int, C.POINTER(C.c_uint64)])
_sig("mczi_zi_server_submit_packed", C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, C.POINTER(C.c_uint64)])
_sig("mczi_zi_server_poll", C.c_int, [_P, C.c_uint64, C.c_int, C.POINTER(C.POINTER(_ZiResult))])
_sig("mczi_zi_result_free", None, [C.POINTER(_ZiResult)])
_sig("mczi_zi_server_stats", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double)])
_sig("mczi_zi_server_lookup_stats", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_double)])
_sig("mczi_zi_server_destroy", None, [_P])
_sig("mczi_synth_generate", C.c_int, [C.POINTER(_SynthSpec), C.POINTER(_SynthReads)])
_sig("mczi_synth_genome", C.c_int, [C.c_uint64, C.c_uint64, C.POINTER(C.POINTER(C.c_uint32))])
_sig("mczi_synth_reads_from", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_uint32, C.c_uint64,
                                        C.c_uint32, C.c_double, C.c_uint64, C.POINTER(_SynthReads)])
_sig("mczi_synth_write_fasta", C.c_int, [C.POINTER(_SynthReads), C.c_char_p, C.c_int64])
_sig("mczi_synth_free", None, [C.POINTER(_SynthReads)])

#: every symbol include/mczi.h declares
EXPORTED = [
    "mczi_last_error", "mczi_set_threads", "mczi_version", "mczi_free", "mczi_validate_config", "mczi_expand_fofns",
    "mczi_free_strings", "mczi_count_inputs", "mczi_count_datasets", "mczi_count_inputs_to_fasta_path",
    "mczi_count_datasets_to_fasta_path", "mczi_count_inputs_to_kff_path", "mczi_count_datasets_to_kff_path",
    "mczi_count_inputs_to_kmer_fasta_path", "mczi_count_datasets_to_kmer_fasta_path",
    "mczi_run_inputs_phase12_direct", "mczi_count_inputs_minimizer_phase_direct", "mczi_write_fasta_path",
    "mczi_write_fasta_after_index", "mczi_write_kff_path", "mczi_simplitig_sequences",
    "mczi_simplitig_sequences_parallel", "mczi_decode_kmer", "mczi_minimizer_run_hashes", "mczi_minimizer_run_hashes_noalloc",
    "mczi_batch_from_files", "mczi_pack_fastx", "mczi_batch_from_packed", "mczi_batch_from_packed_inputs",
    "mczi_write_read_cache",
    "mczi_read_read_cache", "mczi_write_superkmer_partitions", "mczi_read_superkmer_partitions",
    "mczi_count_superkmer_partitions", "mczi_write_counted", "mczi_read_counted", "mczi_batch_info", "mczi_batch_free", "mczi_batch_count",
    "mczi_batch_download", "mczi_batch_autotune", "mczi_copy_probe", "mczi_device_info", "mczi_zi_subtract", "mczi_zi_absent", "mczi_mczi_subtract", "mczi_mczi_absent", "mczi_zi_server_create",
    "mczi_zi_server_create_from_text", "mczi_zi_server_submit", "mczi_zi_server_submit_packed",
    "mczi_zi_server_poll", "mczi_zi_result_free",
    "mczi_zi_server_stats", "mczi_zi_server_lookup_stats", "mczi_zi_server_destroy", "mczi_synth_generate", "mczi_synth_genome",
    "mczi_synth_write_fasta", "mczi_synth_free", "mczi_synth_reads_from",
]


def _check(rc: int) -> None:
    if rc != 0:
        raise MczaError(rc, (_lib.mczi_last_error() or b"").decode())


def _paths(paths) -> tuple:
    enc = [os.fspath(p).encode() for p in paths]
    arr = (C.c_char_p * max(len(enc), 1))(*enc)
    return arr, len(enc)


def _take(ptr, n) -> np.ndarray:
    """The library's malloc'd record array as a numpy array without a copy;
    it is released with mczi_free when the last view of it is gone."""
    addr = C.cast(ptr, C.c_void_p).value
    if not n or not addr:
        if addr:
            _lib.mczi_free(C.c_void_p(addr))
        return np.zeros(0, dtype=COUNTED)
    raw = np.ctypeslib.as_array(C.cast(addr, C.POINTER(C.c_uint8)), shape=(n * COUNTED.itemsize,))
    weakref.finalize(raw, _lib.mczi_free, C.c_void_p(addr))
    return raw.view(COUNTED)


def copy_probe(nbytes: int, reps: int = 10) -> float:
    """Average ms of a plain device copy of nbytes (16-byte vectors)."""
    ms = C.c_double()
    _check(_lib.mczi_copy_probe(nbytes, reps, C.byref(ms)))
    return ms.value


def device_info() -> dict:
    import json
    buf = C.create_string_buffer(4096)
    _check(_lib.mczi_device_info(buf, 4096))
    return json.loads(buf.value.decode())


def set_threads(n: int) -> None:
    """Host worker threads of later calls (0: every core, at most 16)."""
    _check(_lib.mczi_set_threads(n))


def version() -> str:
    return _lib.mczi_version().decode()


@dataclass(frozen=True)
class CounterConfig:
    """CounterConfig (src/lib.rs:48-53)."""
    k: int
    minimizer: int
    threshold: int = 1
    partition_count: int = DEFAULT_PARTITION_COUNT

    def _c(self) -> _Config:
        return _Config(self.k, self.minimizer, self.threshold, self.partition_count)


def validate_config(cfg: CounterConfig, datasets: bool = False) -> None:
    _check(_lib.mczi_validate_config(C.byref(cfg._c()), int(datasets)))


def expand_fofns(fofns) -> list[str]:
    arr, n = _paths(fofns)
    out = C.POINTER(C.c_char_p)()
    nout = C.c_size_t()
    _check(_lib.mczi_expand_fofns(arr, n, C.byref(out), C.byref(nout)))
    res = [out[i].decode() for i in range(nout.value)]
    _lib.mczi_free_strings(out, nout.value)
    return res


def count_inputs(inputs, cfg: CounterConfig) -> np.ndarray:
    """Exact counts > threshold, sorted by encoded k-mer (src/lib.rs:278)."""
    arr, n = _paths(inputs)
    p = C.POINTER(_Counted)()
    nout = C.c_size_t()
    _check(_lib.mczi_count_inputs(arr, n, C.byref(cfg._c()), C.byref(p), C.byref(nout)))
    return _take(p, nout.value)


def count_datasets(inputs, cfg: CounterConfig) -> np.ndarray:
    """Dataset-presence counts > threshold, sorted (src/lib.rs:289)."""
    arr, n = _paths(inputs)
    p = C.POINTER(_Counted)()
    nout = C.c_size_t()
    _check(_lib.mczi_count_datasets(arr, n, C.byref(cfg._c()), C.byref(p), C.byref(nout)))
    return _take(p, nout.value)


def count_inputs_to_fasta_path(inputs, cfg: CounterConfig, out) -> None:
    arr, n = _paths(inputs)
    _check(_lib.mczi_count_inputs_to_fasta_path(arr, n, C.byref(cfg._c()), os.fspath(out).encode()))


def count_datasets_to_fasta_path(inputs, cfg: CounterConfig, out) -> None:
    arr, n = _paths(inputs)
    _check(_lib.mczi_count_datasets_to_fasta_path(arr, n, C.byref(cfg._c()), os.fspath(out).encode()))


def count_inputs_to_kff_path(inputs, cfg: CounterConfig, out) -> None:
    arr, n = _paths(inputs)
    _check(_lib.mczi_count_inputs_to_kff_path(arr, n, C.byref(cfg._c()), os.fspath(out).encode()))


def count_datasets_to_kff_path(inputs, cfg: CounterConfig, out) -> None:
    arr, n = _paths(inputs)
    _check(_lib.mczi_count_datasets_to_kff_path(arr, n, C.byref(cfg._c()), os.fspath(out).encode()))


def count_inputs_to_kmer_fasta_path(inputs, cfg: CounterConfig, out) -> int:
    arr, n = _paths(inputs)
    st = _FastaStats()
    _check(_lib.mczi_count_inputs_to_kmer_fasta_path(arr, n, C.byref(cfg._c()), os.fspath(out).encode(),
                                                      C.byref(st)))
    return st.selected_kmers


def count_datasets_to_kmer_fasta_path(inputs, cfg: CounterConfig, out) -> int:
    arr, n = _paths(inputs)
    st = _FastaStats()
    _check(_lib.mczi_count_datasets_to_kmer_fasta_path(arr, n, C.byref(cfg._c()), os.fspath(out).encode(),
                                                        C.byref(st)))
    return st.selected_kmers


def mczi_subtract(index_inputs, query_inputs, cfg: CounterConfig, out=None, index_datasets: bool = False) -> dict:
    """MCZI in one process: MC-count the index inputs (threshold cfg.threshold;
    index_datasets = one dataset per file), subtract the query k-mers found
    in that set, write the absent k-mers' simplitigs to `out` (None: stats
    only). Returns the query statistics."""
    ia, ni = _paths(index_inputs)
    qa, nq = _paths(query_inputs)
    st = McziStats()
    _check(_lib.mczi_mczi_subtract(ia, ni, int(index_datasets), qa, nq, C.byref(cfg._c()),
                                   None if out is None else os.fspath(out).encode(), C.byref(st)))
    return st.as_dict()


def mczi_absent(index_inputs, query_inputs, cfg: CounterConfig, index_datasets: bool = False):
    """The same subtraction: (sorted absent k-mers, query statistics)."""
    ia, ni = _paths(index_inputs)
    qa, nq = _paths(query_inputs)
    st = McziStats()
    p = C.POINTER(_Counted)()
    n = C.c_size_t()
    _check(_lib.mczi_mczi_absent(ia, ni, int(index_datasets), qa, nq, C.byref(cfg._c()), C.byref(p), C.byref(n),
                                 C.byref(st)))
    return _take(p, n.value), st.as_dict()


def _counted_arr(kmers) -> np.ndarray:
    return np.ascontiguousarray(kmers, dtype=COUNTED)


def _words_call(f, *args):
    wp = C.POINTER(C.c_uint32)()
    op = C.POINTER(C.c_uint64)()
    n = C.c_size_t()
    _check(f(*args, C.byref(wp), C.byref(op), C.byref(n)))
    try:
        off = np.ctypeslib.as_array(op, shape=(n.value + 1,)).copy()
        nw = int((off[-1] + 15) // 16)
        words = np.ctypeslib.as_array(wp, shape=(nw,)).copy() if nw else np.zeros(0, np.uint32)
    finally:
        _lib.mczi_free(C.cast(wp, C.c_void_p))
        _lib.mczi_free(C.cast(op, C.c_void_p))
    return words, off


def write_read_cache(fastx_path, cache_path, min_len: int = 1) -> None:
    """Packed-read cache (.pread) of one FASTA/FASTQ file."""
    _check(_lib.mczi_write_read_cache(os.fspath(fastx_path).encode(), min_len, os.fspath(cache_path).encode()))


def read_read_cache(cache_path, min_len: int = 1):
    """(words, frag_off) of a packed-read cache, as pack_fastx returns them."""
    return _words_call(_lib.mczi_read_read_cache, os.fspath(cache_path).encode(), min_len)


def write_superkmer_partitions(inputs, cfg: CounterConfig, out_dir) -> None:
    """The device scan's super-k-mer records as partition_NNNN.skm files."""
    arr, n = _paths(inputs)
    _check(_lib.mczi_write_superkmer_partitions(arr, n, C.byref(cfg._c()), os.fspath(out_dir).encode()))


def read_superkmer_partitions(paths, min_len: int = 1):
    """Records of .skm files as packed fragments (words, frag_off)."""
    arr, n = _paths(paths)
    return _words_call(_lib.mczi_read_superkmer_partitions, arr, n, min_len)


def count_superkmer_partitions(paths, cfg: CounterConfig) -> np.ndarray:
    """Count the k-mers of .skm partition files on the device (count > x, sorted)."""
    arr, n = _paths(paths)
    p = C.POINTER(_Counted)()
    m = C.c_size_t()
    _check(_lib.mczi_count_superkmer_partitions(arr, n, C.byref(cfg._c()), C.byref(p), C.byref(m)))
    return _take(p, m.value)


def write_counted(path, kmers) -> None:
    a = _counted_arr(kmers)
    _check(_lib.mczi_write_counted(os.fspath(path).encode(), a.ctypes.data if a.size else None, a.size))


def read_counted(path) -> np.ndarray:
    p = C.POINTER(_Counted)()
    m = C.c_size_t()
    _check(_lib.mczi_read_counted(os.fspath(path).encode(), C.byref(p), C.byref(m)))
    return _take(p, m.value)


def pack_fastx(path, min_len: int = 1):
    """Parse and 2-bit pack one FASTA/FASTQ file on the host: (words, frag_off)
    as numpy arrays in the layout Batch.from_packed takes."""
    wp = C.POINTER(C.c_uint32)()
    op = C.POINTER(C.c_uint64)()
    n = C.c_size_t()
    _check(_lib.mczi_pack_fastx(os.fspath(path).encode(), min_len, C.byref(wp), C.byref(op), C.byref(n)))
    try:
        off = np.ctypeslib.as_array(op, shape=(n.value + 1,)).copy()
        nw = int((off[-1] + 15) // 16)
        words = np.ctypeslib.as_array(wp, shape=(nw,)).copy() if nw else np.zeros(0, np.uint32)
    finally:
        _lib.mczi_free(C.cast(wp, C.c_void_p))
        _lib.mczi_free(C.cast(op, C.c_void_p))
    return words, off


def run_inputs_phase12_direct(inputs, cfg: CounterConfig) -> dict:
    """Passes 1 and 2 (minimizer counts, filtered super-k-mer partitioning)
    with their times, distinct minimizers and partition-file bytes."""
    arr, n = _paths(inputs)
    st = Phase12Stats()
    _check(_lib.mczi_run_inputs_phase12_direct(arr, n, C.byref(cfg._c()), C.byref(st)))
    return {f: getattr(st, f) for f, _ in st._fields_}


def count_inputs_minimizer_phase_direct(inputs, cfg: CounterConfig) -> int:
    arr, n = _paths(inputs)
    u = C.c_uint64()
    _check(_lib.mczi_count_inputs_minimizer_phase_direct(arr, n, C.byref(cfg._c()), C.byref(u)))
    return u.value


def write_fasta_path(path, kmers, k: int) -> None:
    a = _counted_arr(kmers)
    _check(_lib.mczi_write_fasta_path(os.fspath(path).encode(), a.ctypes.data if a.size else None, a.size, k))


def write_fasta_after_index(path, kmers, k: int, last_written_index: int) -> int:
    a = _counted_arr(kmers)
    last = C.c_size_t()
    _check(_lib.mczi_write_fasta_after_index(os.fspath(path).encode(), a.ctypes.data if a.size else None, a.size,
                                             k, last_written_index, C.byref(last)))
    return last.value


def _strings(pp, n) -> list:
    try:
        return [pp[i].decode() for i in range(n.value)]
    finally:
        _lib.mczi_free_strings(pp, n.value)


def simplitig_sequences(kmers, k: int) -> list:
    """The greedy simplitigs of the k-mers as strings (simplitig_sequences)."""
    a = _counted_arr(kmers)
    pp, n = C.POINTER(C.c_char_p)(), C.c_size_t()
    _check(_lib.mczi_simplitig_sequences(a.ctypes.data if a.size else None, a.size, k, C.byref(pp), C.byref(n)))
    return _strings(pp, n)


def simplitig_sequences_parallel(kmers, k: int, minimizer: int, partition_count: int = 0) -> list:
    """simplitig_sequences_from_kmers_parallel: the greedy per minimizer partition."""
    a = _counted_arr(kmers)
    pp, n = C.POINTER(C.c_char_p)(), C.c_size_t()
    _check(_lib.mczi_simplitig_sequences_parallel(a.ctypes.data if a.size else None, a.size, k, minimizer,
                                                  partition_count, C.byref(pp), C.byref(n)))
    return _strings(pp, n)


def write_kff(path, kmers, cfg: CounterConfig) -> None:
    a = _counted_arr(kmers)
    _check(_lib.mczi_write_kff_path(os.fspath(path).encode(), a.ctypes.data if a.size else None, a.size,
                                    C.byref(cfg._c())))


def decode_kmer(lo: int, hi: int, k: int) -> str:
    buf = C.create_string_buffer(k + 1)
    _check(_lib.mczi_decode_kmer(lo, hi, k, buf))
    return buf.raw[:k].decode()


def minimizer_run_hashes(packed: bytes | np.ndarray, n_bases: int, k: int, m: int):
    """Device run_hashes: (hashes, starts) of one packed-seq sequence."""
    buf = np.frombuffer(bytes(packed), dtype=np.uint8) if not isinstance(packed, np.ndarray) else packed
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    nw = max(n_bases - k + 1, 1)
    h = np.zeros(nw, dtype=np.uint32)
    s = np.zeros(nw, dtype=np.uint32)
    n = C.c_size_t()
    _check(_lib.mczi_minimizer_run_hashes(buf.ctypes.data, n_bases, k, m, h.ctypes.data, s.ctypes.data,
                                          C.byref(n)))
    return h[: n.value].copy(), s[: n.value].copy()


class Batch:
    """Device-resident segmented read buffer (many reads / files / datasets)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_files(cls, paths, min_len: int = 1, datasets: bool = False) -> "Batch":
        arr, n = _paths(paths)
        h = C.c_void_p()
        _check(_lib.mczi_batch_from_files(arr, n, min_len, int(datasets), C.byref(h)))
        return cls(h)

    @classmethod
    def from_packed(cls, words: np.ndarray, frag_off: np.ndarray, dataset: np.ndarray | None = None,
                    input_frag: np.ndarray | None = None) -> "Batch":
        """Packed reads to a device batch. input_frag (optional) marks where
        each packed input starts (fragment indices, 0 first, n_frag last):
        the plain launch mode then scans one input per launch."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        frag_off = np.ascontiguousarray(frag_off, dtype=np.uint64)
        ds = None if dataset is None else np.ascontiguousarray(dataset, dtype=np.uint32)
        h = C.c_void_p()
        if input_frag is None:
            _check(_lib.mczi_batch_from_packed(words.ctypes.data, frag_off.ctypes.data, frag_off.size - 1,
                                               ds.ctypes.data if ds is not None else None, C.byref(h)))
        else:
            inp = np.ascontiguousarray(input_frag, dtype=np.uint64)
            _check(_lib.mczi_batch_from_packed_inputs(words.ctypes.data, frag_off.ctypes.data, frag_off.size - 1,
                                                      ds.ctypes.data if ds is not None else None,
                                                      inp.ctypes.data, inp.size - 1, C.byref(h)))
        return cls(h)

    def info(self) -> tuple[int, int, int]:
        nb, nf, nd = C.c_uint64(), C.c_uint64(), C.c_uint32()
        _check(_lib.mczi_batch_info(self._h, C.byref(nb), C.byref(nf), C.byref(nd)))
        return nb.value, nf.value, nd.value

    def count(self, cfg: CounterConfig, mode: int = MODE_TWO_PASS, launch: int = LAUNCH_BATCHED,
              order: int = ORDER_ENCODED) -> RunStats:
        st = RunStats()
        _check(_lib.mczi_batch_count(self._h, C.byref(cfg._c()), mode, launch, order, C.byref(st)))
        return st

    def autotune(self, cfg: CounterConfig, mode: int = MODE_TWO_PASS, launch: int = LAUNCH_BATCHED,
                 order: int = ORDER_ENCODED) -> float:
        t = C.c_double()
        _check(_lib.mczi_batch_autotune(self._h, C.byref(cfg._c()), mode, launch, order, C.byref(t)))
        return t.value

    def download(self) -> np.ndarray:
        p = C.POINTER(_Counted)()
        n = C.c_size_t()
        _check(_lib.mczi_batch_download(self._h, C.byref(p), C.byref(n)))
        return _take(p, n.value)

    def close(self) -> None:
        if self._h:
            _lib.mczi_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- synthetic inputs -----------------------------------------------------------
@dataclass
class SynthSpec:
    genome_len: int
    genome_seed: int
    read_seed: int
    n_reads: int = 0
    read_len: int = 150
    sub_rate: float = 0.0
    ins_rate: float = 0.0
    del_rate: float = 0.0
    long_bases: int = 0
    long_median: float = 8000.0
    long_sigma: float = 0.8
    long_min: int = 500
    long_max: int = 100000
    repeat_fraction: float = 0.0
    repeat_len: int = 0
    repeat_copies: int = 0
    n_datasets: int = 0
    snp_rate: float = 0.0

    def _c(self) -> _SynthSpec:
        return _SynthSpec(*[getattr(self, f) for f, _ in _SynthSpec._fields_])


class SynthReads:
    """Packed synthetic reads in host memory (numpy views owned by the library)."""

    def __init__(self, spec: SynthSpec | None = None, _raw: _SynthReads | None = None):
        if _raw is None:
            self._r = _SynthReads()
            _check(_lib.mczi_synth_generate(C.byref(spec._c()), C.byref(self._r)))
        else:
            self._r = _raw
        r = self._r
        self.n_bases = r.n_bases
        self.words = np.ctypeslib.as_array(r.words, shape=(r.n_words,))
        self.frag_off = np.ctypeslib.as_array(r.frag_off, shape=(r.n_frag + 1,))
        self.dataset = np.ctypeslib.as_array(r.dataset, shape=(max(r.n_frag, 1),))[: r.n_frag]

    def codes(self) -> np.ndarray:
        """One packed-seq code per base (uint8)."""
        w = self.words.view(np.uint8)
        b = np.unpackbits(w[:, None], axis=1, bitorder="little").reshape(-1, 4, 2)
        c = (b[:, :, 0] | (b[:, :, 1] << 1)).reshape(-1)
        return c[: self.n_bases].astype(np.uint8)

    def write_fasta(self, path, dataset: int = -1) -> None:
        _check(_lib.mczi_synth_write_fasta(C.byref(self._r), os.fspath(path).encode(), dataset))

    def close(self) -> None:
        if self._r is not None:
            _lib.mczi_synth_free(C.byref(self._r))
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synth_reads_from(genomes, n_reads: int, read_len: int, sub_rate: float, seed: int) -> SynthReads:
    """Reads from packed genomes [(words, length), ...]; read i from genome i % len(genomes)."""
    ws = [np.ascontiguousarray(w, dtype=np.uint32) for w, _ in genomes]
    ptrs = (C.c_void_p * len(ws))(*[w.ctypes.data for w in ws])
    lens = (C.c_uint64 * len(ws))(*[int(n) for _, n in genomes])
    raw = _SynthReads()
    _check(_lib.mczi_synth_reads_from(ptrs, lens, len(ws), n_reads, read_len, sub_rate, seed, C.byref(raw)))
    return SynthReads(_raw=raw)


def synth_genome(genome_len: int, seed: int) -> np.ndarray:
    """Packed words of a SplitMix64 random genome (copy)."""
    p = C.POINTER(C.c_uint32)()
    _check(_lib.mczi_synth_genome(genome_len, seed, C.byref(p)))
    n = (genome_len + 15) // 16 + 32
    arr = np.ctypeslib.as_array(p, shape=(n,)).copy()
    _lib.mczi_free(C.cast(p, C.c_void_p))
    return arr
