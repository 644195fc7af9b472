/*
 * mczi.h — C ABI of the MI355X-native MCZI k-mer counting / subtraction path.
 *
 * Every entry point returns 0 on success and a nonzero code on error; the
 * error text (an anyhow-style context chain, outermost context first) is
 * available from mczi_last_error() on the calling thread. No C++ exception
 * crosses this boundary. Memory returned through an out-pointer belongs to
 * the caller and is released with mczi_free() (or the named *_free call).
 *
 * Reference citations are relative to the reference checkout
 * (src/lib.rs, src/zi.rs, src/main.rs, vendor/simd-minimizers-2.3.1/src/...).
 * The device work (scan, partitioning, counting, filters, ZI lookups) runs as
 * HIP kernels on the current device; calls made from several host threads are
 * serialized per device by the library.
 */
#ifndef MCZI_H
#define MCZI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
enum {
    MCZI_OK = 0,
    MCZI_E_INVALID = 1,  /* configuration or argument rejected */
    MCZI_E_IO = 2,       /* file could not be read or written */
    MCZI_E_DEVICE = 3,   /* HIP runtime / kernel failure */
    MCZI_E_NOMEM = 4,    /* host or device memory exhausted, or over budget */
    MCZI_E_NOTFOUND = 5, /* unknown request / handle */
    MCZI_E_PENDING = 6   /* request not finished yet (non-blocking poll) */
};

/* CounterConfig (src/lib.rs:48-53). k-mers are stored 2 bits per base,
 * A0 C1 G2 T3, first base most significant, in (hi:lo) 128 bits. */
typedef struct {
    uint32_t k;
    uint32_t minimizer;
    uint64_t threshold;
    uint32_t partition_count;
} mczi_counter_config;

/* CountedKmer (src/lib.rs:63-66): encoded = (hi << 64) | lo. */
typedef struct {
    uint64_t lo;
    uint64_t hi;
    uint64_t count;
} mczi_counted_kmer;

/* KmerFastaWriteStats (src/lib.rs:103-108), without the per-phase vector. */
typedef struct {
    uint64_t selected_kmers;
    uint64_t minimizers_above_threshold;
    double total_seconds;
} mczi_kmer_fasta_stats;

/* DEFAULT_PARTITION_COUNT (src/lib.rs:29). */
#define MCZI_DEFAULT_PARTITION_COUNT 1024u

const char *mczi_last_error(void);
/* Host worker threads of later calls (ingest, writers, upload staging): the
 * global Rayon pool size the MC and ZI command lines set from -t/--threads
 * (src/main.rs:73-79; ThreadPoolBuilder::num_threads). 0 restores the
 * default: every host core, at most 16. */
int mczi_set_threads(uint32_t n);
const char *mczi_version(void);
void mczi_free(void *p);

/* validate_config (src/lib.rs:110-117) / validate_dataset_presence_config
 * (src/lib.rs:141-152). datasets != 0 selects the fof rules. */
int mczi_validate_config(const mczi_counter_config *cfg, int datasets);

/* expand_fofns (src/lib.rs:242-276). *out is an array of n_out C strings,
 * released with mczi_free_strings. */
int mczi_expand_fofns(const char *const *fofns, size_t n, char ***out, size_t *n_out);
void mczi_free_strings(char **strs, size_t n);

/* count_inputs (src/lib.rs:278-287): exact counts > threshold, sorted
 * ascending by encoded k-mer. */
int mczi_count_inputs(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                      mczi_counted_kmer **out, size_t *n_out);
/* count_datasets (src/lib.rs:289-304): count = number of datasets (one input
 * file each) holding the k-mer, emitted when > threshold, sorted. */
int mczi_count_datasets(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                        mczi_counted_kmer **out, size_t *n_out);

/* count_inputs_to_fasta_path (src/lib.rs:427-440): simplitig FASTA. */
int mczi_count_inputs_to_fasta_path(const char *const *paths, size_t n,
                                    const mczi_counter_config *cfg, const char *out_path);
/* count_datasets_to_fasta_path (src/lib.rs:442-455). */
int mczi_count_datasets_to_fasta_path(const char *const *paths, size_t n,
                                      const mczi_counter_config *cfg, const char *out_path);
/* count_inputs_to_kff_path (src/lib.rs:457-507): k <= 32 takes the KMC-style
 * single-pass u8 path (src/lib.rs:599-635), k > 32 the 8-byte-count path. */
int mczi_count_inputs_to_kff_path(const char *const *paths, size_t n,
                                  const mczi_counter_config *cfg, const char *out_path);
/* count_datasets_to_kff_path (src/lib.rs:509-522): 4-byte counts. */
int mczi_count_datasets_to_kff_path(const char *const *paths, size_t n,
                                    const mczi_counter_config *cfg, const char *out_path);
/* count_inputs_to_kmer_fasta_path_with_stats (src/lib.rs:314-355). */
int mczi_count_inputs_to_kmer_fasta_path(const char *const *paths, size_t n,
                                         const mczi_counter_config *cfg, const char *out_path,
                                         mczi_kmer_fasta_stats *stats);
/* count_datasets_to_kmer_fasta_path_with_stats (src/lib.rs:365-371). */
int mczi_count_datasets_to_kmer_fasta_path(const char *const *paths, size_t n,
                                           const mczi_counter_config *cfg,
                                           const char *out_path, mczi_kmer_fasta_stats *stats);

/* Phase12Stats (src/lib.rs:69-75): seconds and pass-1 / partition figures. */
typedef struct {
    double phase1, phase2, total;
    uint64_t unique_minimizer_hashes;
    uint64_t partition_bytes;
} mczi_phase12_stats;
/* run_inputs_phase12_direct (src/lib.rs:160-165; the production minimizer
 * order): pass 1 (minimizer counts) and pass 2 (filtered super-k-mer
 * partitioning) on the device; partition_bytes is the size the reference's
 * partition files would have. The SimdValueHash and AntiLex orders of the
 * other benchmark variants are not provided. */
int mczi_run_inputs_phase12_direct(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                   mczi_phase12_stats *stats);
/* count_inputs_minimizer_phase_direct / _packed_direct (src/lib.rs:1836-1854):
 * pass 1 only; *unique gets the number of distinct minimizer keys. */
int mczi_count_inputs_minimizer_phase_direct(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                             uint64_t *unique);

/* write_fasta_path (src/lib.rs:1203-1207) and write_fasta_after_index
 * (src/lib.rs:1194-1201, here writing to a path); last_index may be NULL. */
int mczi_write_fasta_path(const char *path, const mczi_counted_kmer *kmers, size_t n, uint32_t k);
int mczi_write_fasta_after_index(const char *path, const mczi_counted_kmer *kmers, size_t n,
                                 uint32_t k, size_t last_written_index, size_t *last_index);
/* write_kff (src/lib.rs:1278-1293): 8-byte counts, ordered=1. */
/* simplitig_sequences (src/lib.rs:1209-1211): the simplitigs of the greedy
 * cover above as NUL-terminated ASCII strings, in its order. */
int mczi_simplitig_sequences(const mczi_counted_kmer *kmers, size_t n, uint32_t k, char ***seqs,
                             size_t *n_seqs);
/* simplitig_sequences_from_kmers_parallel (src/lib.rs:1217-1264): k-mers split
 * into partition_count partitions (a power of two up to 1024; 0 = the
 * reference's default: 16 per worker thread, or MC_SIMPLITIG_PARTITIONS) by the
 * hash of their smallest canonical minimizer-size m-mer, the greedy run on
 * each, results concatenated in partition order. Both lists are released with
 * mczi_free_strings. */
int mczi_simplitig_sequences_parallel(const mczi_counted_kmer *kmers, size_t n, uint32_t k,
                                      uint32_t minimizer, uint32_t partition_count, char ***seqs,
                                      size_t *n_seqs);
int mczi_write_kff_path(const char *path, const mczi_counted_kmer *kmers, size_t n,
                        const mczi_counter_config *cfg);
/* decode_kmer (src/lib.rs:3056-3069); out must hold k bytes (no NUL added). */
int mczi_decode_kmer(uint64_t lo, uint64_t hi, uint32_t k, char *out);

/* ---- kernel-level boundary ------------------------------------------------
 * Builder::run_hashes with super_kmers (vendor/simd-minimizers-2.3.1/src/
 * lib.rs:569-616) as consumed by for_each_minimizer_run_on_seq
 * (src/lib.rs:3694-3752): canonical minimizers (hash length m, window
 * w = k - m + 1) of ONE packed sequence, computed on the device.
 * packed: packed-seq bytes (A0 C1 T2 G3, 4 bases per byte, LSB first).
 * hashes/starts must hold n_bases - k + 1 entries; *n_runs gets the number
 * of runs (deduplicated selected hashes and the first window of each). */
int mczi_minimizer_run_hashes(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m,
                              uint32_t *hashes, uint32_t *starts, size_t *n_runs);
/* No counterpart in the reference: the same as mczi_minimizer_run_hashes
 * (same arguments, outputs and errors) for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): the
 * device memory of a call is a fixed array the library holds on each GPU it
 * runs on (about 69 MiB, kept until the process ends), so no device
 * allocation is made. n_bases is at most 16777216 (2^24): longer sequences
 * are rejected with MCZI_E_INVALID. Calls on one GPU share that array; the
 * library serializes them, and calls that bypass its lock and overlap on one
 * GPU would race. */
int mczi_minimizer_run_hashes_noalloc(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m,
                                      uint32_t *hashes, uint32_t *starts, size_t *n_runs);

/* ---- batched device API ---------------------------------------------------
 * A device-resident segmented buffer of reads (many files / datasets packed
 * together). Fragments are ACGT-only runs; fragment i spans bases
 * [frag_off[i], frag_off[i+1]) of the packed stream, which is 16 bases per
 * little-endian uint32 word (the packed-seq byte layout read as words).
 * dataset (optional) gives the fof dataset of each fragment. */
typedef struct mczi_batch mczi_batch;

int mczi_batch_from_files(const char *const *paths, size_t n, uint32_t min_len,
                          int one_dataset_per_file, mczi_batch **out);
/* Host-side ingest of one FASTA/FASTQ file into the packed layout above
 * (for_fastx_packed_batches, src/lib.rs:3264-3308): fragments split at
 * non-ACGT, shorter than min_len dropped. *words holds ceil(n_bases / 16)
 * words and *frag_off n_frag + 1 offsets; both are released with mczi_free. */
int mczi_pack_fastx(const char *path, uint32_t min_len, uint32_t **words, uint64_t **frag_off,
                    size_t *n_frag);
/* On-disk formats of the reference's partitioned path, for interop and spill
 * (big-endian integers; packed-seq codes A0 C1 T2 G3, 4 bases per byte from
 * bit 0, unused high bits of a last byte zero):
 * - packed-read cache (write_packed_read_cache_batch / for_packed_read_cache_
 *   batches, src/lib.rs:3394-3450): "MCRD0001", then per fragment a u32
 *   length and its packed bytes. The reader returns the layout of
 *   mczi_pack_fastx; fragments shorter than min_len are skipped.
 * - super-k-mer partitions (partition_NNNN.skm; append_record_len,
 *   add_packed_range and read_superkmer_partition, src/lib.rs:3209-3216,
 *   4893-4917, 2692-2711): per record a length (one byte < 255, or 0xFF and a
 *   u32) and its packed bases. mczi_write_superkmer_partitions runs the device
 *   scan and writes the records of each partition spread_hash_u32(key) % P to
 *   dir/partition_NNNN.skm (record splits differ from the reference's; the
 *   k-mer multiset of each partition is the same). mczi_count_superkmer_
 *   partitions counts such files on the device, as the reference's phase 3
 *   (count > x, sorted). mczi_read_superkmer_partitions returns their records
 *   as packed fragments.
 * - counted k-mers (.cnt; process_superkmer_partition / emit_counted_partition,
 *   src/lib.rs:2413-2421, 2657-2678): per k-mer u128 encoding and u64 count. */
int mczi_write_read_cache(const char *fastx_path, uint32_t min_len, const char *cache_path);
int mczi_read_read_cache(const char *cache_path, uint32_t min_len, uint32_t **words, uint64_t **frag_off,
                         size_t *n_frag);
int mczi_write_superkmer_partitions(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                    const char *dir);
int mczi_read_superkmer_partitions(const char *const *paths, size_t n, uint32_t min_len, uint32_t **words,
                                   uint64_t **frag_off, size_t *n_frag);
int mczi_count_superkmer_partitions(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                    mczi_counted_kmer **out, size_t *n_out);
int mczi_write_counted(const char *path, const mczi_counted_kmer *kmers, size_t n);
int mczi_read_counted(const char *path, mczi_counted_kmer **out, size_t *n_out);
int mczi_batch_from_packed(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                           const uint32_t *dataset, mczi_batch **out);
/* Same, with the inputs the caller packed together marked: input c holds
 * fragments [input_frag[c], input_frag[c+1]) (input_frag[0] = 0,
 * input_frag[n_inputs] = n_frag, non-decreasing). The plain launch mode then
 * scans one input per launch, as the reference reads one file at a time in
 * batches of at most BATCH_BASES (src/lib.rs:28, for_fastx_packed_batches
 * src/lib.rs:3264-3308); inputs above 128 Mi bases are split further. The
 * batched mode and the results do not depend on the marks. */
int mczi_batch_from_packed_inputs(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                                  const uint32_t *dataset, const uint64_t *input_frag, size_t n_inputs,
                                  mczi_batch **out);
int mczi_batch_info(const mczi_batch *b, uint64_t *n_bases, uint64_t *n_frag,
                    uint32_t *n_datasets);
void mczi_batch_free(mczi_batch *b);

/* Counting modes. The two-pass modes run the reference's pass 1 on the
 * device (minimizer counts saturating at x + 1, src/lib.rs:3462-3481 and
 * 4171-4251; or per-dataset minimizer presence, src/lib.rs:4253-4370) and
 * stage only the super-k-mers whose minimizer passes (src/lib.rs:3483-3523,
 * 4090-4128). The single-pass modes stage every super-k-mer (the KMC-style
 * path, src/lib.rs:599-635). Both give the same k-mers and counts: every
 * occurrence of a k-mer has the same minimizer, whose count (presence) is at
 * least the k-mer's (min(c, x + 1)), so the filter only drops k-mers that
 * fail count > x. The two-pass modes also report the minimizer statistics. */
enum {
    MCZI_MODE_TWO_PASS = 0,        /* pass 1 + filtered super-k-mers (FASTA path) */
    MCZI_MODE_SINGLE_PASS = 1,     /* all super-k-mers (KMC-style KFF path) */
    MCZI_MODE_DATASETS = 2,        /* fof dataset presence, all super-k-mers */
    MCZI_MODE_DATASETS_TWO_PASS = 3 /* fof: dataset minimizer presence pass, then filtered super-k-mers */
};
/* Launch structure. */
enum {
    MCZI_LAUNCH_BATCHED = 0, /* one persistent launch per pass over the whole buffer (graph) */
    MCZI_LAUNCH_PLAIN = 1    /* per-file path: with several inputs (files / input marks) and the
                                encoded order (k <= 32; fof: one dataset per input), each input
                                is counted on its own and the per-file counts are merged on the
                                device; otherwise one scan launch per input chunk (per file /
                                128 Mi bases) and one count. Same output as BATCHED. */
};
/* Output order. */
enum {
    MCZI_ORDER_ENCODED = 0,  /* ascending encoded k-mer (count_inputs) */
    MCZI_ORDER_KFF_U8 = 1,   /* reference KFF u8 order: partition, encoded & 255, encoded */
    MCZI_ORDER_NONE = 2
};

typedef struct {
    uint64_t n_bases;            /* bases in the batch */
    uint64_t n_windows;          /* k-mer instances scanned */
    uint64_t n_pieces;           /* minimizer runs (pieces) found */
    uint64_t n_records;          /* super-k-mer records written to partitions */
    uint64_t n_distinct;         /* distinct k-mers counted */
    uint64_t n_overflow;         /* distinct k-mers that spilled from LDS to HBM */
    uint64_t n_out;              /* k-mers emitted */
    uint64_t tile_windows;       /* windows per scan tile (pieces are also cut at tile starts) */
    uint64_t partition_bytes;    /* bytes the staged super-k-mers take in the reference's
                                    partition files (src/lib.rs:3209-3216, 575) */
    uint32_t internal_partitions;
    uint32_t lds_slots;
    double ms_scan1, ms_scan2, ms_count, ms_sort, ms_total; /* device time per phase */
    /* two-pass modes: distinct minimizer keys of pass 1 (unique_hashes(),
     * src/lib.rs:4630-4632), and the keys that pass the filter (count > x;
     * fof: present in more than x datasets, the retained minimizers of
     * src/lib.rs:815); ms_pass1 = pass-1 device time */
    uint64_t unique_minimizers;
    uint64_t minimizers_retained;
    double ms_pass1;
    /* partition-group passes the count took (1: one pass; more when one
     * pass would not fit in device memory, or MCZI_GROUP_PASSES asks) */
    uint64_t partition_groups;
    /* scan blocks per CU of the persistent scan grid (tuned; see
     * mczi_batch_autotune) */
    uint64_t scan_blocks_per_cu;
} mczi_run_stats;

/* Count the k-mers of a batch on the device. Results stay on the device
 * until mczi_batch_download; stats may be NULL. With stats, partition_bytes
 * is exact (runs cut by the scan's tiles are joined on the host). */
int mczi_batch_count(mczi_batch *b, const mczi_counter_config *cfg, int mode, int launch,
                     int order, mczi_run_stats *stats);
int mczi_batch_download(mczi_batch *b, mczi_counted_kmer **out, size_t *n_out);
/* Autotuner (no counterpart in the reference): times the count over the LDS
 * table size (2^10, 2^11, 2^12 slots; k <= 32), the expected k-mer instances
 * per internal partition (which sets the partition count) and the scan's
 * blocks per CU, keeps the fastest for later mczi_batch_count calls on this
 * batch, and appends it to the tuning cache (MCZI_TUNE_CACHE, else
 * $HOME/.cache/mczi/autotune.tsv), which later counts of the same device,
 * k, m, mode, order and batch shape use. The result does not change the
 * output. *instances_per_partition gets the chosen fill. */
int mczi_batch_autotune(mczi_batch *b, const mczi_counter_config *cfg, int mode, int launch, int order,
                        double *instances_per_partition);

/* Device copy of `bytes` bytes (16-byte vector loads and stores), `reps`
 * times; *ms_per_copy gets the average duration. No counterpart in the
 * reference: it calibrates HBM byte counters and measures the box's copy
 * bandwidth for roofline reporting. */
int mczi_copy_probe(uint64_t bytes, int reps, double *ms_per_copy);
/* Device description as JSON text (name, gcnArchName, CUs, clocks, memory,
 * LDS limits); no counterpart in the reference. */
int mczi_device_info(char *buf, size_t len);

/* ---- ZI (src/zi.rs) -------------------------------------------------------- */
/* ZI main path (src/zi.rs:64-100): absent query k-mers as simplitig FASTA.
 * m == 0 selects default_minimizer_size (src/zi.rs:109-112). threads > 0
 * sets the host worker threads as mczi_set_threads does (0 keeps them).
 * ram_limit_gib bounds the device memory of one query pass: query fragments
 * beyond it are looked up in several passes whose absent sets are merged
 * (the reference bounds SSHash construction memory with it). */
int mczi_zi_subtract(const char *index_path, const char *const *query_paths, size_t n,
                     uint32_t k, uint32_t m, const char *out_path, uint32_t threads,
                     uint32_t ram_limit_gib);
/* collect_query_kmers_absent_from_index (src/zi.rs:152-175) as a library
 * call: sorted, deduplicated absent canonical k-mers (count = 1). */
int mczi_zi_absent(const char *index_path, const char *const *query_paths, size_t n,
                   uint32_t k, mczi_counted_kmer **out, size_t *n_out);

/* ---- ZI serving: a device-resident index and a continuous-batching loop ----
 * No counterpart in the reference (which runs one query set per process).
 * Requests are FASTA/FASTQ text (or packed reads); each result is the sorted
 * absent set and, on request, its simplitig FASTA text. The absent arrays
 * are host memory owned by the library (pinned while the host lets the
 * process pin more, else ordinary): release the whole result with
 * mczi_zi_result_free (never free()). */
typedef struct mczi_zi_server mczi_zi_server;
typedef struct {
    uint64_t request_id;
    uint64_t n_windows;   /* query k-mers looked up */
    uint64_t n_absent;    /* distinct absent k-mers */
    uint64_t *absent_lo;  /* sorted absent k-mers, low 64 bits */
    uint64_t *absent_hi;  /* high 64 bits (k > 32), else NULL */
    char *fasta;          /* simplitig FASTA text (NUL-terminated) */
    size_t fasta_len;
    double queue_ms;      /* submit -> admitted */
    double latency_ms;    /* submit -> result ready */
} mczi_zi_result;

int mczi_zi_server_create(const char *index_path, uint32_t k, uint32_t m,
                          uint64_t hbm_budget_bytes, uint32_t max_inflight,
                          mczi_zi_server **out);
/* Build the server from an in-memory index (FASTA text). */
int mczi_zi_server_create_from_text(const char *index_fasta, size_t len, uint32_t k, uint32_t m,
                                    uint64_t hbm_budget_bytes, uint32_t max_inflight,
                                    mczi_zi_server **out);
int mczi_zi_server_submit(mczi_zi_server *s, const char *fastx, size_t len, int make_fasta,
                          uint64_t *request_id);
/* Same, for reads already packed (16 bases per u32 word, fragment offsets
 * as in mczi_batch_from_packed); fragments shorter than k are skipped. */
int mczi_zi_server_submit_packed(mczi_zi_server *s, const uint32_t *words, const uint64_t *frag_off,
                                 size_t n_frag, int make_fasta, uint64_t *request_id);
/* wait != 0 blocks until the request is done; otherwise MCZI_E_PENDING. */
int mczi_zi_server_poll(mczi_zi_server *s, uint64_t request_id, int wait,
                        mczi_zi_result **result);
void mczi_zi_result_free(mczi_zi_result *r);
int mczi_zi_server_stats(mczi_zi_server *s, uint64_t *index_kmers, uint64_t *completed,
                         uint64_t *lookups, double *busy_ms);
/* Lookups done so far and the summed duration of the lookup kernels that did
 * them (HIP events on the slot streams; no counterpart in the reference). */
int mczi_zi_server_lookup_stats(mczi_zi_server *s, uint64_t *lookups, double *lookup_kernel_ms);
void mczi_zi_server_destroy(mczi_zi_server *s);

/* MCZI in one process (src/mczi.rs:238-451), without GGCAT and SSHash: the
 * index inputs are counted by MC (cfg: k odd in [3, 63], minimizer in (0, k),
 * threshold x; index_datasets != 0 = one dataset per file, fof presence,
 * --index-fofn) and the k-mers above the threshold become the device
 * membership index directly; the query k-mers absent from it are subtracted
 * as in ZI. Logs MCZI phase lines and the MCZI_STAT query statistics
 * (src/mczi.rs:1600-1650). out_path receives the simplitig FASTA of the absent
 * canonical k-mers (the greedy writer above: the reference compacts the same
 * set with GGCAT, so the k-mer set is identical and the layout differs);
 * out_path NULL = --output-mode no-output. */
typedef struct {
    uint64_t index_kmers;                     /* index k-mers above the threshold */
    uint64_t query_kmers_scanned;             /* query windows */
    uint64_t query_kmers_filtered;            /* windows whose k-mer is in the index */
    uint64_t query_kmers_not_filtered;        /* windows whose k-mer is absent */
    uint64_t query_unique_kmers_not_filtered; /* distinct absent canonical k-mers */
} mczi_mczi_stats;
int mczi_mczi_subtract(const char *const *index_paths, size_t n_index, int index_datasets,
                       const char *const *query_paths, size_t n_query, const mczi_counter_config *cfg,
                       const char *out_path, mczi_mczi_stats *stats);
/* The same subtraction returning the sorted absent k-mers (count 1 each),
 * released with mczi_free. */
int mczi_mczi_absent(const char *const *index_paths, size_t n_index, int index_datasets,
                     const char *const *query_paths, size_t n_query, const mczi_counter_config *cfg,
                     mczi_counted_kmer **out, size_t *n_out, mczi_mczi_stats *stats);

/* ---- synthetic inputs (no counterpart in the reference; test/bench data) --
 * SplitMix64 genome and reads (bench/simulate_long_reads.rs:9-30, 166-180),
 * extended with random strand, iid errors, indels and repeats. Produces
 * packed words + fragment offsets + read-to-dataset ids in host memory. */
typedef struct {
    uint64_t genome_len;
    uint64_t genome_seed;
    uint64_t read_seed;
    uint64_t n_reads;        /* short reads (or 0) */
    uint32_t read_len;       /* short-read length */
    double sub_rate, ins_rate, del_rate;
    uint64_t long_bases;     /* total bases of lognormal long reads (or 0) */
    double long_median, long_sigma;
    uint32_t long_min, long_max;
    double repeat_fraction;  /* genome fraction duplicated as repeats */
    uint32_t repeat_len, repeat_copies;
    uint32_t n_datasets;     /* > 0: dataset variants of the genome (fof) */
    double snp_rate;         /* per-dataset SNP rate */
} mczi_synth_spec;

typedef struct {
    uint32_t *words;
    uint64_t n_words;
    uint64_t *frag_off;
    uint64_t n_frag;
    uint32_t *dataset;
    uint64_t n_bases;
} mczi_synth_reads;

int mczi_synth_generate(const mczi_synth_spec *spec, mczi_synth_reads *out);
int mczi_synth_genome(uint64_t genome_len, uint64_t seed, uint32_t **words);
/* Reads sampled from given genomes (packed words): read i comes from genome
 * i % n_genomes; same error / strand model as mczi_synth_generate. */
int mczi_synth_reads_from(const uint32_t *const *genome_words, const uint64_t *genome_len,
                          uint32_t n_genomes, uint64_t n_reads, uint32_t read_len, double sub_rate,
                          uint64_t seed, mczi_synth_reads *out);
/* Write reads as FASTA (one record per read); dataset_filter < 0 for all. */
int mczi_synth_write_fasta(const mczi_synth_reads *r, const char *path, int64_t dataset_filter);
void mczi_synth_free(mczi_synth_reads *r);

#ifdef __cplusplus
}
#endif
#endif
