/*
 * CPU oracle for the MCZI k-mer counting and subtraction path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (mczi_amd/libmczi.so) never links, loads or calls it.
 *
 * It is a plain-C restatement of the reference algorithm. Every function
 * cites the reference file:line it follows (paths relative to the reference
 * checkout: src/lib.rs, src/zi.rs, vendor/simd-minimizers-2.3.1/src/...).
 *
 * Base codes used throughout are packed-seq codes A0 C1 T2 G3, one byte per
 * base (the reference's packed buffers hold the same codes, 4 per byte,
 * LSB-first). MC k-mer encoding is A0 C1 G2 T3, first base most significant
 * (code ^ (code >> 1) converts, src/lib.rs:3192-3197).
 *
 * Pinning status of each function is listed in oracle.c's header comment.
 */
#ifndef MCZI_ORACLE_H
#define MCZI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t lo;    /* low 64 bits of the encoded k-mer */
    uint64_t hi;    /* high 64 bits (k > 32 only) */
    uint64_t count;
} or_counted;

/* A set of fragments: codes[off[i] .. off[i+1]) is fragment i. */
typedef struct {
    const uint8_t *codes;
    const uint64_t *off;
    size_t nfrag;
    const uint32_t *dataset; /* optional per-fragment dataset id (fof) */
} or_frags;

/* ---- hashing ---------------------------------------------------------- */
uint32_t or_nthash32(const uint8_t *codes, int m, int canonical);
uint32_t or_spread_hash_u32(uint32_t v);
uint64_t or_spread_hash_u64(uint64_t v);

/* ---- scan (simd-minimizers scalar spec) ------------------------------- */
/* Selected hash per window; out has n-k+1 entries. Returns windows. */
size_t or_window_hashes(const uint8_t *codes, size_t n, int k, int m, uint32_t *out);
/* Runs: deduplicated selected hashes + first window of each run. */
size_t or_minimizer_runs(const uint8_t *codes, size_t n, int k, int m,
                         uint32_t *run_hash, uint32_t *run_start);
/* Positions API (dedup of per-window selected positions), canonical or forward. */
size_t or_minimizer_positions(const uint8_t *codes, size_t n, int k_hash, int w,
                              int canonical, uint32_t *out);

/* ---- MC counting ------------------------------------------------------- */
int or_count_bruteforce(const or_frags *f, int k, uint64_t threshold,
                        or_counted **out, size_t *n_out);
int or_count_two_pass(const or_frags *f, int k, int m, uint64_t threshold,
                      size_t partitions, or_counted **out, size_t *n_out,
                      uint64_t *partition_bytes, uint64_t *minimizers_kept);
/* The same two-pass count on `threads` host threads, as the reference runs
 * it under Rayon: per-thread minimizer counts merged per shard, per-thread
 * partition buffers, partitions counted in parallel. Same output. */
int or_count_two_pass_mt(const or_frags *f, int k, int m, uint64_t threshold,
                         size_t partitions, int threads, or_counted **out, size_t *n_out);
/* KMC-style single-pass KFF u8 path: full KFF file bytes. */
int or_kff_u8_kmc(const or_frags *f, int k, int m, uint64_t threshold,
                  size_t partitions, uint8_t **bytes, size_t *n_bytes,
                  uint64_t *records);
/* fof dataset presence (counts = number of datasets holding the k-mer). */
int or_count_datasets(const or_frags *f, size_t n_datasets, int k, int m,
                      uint64_t threshold, size_t partitions,
                      or_counted **out, size_t *n_out, uint64_t *minimizers_kept);
/* The same on `threads` host threads (datasets in order, each dataset's
 * reads in parallel, as the reference's Rayon path). Same output. */
int or_count_datasets_mt(const or_frags *f, size_t n_datasets, int k, int m,
                         uint64_t threshold, size_t partitions, int threads,
                         or_counted **out, size_t *n_out, uint64_t *minimizers_kept);
int or_count_datasets_bruteforce(const or_frags *f, size_t n_datasets, int k,
                                 uint64_t threshold, or_counted **out, size_t *n_out);

/* ---- ZI ---------------------------------------------------------------- */
int or_zi_absent(const or_frags *index, const or_frags *query, int k,
                 or_counted **out, size_t *n_out);
/* ZI's one-thread query loop (hash-set index, per-window lookups), with the
 * index build and the lookups timed apart. */
int or_zi_lookup_timed(const or_frags *index, const or_frags *query, int k, uint64_t *lookups,
                       uint64_t *absent_unique, double *build_s, double *query_s);

/* Exact occurrence counts of 128-bit canonical query k-mers (33 <= k <= 64)
 * over packed reads (16 bases per u32 word). */
int or_spot_counts128(const uint32_t *words, const uint64_t *off, size_t n_frag, int k, const uint64_t *qlo,
                      const uint64_t *qhi, size_t n_query, uint64_t *counts, int threads);

/* ---- writers ----------------------------------------------------------- */
/* Simplitig FASTA (sorted input). Returns index of last simplitig written. */
int or_write_simplitigs(const char *path, const or_counted *kmers, size_t n,
                        int k, size_t first_index, size_t *last_index);
int or_write_kmer_fasta(const char *path, const or_counted *kmers, size_t n, int k);
int or_write_kff(const char *path, const or_counted *kmers, size_t n, int k, int m,
                 uint64_t threshold, int count_size, int ordered);

/* ---- input ------------------------------------------------------------- */
int or_parse_fastx(const char *path, int min_len, uint8_t **codes, uint64_t **off,
                   size_t *nfrag);

/* ---- full-size spot checks -------------------------------------------- */
/* Exact occurrence count and dataset-presence mask of each query k-mer
 * (canonical MC encoding, k <= 32) over reads packed 16 bases per u32 word
 * (packed-seq codes, base i at bits 2(i%16) of word i/16), fragment i =
 * bases [off[i], off[i+1]). dataset may be NULL (all 0); masks may be NULL.
 * Counts every window exactly like or_count_bruteforce, on `threads`
 * threads over fragment ranges. */
int or_spot_counts(const uint32_t *words, const uint64_t *off, size_t n_frag, const uint32_t *dataset, int k,
                   const uint64_t *query, size_t n_query, uint64_t *counts, uint32_t *masks, int threads);

/* ---- small helpers ----------------------------------------------------- */
uint64_t or_canonical_u64(const uint8_t *codes, int k); /* k <= 32 */
void or_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
