// Host-side internals of libmczi: errors, device context, packed reads and
// the MC / ZI drivers. Everything here stays behind the C ABI in mczi.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <atomic>
#include <functional>
#include <map>
#include <set>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mczi.h"

namespace mczi {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define MCZI_HIP(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw ::mczi::Error(MCZI_E_DEVICE, std::string("HIP error '") +                \
                                                   hipGetErrorString(e_) + "' in " #x);   \
    } while (0)

inline void require(bool ok, int code, const std::string &msg) {
    if (!ok) throw Error(code, msg);
}
// Literal messages: the string is only built on failure (checks run inside
// loops over millions of fragments).
inline void require(bool ok, int code, const char *msg) {
    if (!ok) throw Error(code, msg);
}

// ---------------------------------------------------------------------------
// Host worker threads: the limit set by mczi_set_threads (the reference's
// global Rayon pool size, src/main.rs:73-79), else up to 16 (the host cores
// one GPU gets); one for small jobs.
extern std::atomic<unsigned> g_host_threads;
inline unsigned host_workers(size_t n) {
    if (n < (size_t(1) << 16)) return 1;
    const unsigned lim = g_host_threads.load(std::memory_order_relaxed);
    if (lim) return lim;
    const unsigned hc = std::thread::hardware_concurrency();
    return hc == 0 ? 1u : (hc < 16u ? hc : 16u);
}

// f(begin, end) over [0, n) in chunks of `grain`, on up to nt threads.
template <class F> void parallel_ranges(size_t n, size_t grain, unsigned nt, F &&f) {
    std::atomic<size_t> next{0};
    auto run = [&] {
        for (;;) {
            const size_t b = next.fetch_add(grain, std::memory_order_relaxed);
            if (b >= n) return;
            f(b, b + grain < n ? b + grain : n);
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt && grain < n; t++) {
        try {
            th.emplace_back(run);
        } catch (...) {
            break;  // fewer workers; the chunks still all get done
        }
    }
    run();
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------------------
// Device context: one stream for the synchronous API, a caching allocator.
class Device {
  public:
    static Device &get();
    hipStream_t stream() const { return stream_; }
    int cu_count() const { return cus_; }
    int device() const { return dev_; }
    // Cached device allocation (rounded to 2 MiB); contents undefined.
    // Released blocks stay cached up to kDeviceCacheBytes (beyond that they
    // are freed, so that one large job does not keep the device's memory
    // from the next); a failed allocation frees cached blocks, largest
    // first, and tries again.
    void *alloc(size_t bytes);
    void release(void *p);
    size_t pooled_bytes();  // released device blocks kept for reuse
    // Pinned host allocation: device <-> host copies into it run as plain
    // DMA. Blocks up to kSlabBytes come from a buddy allocator over 1 GiB
    // slabs of hipHostMalloc memory (power-of-two blocks from 64 KiB; a
    // released block merges with its free buddy), so a steady stream of
    // requests of any mix of sizes reuses pinned pages instead of pinning
    // fresh ones (about 29 ms per GiB, and it stalls the threads copying into
    // other pinned blocks meanwhile). Slabs stay pinned until the process
    // ends, at most kSlabCapBytes of them. Larger blocks are pinned on their
    // own and cached by size up to kPinnedCacheBytes.
    // With allow_pageable, a request the host cannot pin (a process may pin
    // only so much) is served from ordinary memory instead of failing;
    // host_release frees every kind.
    void *host_alloc(size_t bytes, bool allow_pageable = false);
    void host_release(void *p);
    // Pins slabs until at least `bytes` of them exist (a server's staging
    // arena, taken once at creation rather than on the request path).
    void host_reserve(size_t bytes);
    size_t host_slab_bytes();
    // Host -> device copy of pageable memory: the source is copied by
    // several host threads into two alternating pinned staging blocks, each
    // sent by DMA while the next one fills. Synchronous on `s`.
    void upload(void *dst, const void *src, size_t bytes, hipStream_t s);
    // Device -> host copy into pageable memory, the same way round: DMA into
    // alternating pinned blocks, each copied out by several threads while the
    // next one arrives. Synchronous on `s`.
    void download(void *dst, const void *src, size_t bytes, hipStream_t s);
    std::mutex &call_mutex() { return call_mu_; }

  private:
    Device();
    hipStream_t stream_ = nullptr;
    int cus_ = 0;
    int dev_ = 0;
    std::mutex mu_;
    std::mutex call_mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t pooled_ = 0;  // bytes in free_
    static constexpr size_t kDeviceCacheBytes = 128ull << 30;
    static constexpr size_t kPinnedCacheBytes = 16ull << 30;
    static constexpr size_t kMinBlock = 64u << 10;
    static constexpr int kOrders = 15;  // 64 KiB << 14 = 1 GiB
    static constexpr size_t kSlabBytes = kMinBlock << (kOrders - 1);
    static constexpr size_t kSlabCapBytes = 48ull << 30;
    std::mutex hmu_;
    std::multimap<size_t, void *> hfree_;
    std::map<void *, size_t> hlive_;
    std::map<void *, size_t> hpage_;  // ordinary memory handed out in place of pinned
    size_t hcached_ = 0;
    // buddy arena (under hmu_)
    std::set<char *> bfree_[kOrders];
    std::unordered_map<char *, int> blive_;  // live block -> order
    std::map<char *, size_t> slabs_;         // slab base -> bytes
    size_t slab_bytes_ = 0;
    bool add_slab_locked();
    char *buddy_take_locked(int order);
};

// RAII device buffer from the cache.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { reset(bytes); }
    ~DevBuf() { reset(0); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { reset(0); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void reset(size_t bytes) {
        if (p_) Device::get().release(p_);
        p_ = nullptr;
        n_ = 0;
        if (bytes) { p_ = Device::get().alloc(bytes); n_ = bytes; }
    }
    // Grow-only reuse: keeps the buffer when it is already big enough.
    void ensure(size_t bytes) { if (bytes > n_) reset(bytes); }
    template <typename T> T *as() const { return static_cast<T *>(p_); }
    size_t size() const { return n_; }
    void swap(DevBuf &o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }

  private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// Allocator whose resize(n) leaves new elements uninitialised (resize(n, v)
// still fills): for large buffers that are written in full right after.
template <class T> struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) {}
    template <class U, class... A> void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
    template <class U> void construct(U *p) { ::new ((void *)p) U; }
};

// ---------------------------------------------------------------------------
// Packed reads in host memory: 16 bases per u32 word, fragments back to back.
struct HostReads {
    std::vector<uint32_t, NoInitAlloc<uint32_t>> words;  // padded with zero words at the end
    std::vector<uint64_t> off{0};    // fragment boundaries (n_frag + 1)
    std::vector<uint32_t> dataset;   // dataset of each fragment (may be empty)
    std::vector<uint64_t> chunk_frag{0}; // fragment index where each input chunk starts
    std::vector<uint64_t> input_frag{0}; // fragment index where each input (file) starts
    uint32_t n_datasets = 1;
    uint64_t n_bases() const { return off.back(); }
    uint64_t n_frag() const { return off.size() - 1; }
    void append_base(uint32_t code) {
        uint64_t n = cur_;
        if ((n >> 4) >= words.size()) words.push_back(0);
        words[n >> 4] |= code << ((n & 15u) * 2u);
        cur_++;
    }
    // Close the fragment being appended; drops it when shorter than min_len.
    void close_fragment(uint64_t min_len, uint32_t ds) {
        uint64_t start = off.back();
        if (cur_ > start && cur_ - start >= min_len) {
            off.push_back(cur_);
            dataset.push_back(ds);
        } else {
            // drop it: clear the bits written for its bases
            for (uint64_t i = start; i < cur_; i++) words[i >> 4] &= ~(3u << ((i & 15u) * 2u));
            cur_ = start;
            words.resize((cur_ + 15) / 16);
        }
    }
    void end_chunk() {
        if (chunk_frag.back() != n_frag()) chunk_frag.push_back(n_frag());
    }
    // End of one input (file); an input with no fragment leaves no entry.
    void end_input() {
        end_chunk();
        if (input_frag.back() != n_frag()) input_frag.push_back(n_frag());
    }
    void finish() {
        words.resize((cur_ + 15) / 16 + 32, 0);
        end_chunk();
    }
    uint64_t cur_ = 0;
};

// Empty fragments (frag_off[i] == frag_off[i + 1]) hold no window, but the
// device scan assumes at most one fragment start per base position (it keeps
// each window's fragment as a small offset from its thread's first). Returns
// the indices of the non-empty fragments when some are empty, else an empty
// vector (nothing to drop).
inline std::vector<uint64_t> nonempty_fragments(const uint64_t *frag_off, size_t n_frag) {
    size_t empty = 0;
    for (size_t i = 0; i < n_frag; i++) empty += frag_off[i + 1] == frag_off[i];
    std::vector<uint64_t> keep;
    if (!empty) return keep;
    keep.reserve(n_frag - empty);
    for (size_t i = 0; i < n_frag; i++)
        if (frag_off[i + 1] != frag_off[i]) keep.push_back(i);
    return keep;
}

// Parse FASTA/FASTQ (headers and qualities ignored, split at non-ACGT,
// fragments shorter than min_len dropped; src/lib.rs:3264-3308).
void parse_fastx_file(const std::string &path, uint64_t min_len, uint32_t dataset, HostReads &out);

// ---- compression (compress.cpp) ----
enum class InputCompression { None, Gzip, Xz, Zstd };
// By magic bytes: gzip 1f 8b, xz fd '7zXZ' 00, zstd 28 b5 2f fd.
InputCompression input_compression(const uint8_t *p, size_t len);
void decompress_input(const uint8_t *p, size_t len, InputCompression c, const std::string &what, std::string &out);
// 1 gzip, 2 xz, 3 zstd by the path's extension, else 0.
int output_compression_kind(const std::string &path);
// Streams compressed output of one kind into an open file.
class Compressor {
  public:
    Compressor(FILE *fp, const std::string &path, int kind);
    ~Compressor();
    Compressor(const Compressor &) = delete;
    Compressor &operator=(const Compressor &) = delete;
    void write(const char *p, size_t n, bool finish);

  private:
    struct Impl;
    Impl *m_;
};
void parse_fastx_text(const char *text, size_t len, uint64_t min_len, uint32_t dataset,
                      HostReads &out);

// ---------------------------------------------------------------------------
// Device-resident segmented read buffer and the MC counting state.
struct McPlan;

struct Batch {
    DevBuf words, off, dataset;
    uint64_t n_bases = 0, n_frag = 0, n_words = 0;
    uint32_t n_datasets = 1;
    bool has_dataset = false;
    std::vector<uint64_t> chunk_frag;
    // fragment lengths: how many of each length below 64, and the count and
    // base total of the longer ones (window count for any k <= 64)
    uint64_t len_hist[64] = {};
    uint64_t long_cnt = 0, long_sum = 0;
    // Inputs (files, or the caller's input marks): fragment boundaries (each
    // input starts a chunk), the dataset of each input when all its
    // fragments share one (else NO_DATASET), and per input the fragment
    // length summary above ([0..63] lengths below 64, [64] count and [65]
    // bases of the longer ones).
    static constexpr uint32_t NO_DATASET = 0xFFFFFFFFu;
    std::vector<uint64_t> input_frag;
    std::vector<uint32_t> input_ds;
    std::vector<std::array<uint64_t, 66>> input_hist;
    // tile schedule (built with the batch) and per-run state
    std::unique_ptr<McPlan> plan;
    // results of the last count (device)
    DevBuf out_kmer, out_count, out_part, out_hi;  // out_hi: high words for k > 32
    uint64_t n_out = 0;
    uint32_t k = 0;
    // (lo, hi) -> what the 32-bit device count of that k-mer lacks (counts
    // of 2^32 and more; added on download)
    std::map<std::pair<uint64_t, uint64_t>, uint64_t> carries;
    mczi_run_stats last{};
    ~Batch();
};

std::unique_ptr<Batch> make_batch(HostReads &&reads);
std::unique_ptr<Batch> make_batch_packed(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                                         const uint32_t *dataset, uint32_t n_datasets,
                                         std::vector<uint64_t> &&chunk_frag, std::vector<uint64_t> &&input_frag);
void batch_count(Batch &b, const mczi_counter_config &cfg, int mode, int launch, int order,
                 mczi_run_stats *stats);
void batch_download(Batch &b, std::vector<mczi_counted_kmer> &out);
void batch_download(Batch &b, mczi_counted_kmer *out);  // b.n_out records
// The super-k-mer records of the last count, sorted by internal partition:
// rec_words u64 words each (2 for k <= 32: bases 0-55 at 2 bits, MC codes,
// base count at bits 112-119; 4 for k > 32: bases 0-111, count at bits
// 240-247), with each record's internal partition (of p_int; the user
// partition is that modulo the partition count).
void batch_download_records(Batch &b, std::vector<uint64_t> &rec, std::vector<uint32_t> &part, uint32_t &rec_words,
                            uint32_t &p_int);
// Only the emitted k-mers' low (and, for k > 32, high) words.
void batch_download_keys(Batch &b, uint64_t *lo, uint64_t *hi);
double batch_autotune(Batch &b, const mczi_counter_config &cfg, int mode, int launch, int order);
// Pass 1 alone (minimizer counts): the number of distinct minimizer keys.
uint64_t batch_minimizer_count(Batch &b, const mczi_counter_config &cfg);
void free_plan(McPlan *p);

// Plain device copy of `bytes` (average ms per launch over `reps`).
double device_copy_probe(uint64_t bytes, int reps);

// The same without any device allocation (a fixed file-scope scratch array;
// n_bases <= 2^24, the caller checks the arguments).
void device_run_hashes_noalloc(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m, uint32_t *hashes,
                               uint32_t *starts, size_t *n_runs);
constexpr uint64_t NOALLOC_RUN_HASHES_MAX_BASES = 1ull << 24;

// Device scan boundary for one sequence (run_hashes).
void device_run_hashes(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m,
                       std::vector<uint32_t> &hashes, std::vector<uint32_t> &starts);

// ---------------------------------------------------------------------------
// Output writers (host end of the path).
void write_simplitig_fasta(const std::string &path, const mczi_counted_kmer *kmers, size_t n,
                           uint32_t k, size_t first_index, size_t *last_index);
void write_kmer_fasta(const std::string &path, const std::vector<mczi_counted_kmer> &kmers,
                      uint32_t k);
void write_kff(const std::string &path, const std::vector<mczi_counted_kmer> &kmers,
               const mczi_counter_config &cfg, uint64_t count_size, bool ordered,
               uint64_t saturate_at);
std::string simplitig_fasta_text(const mczi_counted_kmer *kmers, size_t n, uint32_t k);
// simplitig_sequences (src/lib.rs:1209-1211): the greedy's simplitigs as
// ASCII strings, in its order.
std::vector<std::string> simplitig_sequences(const mczi_counted_kmer *kmers, size_t n, uint32_t k);
// simplitig_sequences_from_kmers_parallel (src/lib.rs:1217-1264): k-mers split
// by canonical minimizer into partition_count partitions (0: the reference's
// default), the greedy run on each, results concatenated in partition order.
std::vector<std::string> simplitig_sequences_parallel(const mczi_counted_kmer *kmers, size_t n, uint32_t k,
                                                      uint32_t minimizer, uint32_t partition_count);
// Greedy simplitig cover (src/lib.rs:1478-1533) of `kmers` (any order,
// duplicates collapse) as FASTA text handed to `put` in chunks; headers count
// from first_index + 1. Returns the last index written.
size_t emit_simplitigs(const mczi_counted_kmer *kmers, size_t n, uint32_t k, size_t first_index,
                       const std::function<void(const char *, size_t)> &put);
// The same from n keys that fill(lo, hi) writes as 64-bit words (hi is null
// for k <= 32).
// link_fill (optional) supplies the forced-link words instead of the host
// neighbour pass: device_forced_links for keys resident in HBM.
size_t emit_simplitigs_keys(size_t n, uint32_t k, const std::function<void(uint64_t *, uint64_t *)> &fill,
                            size_t first_index, const std::function<void(const char *, size_t)> &put,
                            const std::function<void(uint64_t *)> &link_fill = nullptr);
void write_simplitig_fasta_keys(const std::string &path, size_t n, uint32_t k,
                                const std::function<void(uint64_t *, uint64_t *)> &fill,
                                const std::function<void(uint64_t *)> &link_fill = nullptr);
// Forced links of n sorted unique k-mers in HBM (lo; hi for k > 32) into
// link_host[0 .. 2n) (graph_kernels.hip).
void device_forced_links(const unsigned long long *lo, const unsigned long long *hi, uint64_t n, uint32_t k,
                         uint64_t *link_host);

// ---------------------------------------------------------------------------
// On-disk formats of the reference's partitioned path (interop.cpp).
// Packed-read cache (.pread): "MCRD0001", then per fragment a u32 BE length
// and ceil(len / 4) bytes of packed-seq codes, 4 bases per byte from bit 0.
void write_read_cache(const std::string &path, const HostReads &r);
void read_read_cache(const std::string &path, uint64_t min_len, HostReads &out);
// Super-k-mer partition files (partition_NNNN.skm): per record a length
// (one byte, or 0xFF and a u32 BE) and its packed bases.
void write_superkmer_partitions(const std::string &dir, uint32_t partitions, const std::vector<uint64_t> &rec,
                                const std::vector<uint32_t> &part, uint32_t rec_words);
void read_superkmer_file(const std::string &path, uint64_t min_len, HostReads &out);
// Counted k-mer records (.cnt): u128 BE k-mer, u64 BE count.
void write_counted_file(const std::string &path, const mczi_counted_kmer *kmers, size_t n);
std::vector<mczi_counted_kmer> read_counted_file(const std::string &path);

// Phase log lines on stderr, in the reference's formats (src/lib.rs:994-1062).
void log_phase(const char *name, double seconds);
// "phaseN detail" of a phase name (display_phase_name, src/lib.rs:1027-1035).
std::string phase_display_name(const char *name);

}  // namespace mczi

namespace mczi {
// Synthetic inputs (synth.cpp)
void synth_generate(const mczi_synth_spec &spec, mczi_synth_reads &out);
void synth_write_fasta(const mczi_synth_reads &r, const std::string &path, int64_t ds_filter);
void synth_genome_words(uint64_t len, uint64_t seed, uint32_t **words);
void synth_reads_from(const uint32_t *const *gw, const uint64_t *gl, uint32_t ng, uint64_t n_reads,
                      uint32_t read_len, double sub, uint64_t seed, mczi_synth_reads &out);
}  // namespace mczi
