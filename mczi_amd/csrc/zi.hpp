// ZI: device-resident k-mer index, streaming subtraction, serving loop.
#pragma once

#include <condition_variable>
#include <deque>
#include <thread>

#include "internal.hpp"

namespace mczi {

// Canonical k-mer set of an index, in HBM: open addressing, linear probing.
// k <= 32: 64-bit keys in `lo` (empty = all ones); k <= 63: (hi, lo) pairs,
// slot claimed through `hi` (a canonical k-mer of k <= 63 has hi < 2^62).
struct ZiIndex {
    uint32_t k = 0, m = 0;
    bool wide = false;
    DevBuf lo, hi;
    uint64_t mask = 0;
    uint64_t n_kmers = 0;  // index windows inserted
};

void zi_build_index(ZiIndex &idx, HostReads &&reads, uint32_t k, uint32_t m, hipStream_t s);

// One batch of query requests (fragments tagged with their request slot).
// Sorted, unique absent k-mers of one request, in malloc'd arrays (ownership
// passes to the caller's result).
struct ZiKeys {
    uint64_t *lo = nullptr, *hi = nullptr;
    size_t n = 0;
    ZiKeys() = default;
    ZiKeys(const ZiKeys &) = delete;
    ZiKeys &operator=(const ZiKeys &) = delete;
    ZiKeys(ZiKeys &&o) noexcept : lo(o.lo), hi(o.hi), n(o.n) { o.lo = o.hi = nullptr; o.n = 0; }
    ZiKeys &operator=(ZiKeys &&o) noexcept {
        if (this != &o) {
            free(lo);
            free(hi);
            lo = o.lo;
            hi = o.hi;
            n = o.n;
            o.lo = o.hi = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~ZiKeys() {
        free(lo);
        free(hi);
    }
    std::vector<mczi_counted_kmer> counted() const {
        std::vector<mczi_counted_kmer> v(n);
        for (size_t i = 0; i < n; i++) v[i] = mczi_counted_kmer{lo[i], hi ? hi[i] : 0, 1};
        return v;
    }
};

struct ZiQueryResult {
    std::vector<ZiKeys> absent;     // per request
    std::vector<uint64_t> windows;  // per request
};

// Device memory a batch of `bases` query bases in `frags` fragments needs.
uint64_t zi_batch_bytes(uint64_t bases, uint64_t frags, bool wide);

struct ZiWork;  // device buffers of one in-flight batch
std::unique_ptr<ZiWork> zi_make_work();
void zi_free_work(ZiWork *w);
// Enqueue upload + lookup + segmented sort + unique on stream s (async).
void zi_enqueue(ZiWork &w, const ZiIndex &idx, HostReads &&q, std::vector<uint32_t> &&frag_req, uint32_t n_req,
                hipStream_t s);
// Wait for the batch and collect per-request results.
void zi_collect(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s, ZiQueryResult &out);

}  // namespace mczi
