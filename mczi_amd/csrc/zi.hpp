// ZI: device-resident k-mer index, streaming subtraction, serving loop.
#pragma once

#include <condition_variable>
#include <deque>
#include <thread>

#include "internal.hpp"

namespace mczi {

// Canonical k-mer set of an index, in HBM: open addressing, linear probing.
// k <= 32: one 64-bit key per slot (empty = all ones); k <= 63: 16 bytes per
// slot, the high word then the low word, the slot claimed through the high
// word (a canonical k-mer of k <= 63 has hi < 2^62), so that one read brings
// both words of a probed slot.
struct ZiIndex {
    uint32_t k = 0, m = 0;
    bool wide = false;
    DevBuf slots;
    uint64_t mask = 0;
    uint64_t n_kmers = 0;  // index windows inserted
};

void zi_build_index(ZiIndex &idx, HostReads &&reads, uint32_t k, uint32_t m, hipStream_t s);
// The index from n sorted unique canonical k-mers already in HBM (lo; hi for
// k > 32), e.g. the output of an MC count.
void zi_build_index_from_keys(ZiIndex &idx, const unsigned long long *lo, const unsigned long long *hi, uint64_t n,
                              uint32_t k, uint32_t m, hipStream_t s);

// One batch of query requests (fragments tagged with their request slot).
// Sorted, unique absent k-mers of one request, in pinned host arrays from
// the device's host pool (ownership passes to the caller's result, which
// returns them with Device::host_release).
struct ZiKeys {
    uint64_t *lo = nullptr, *hi = nullptr;
    size_t n = 0;
    ZiKeys() = default;
    ZiKeys(const ZiKeys &) = delete;
    ZiKeys &operator=(const ZiKeys &) = delete;
    ZiKeys(ZiKeys &&o) noexcept : lo(o.lo), hi(o.hi), n(o.n) { o.lo = o.hi = nullptr; o.n = 0; }
    ZiKeys &operator=(ZiKeys &&o) noexcept {
        if (this != &o) {
            release();
            lo = o.lo;
            hi = o.hi;
            n = o.n;
            o.lo = o.hi = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~ZiKeys() { release(); }
    void release() {
        if (lo || hi) {
            Device &d = Device::get();
            d.host_release(lo);
            d.host_release(hi);
        }
        lo = hi = nullptr;
        n = 0;
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
    std::vector<uint64_t> absent_windows;  // per request, repeats counted
    double wait_ms = 0;             // host wait for the batch's device work
    double alloc_ms = 0;            // host time spent allocating the result arrays
};

// Device memory a batch of `bases` query bases in `frags` fragments needs.
uint64_t zi_batch_bytes(uint64_t bases, uint64_t frags, bool wide);

struct ZiWork;  // device buffers of one in-flight batch
std::unique_ptr<ZiWork> zi_make_work();
void zi_free_work(ZiWork *w);
// A batch of query requests laid out in one device word buffer: fragment
// offsets in bases (requests start on word boundaries; a gap before a
// request is a fragment of no request), the request slot of each fragment,
// and the host word ranges copied to their place (pinned sources copy as
// plain DMA). Sources must stay valid until the batch is collected.
struct ZiWordSpan {
    const uint32_t *src;
    uint64_t n_words;
    uint64_t dst_word;
};
struct ZiBatchInput {
    std::vector<uint64_t> off{0};
    std::vector<uint32_t> frag_req;
    std::vector<ZiWordSpan> spans;
    uint64_t n_words = 0;
};
// Enqueue upload + lookup on stream s (async).
void zi_enqueue(ZiWork &w, const ZiIndex &idx, ZiBatchInput &&in, uint32_t n_req, hipStream_t s);
// Wait for the lookup, then sort and deduplicate each request's absent
// k-mers on the device; returns with their counts known (host wait in ms).
double zi_finish(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s);
// Copy each request's unique absent k-mers out (after zi_finish).
void zi_copy_out(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s, ZiQueryResult &out);
// zi_finish + zi_copy_out.
void zi_collect(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s, ZiQueryResult &out);

// All query files in one batch against idx: query windows, absent windows
// (repeats counted) and the sorted unique absent k-mers.
struct ZiScan {
    ZiKeys absent;
    uint64_t windows = 0, absent_windows = 0;
};
ZiScan zi_subtract_files(const ZiIndex &idx, const char *const *qp, size_t n, hipStream_t s);

}  // namespace mczi
