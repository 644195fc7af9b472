// Device context, caching allocator, FASTA/FASTQ ingest and phase logs.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <emmintrin.h>
#include <cctype>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>

#include <thread>

#include "internal.hpp"

namespace mczi {

std::atomic<unsigned> g_host_threads{0};

Device &Device::get() {
    static Device d;
    return d;
}

Device::Device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        throw Error(MCZI_E_DEVICE,
                    "no HIP device available: the MCZI device path needs an MI355X (gfx950) GPU");
    MCZI_HIP(hipGetDevice(&dev_));
    hipDeviceProp_t prop;
    MCZI_HIP(hipGetDeviceProperties(&prop, dev_));
    cus_ = prop.multiProcessorCount;
    MCZI_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

static size_t round_alloc(size_t b) {
    const size_t g = 2u << 20;
    return (b + g - 1) / g * g;
}

void *Device::alloc(size_t bytes) {
    const size_t want = round_alloc(bytes);
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(want);
        if (it != free_.end() && it->first <= want * 2) {
            void *p = it->second;
            live_[p] = it->first;
            pooled_ -= it->first;
            free_.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(mu_);
        // free cached blocks, largest first, until the request fits; then all
        size_t freed = 0;
        while (!free_.empty() && freed < want) {
            auto last = std::prev(free_.end());
            freed += last->first;
            pooled_ -= last->first;
            (void)hipFree(last->second);
            free_.erase(last);
        }
        p = nullptr;
        e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            for (auto &kv : free_) (void)hipFree(kv.second);
            free_.clear();
            pooled_ = 0;
            p = nullptr;
            e = hipMalloc(&p, want);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(MCZI_E_NOMEM, "device allocation of " + std::to_string(want) +
                                          " bytes failed: " + hipGetErrorString(e));
        }
    }
    std::lock_guard<std::mutex> lk(mu_);
    live_[p] = want;
    return p;
}

void Device::release(void *p) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    const size_t n = it->second;
    live_.erase(it);
    if (pooled_ + n > kDeviceCacheBytes) {  // the cache is full: give the block back
        (void)hipFree(p);
        return;
    }
    free_.emplace(n, p);
    pooled_ += n;
}

size_t Device::pooled_bytes() {
    std::lock_guard<std::mutex> lk(mu_);
    return pooled_;
}

bool Device::add_slab_locked() {
    if (slab_bytes_ + kSlabBytes > kSlabCapBytes) return false;
    void *p = nullptr;
    if (hipHostMalloc(&p, kSlabBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    slabs_[(char *)p] = kSlabBytes;
    slab_bytes_ += kSlabBytes;
    bfree_[kOrders - 1].insert((char *)p);
    return true;
}

// Smallest free block of at least `order`, split down to it (the upper
// halves go to the free lists); nullptr when no slab can be added.
char *Device::buddy_take_locked(int order) {
    int j = order;
    while (j < kOrders && bfree_[j].empty()) j++;
    if (j == kOrders) {
        if (!add_slab_locked()) return nullptr;
        j = kOrders - 1;
    }
    char *p = *bfree_[j].begin();
    bfree_[j].erase(bfree_[j].begin());
    while (j > order) {
        j--;
        bfree_[j].insert(p + (kMinBlock << j));
    }
    blive_[p] = order;
    return p;
}

void Device::host_reserve(size_t bytes) {
    std::lock_guard<std::mutex> lk(hmu_);
    while (slab_bytes_ < bytes && add_slab_locked()) {
    }
}

size_t Device::host_slab_bytes() {
    std::lock_guard<std::mutex> lk(hmu_);
    return slab_bytes_;
}

void *Device::host_alloc(size_t bytes, bool allow_pageable) {
    size_t want = kMinBlock;
    while (want < bytes) want <<= 1;
    if (want <= kSlabBytes) {
        int order = 0;
        while ((kMinBlock << order) < want) order++;
        std::lock_guard<std::mutex> lk(hmu_);
        if (char *p = buddy_take_locked(order)) return p;
    }
    {
        // a cached block up to 4x the size is reused: pinning fresh pages
        // costs far more than the memory a larger block holds
        std::lock_guard<std::mutex> lk(hmu_);
        auto it = hfree_.lower_bound(want);
        if (it != hfree_.end() && it->first <= 4 * want) {
            void *p = it->second;
            const size_t have = it->first;
            hfree_.erase(it);
            hcached_ -= have;
            hlive_[p] = have;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(hmu_);
            for (auto &kv : hfree_) (void)hipHostFree(kv.second);
            hfree_.clear();
            hcached_ = 0;
        }
        p = nullptr;
        e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (allow_pageable) {
                p = malloc(want);
                if (!p) throw Error(MCZI_E_NOMEM, "host allocation of " + std::to_string(want) + " bytes failed");
                std::lock_guard<std::mutex> lk(hmu_);
                hpage_[p] = want;
                return p;
            }
            throw Error(MCZI_E_NOMEM, "pinned host allocation of " + std::to_string(want) +
                                          " bytes failed: " + hipGetErrorString(e));
        }
    }
    std::lock_guard<std::mutex> lk(hmu_);
    hlive_[p] = want;
    return p;
}

void Device::upload(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    constexpr size_t kChunk = 64u << 20;
    constexpr int kThreads = 8;
    if (bytes <= (4u << 20)) {  // small: the runtime's own staging is fine
        MCZI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        MCZI_HIP(hipStreamSynchronize(s));
        return;
    }
    char *stage[2] = {(char *)host_alloc(kChunk), (char *)host_alloc(kChunk)};
    hipEvent_t ev[2];
    MCZI_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    MCZI_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    bool used[2] = {false, false};
    try {
        for (size_t off = 0, i = 0; off < bytes; off += kChunk, i++) {
            const int b = (int)(i & 1);
            const size_t n = std::min(kChunk, bytes - off);
            if (used[b]) MCZI_HIP(hipEventSynchronize(ev[b]));
            const char *from = (const char *)src + off;
            std::vector<std::thread> th;
            const size_t per = (n + kThreads - 1) / kThreads;
            for (int t = 0; t < kThreads; t++) {
                const size_t a0 = (size_t)t * per;
                if (a0 >= n) break;
                const size_t len = std::min(per, n - a0);
                th.emplace_back([=] { memcpy(stage[b] + a0, from + a0, len); });
            }
            for (auto &x : th) x.join();
            MCZI_HIP(hipMemcpyAsync((char *)dst + off, stage[b], n, hipMemcpyHostToDevice, s));
            MCZI_HIP(hipEventRecord(ev[b], s));
            used[b] = true;
        }
        MCZI_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        host_release(stage[0]);
        host_release(stage[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    host_release(stage[0]);
    host_release(stage[1]);
}

void Device::download(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    constexpr size_t kChunk = 64u << 20;
    constexpr int kThreads = 8;
    if (bytes <= (4u << 20)) {
        MCZI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipStreamSynchronize(s));
        return;
    }
    char *stage[2] = {(char *)host_alloc(kChunk), (char *)host_alloc(kChunk)};
    hipEvent_t ev[2];
    MCZI_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    MCZI_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    auto drain = [&](int b, size_t off) {  // copy chunk `off` out of stage[b]
        MCZI_HIP(hipEventSynchronize(ev[b]));
        const size_t n = std::min(kChunk, bytes - off);
        char *to = (char *)dst + off;
        std::vector<std::thread> th;
        const size_t per = (n + kThreads - 1) / kThreads;
        for (int t = 0; t < kThreads; t++) {
            const size_t a0 = (size_t)t * per;
            if (a0 >= n) break;
            const size_t len = std::min(per, n - a0);
            th.emplace_back([=] { memcpy(to + a0, stage[b] + a0, len); });
        }
        for (auto &x : th) x.join();
    };
    try {
        size_t i = 0;
        for (size_t off = 0; off < bytes; off += kChunk, i++) {
            const int b = (int)(i & 1);
            const size_t n = std::min(kChunk, bytes - off);
            MCZI_HIP(hipMemcpyAsync(stage[b], (const char *)src + off, n, hipMemcpyDeviceToHost, s));
            MCZI_HIP(hipEventRecord(ev[b], s));
            if (i > 0) drain(1 - b, off - kChunk);
        }
        drain((int)((i - 1) & 1), (i - 1) * kChunk);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        host_release(stage[0]);
        host_release(stage[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    host_release(stage[0]);
    host_release(stage[1]);
}

void Device::host_release(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(hmu_);
    auto bl = blive_.find((char *)p);
    if (bl != blive_.end()) {
        int o = bl->second;
        blive_.erase(bl);
        char *q = (char *)p;
        auto sl = slabs_.upper_bound(q);
        --sl;  // the slab holding q
        char *const base = sl->first;
        while (o < kOrders - 1) {
            char *const buddy = base + ((size_t)(q - base) ^ (kMinBlock << o));
            auto f = bfree_[o].find(buddy);
            if (f == bfree_[o].end()) break;
            bfree_[o].erase(f);
            q = std::min(q, buddy);
            o++;
        }
        bfree_[o].insert(q);
        return;
    }
    auto it = hlive_.find(p);
    if (it == hlive_.end()) {
        auto pg = hpage_.find(p);
        if (pg != hpage_.end()) {
            hpage_.erase(pg);
            free(p);
        }
        return;
    }
    const size_t n = it->second;
    hlive_.erase(it);
    if (hcached_ + n > kPinnedCacheBytes) {
        (void)hipHostFree(p);
        return;
    }
    hfree_.emplace(n, p);
    hcached_ += n;
}

// ---------------------------------------------------------------------------
// FASTA / FASTQ ingest. Base codes are packed-seq codes A0 C1 T2 G3; lower
// case acgt count as bases (src/zi.rs:240-248 and src/lib.rs:3015-3026 map
// both cases); any other character ends the current fragment.

namespace {
struct CodeTable {
    uint8_t t[256];
    CodeTable() {
        memset(t, 5, sizeof t);
        t[(int)'A'] = t[(int)'a'] = 0;
        t[(int)'C'] = t[(int)'c'] = 1;
        t[(int)'T'] = t[(int)'t'] = 2;
        t[(int)'G'] = t[(int)'g'] = 3;
        t[(int)'\n'] = t[(int)'\r'] = 4;
    }
};
const CodeTable kCodes;

struct Packer {
    HostReads &r;
    uint64_t min_len;
    uint32_t ds;
    std::vector<uint8_t> frag;
    Packer(HostReads &rr, uint64_t ml, uint32_t d) : r(rr), min_len(ml), ds(d) { frag.reserve(1 << 16); }
    void close() {
        if (!frag.empty() && frag.size() >= min_len) {
            uint64_t n = r.cur_;
            const uint8_t *c = frag.data();
            size_t len = frag.size();
            r.words.resize((n + len + 15) / 16 + 1, 0);
            size_t i = 0;
            // head until word aligned
            while (i < len && ((n + i) & 15u)) {
                r.words[(n + i) >> 4] |= (uint32_t)c[i] << (((n + i) & 15u) * 2u);
                i++;
            }
            for (; i + 16 <= len; i += 16) {
                uint32_t w = 0;
                for (int j = 0; j < 16; j++) w |= (uint32_t)c[i + j] << (2 * j);
                r.words[(n + i) >> 4] = w;
            }
            for (; i < len; i++) r.words[(n + i) >> 4] |= (uint32_t)c[i] << (((n + i) & 15u) * 2u);
            r.cur_ = n + len;
            r.words.resize((r.cur_ + 15) / 16);
            r.off.push_back(r.cur_);
            r.dataset.push_back(ds);
        }
        frag.clear();
    }
};
}  // namespace

// ---------------------------------------------------------------------------
// FASTA on many threads, with the same fragments and bases as the one-pass
// parser below. Pass 1 scans byte chunks on their own into fragment pieces: a
// chunk starts inside a header line when a '>' follows the last '\n' before
// it. The pieces are stitched across chunk ends and filtered by length; pass
// 2 scans again and packs every kept base at its final position.
namespace {

enum : uint8_t { EV_NONE = 0, EV_BASE = 1, EV_CLOSE = 2 };
constexpr uint64_t DROPPED = ~0ull;

struct FastaChunk {
    size_t a = 0, b = 0;
    bool header = false;      // starts inside a header line
    uint8_t first = EV_NONE;  // first base or fragment end seen in the chunk
    bool open_end = false;    // the last piece runs on past the chunk end
    uint64_t frag0 = 0, off0 = 0;  // global fragment of piece 0, its offset in it
    std::vector<uint64_t> piece;   // piece lengths; later, output positions
};

// 16 bytes that are all of ACGTacgt.
inline bool all_bases16(const char *p) {
    const __m128i v = _mm_or_si128(_mm_loadu_si128((const __m128i *)p), _mm_set1_epi8(0x20));
    const __m128i m = _mm_or_si128(_mm_or_si128(_mm_cmpeq_epi8(v, _mm_set1_epi8('a')), _mm_cmpeq_epi8(v, _mm_set1_epi8('c'))),
                                   _mm_or_si128(_mm_cmpeq_epi8(v, _mm_set1_epi8('g')), _mm_cmpeq_epi8(v, _mm_set1_epi8('t'))));
    return _mm_movemask_epi8(m) == 0xFFFF;
}

// Codes (A0 C1 T2 G3, which is (ascii >> 1) & 3 in both cases) of 8 bases,
// base i at bits 2i.
inline uint32_t codes8(const char *p) {
    uint64_t v;
    memcpy(&v, p, 8);
    uint64_t x = (v >> 1) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    return (uint32_t)((x | (x >> 24)) & 0xFFFFu);
}

void fasta_scan(const char *buf, FastaChunk &c) {
    bool hdr = c.header;
    uint64_t cur = 0;
    for (size_t i = c.a; i < c.b; i++) {
        const char ch = buf[i];
        if (hdr) {
            if (ch == '\n') hdr = false;
            continue;
        }
        const uint8_t code = kCodes.t[(uint8_t)ch];
        if (code < 4) {
            if (c.first == EV_NONE) c.first = EV_BASE;
            cur++;
            while (i + 17 <= c.b && all_bases16(buf + i + 1)) {
                cur += 16;
                i += 16;
            }
            continue;
        }
        if (code == 4) continue;
        // '>' and every other character end the fragment
        if (c.first == EV_NONE) c.first = EV_CLOSE;
        if (cur) c.piece.push_back(cur);
        cur = 0;
        if (ch == '>') hdr = true;
    }
    c.open_end = cur > 0;
    if (cur) c.piece.push_back(cur);
}

// Pack the chunk's kept bases; c.piece now holds each piece's output
// position (or DROPPED). Output positions rise through the chunk, so each
// word is gathered in one go: stored when all its 16 bases are this chunk's,
// OR-ed atomically when it is shared with a neighbour.
void fasta_pack(const char *buf, const FastaChunk &c, uint32_t *words) {
    bool hdr = c.header, in_piece = false, keep = false;
    size_t j = 0;
    uint64_t pos = 0, wi = ~0ull;
    uint32_t acc = 0;
    int filled = 0;
    auto flush = [&] {
        if (wi == ~0ull) return;
        if (filled == 16)
            words[wi] = acc;
        else
            __atomic_fetch_or(&words[wi], acc, __ATOMIC_RELAXED);
    };
    // append n <= 8 codes (2 bits each) at pos
    auto put = [&](uint32_t bits, int n) {
        while (n > 0) {
            if ((pos >> 4) != wi) {
                flush();
                wi = pos >> 4;
                acc = 0;
                filled = 0;
            }
            const int at = (int)(pos & 15u), take = std::min(n, 16 - at);
            const uint32_t part = take == 16 ? bits : (bits & ((1u << (2 * take)) - 1u));
            acc |= part << (2 * at);
            filled += take;
            pos += (uint64_t)take;
            bits = take >= 16 ? 0 : bits >> (2 * take);
            n -= take;
        }
    };
    for (size_t i = c.a; i < c.b; i++) {
        const char ch = buf[i];
        if (hdr) {
            if (ch == '\n') hdr = false;
            continue;
        }
        const uint8_t code = kCodes.t[(uint8_t)ch];
        if (code < 4) {
            if (!in_piece) {
                in_piece = true;
                keep = c.piece[j] != DROPPED;
                pos = c.piece[j];
            }
            if (keep) {
                put(code, 1);
                while (i + 17 <= c.b && all_bases16(buf + i + 1)) {
                    put(codes8(buf + i + 1), 8);
                    put(codes8(buf + i + 9), 8);
                    i += 16;
                }
            } else {
                while (i + 17 <= c.b && all_bases16(buf + i + 1)) i += 16;
            }
            continue;
        }
        if (code == 4) continue;
        if (in_piece) j++;
        in_piece = false;
        if (ch == '>') hdr = true;
    }
    flush();
}

size_t parse_chunk_bytes(size_t len, unsigned nt) {
    if (const char *e = getenv("MCZI_PARSE_CHUNK_BYTES")) {
        const long long v = atoll(e);
        if (v > 0) return (size_t)v;
    }
    return std::max<size_t>(size_t(4) << 20, len / (4 * (size_t)nt) + 1);
}

// buf[0] is the first record's '>'.
void parse_fasta_parallel(const char *buf, size_t len, uint64_t min_len, uint32_t dataset,
                          HostReads &out, size_t chunk) {
    const unsigned nt = host_workers(len);
    std::vector<FastaChunk> ch((len + chunk - 1) / chunk);
    for (size_t c = 0; c < ch.size(); c++) {
        ch[c].a = c * chunk;
        ch[c].b = std::min(len, (c + 1) * chunk);
    }
    parallel_ranges(ch.size(), 1, nt, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; c++) {
            FastaChunk &k = ch[c];
            if (k.a > 0) {
                const char *nl = (const char *)memrchr(buf, '\n', k.a);
                const size_t from = nl ? (size_t)(nl - buf) + 1 : 0;
                k.header = from < k.a && memchr(buf + from, '>', k.a - from) != nullptr;
            }
            fasta_scan(buf, k);
        }
    });
    // stitch pieces into fragments
    size_t n_piece = 0;
    for (auto &k : ch) n_piece += k.piece.size();
    std::vector<uint64_t> flen;
    flen.reserve(n_piece);
    out.off.reserve(out.off.size() + n_piece);
    out.dataset.reserve(out.dataset.size() + n_piece);
    bool open = false;
    for (auto &k : ch) {
        if (k.piece.empty()) {
            if (k.first == EV_CLOSE) open = false;
            continue;
        }
        size_t j0 = 0;
        if (open && k.first == EV_BASE) {
            k.frag0 = flen.size() - 1;
            k.off0 = flen.back();
            flen.back() += k.piece[0];
            j0 = 1;
        } else {
            k.frag0 = flen.size();
        }
        for (size_t j = j0; j < k.piece.size(); j++) flen.push_back(k.piece[j]);
        open = k.open_end;
    }
    // keep fragments of at least min_len bases, placed after out.cur_:
    // flen[f] becomes the fragment's output position (or DROPPED)
    const size_t nf = flen.size();
    const size_t nblk = std::max<size_t>(1, std::min<size_t>(4 * nt, nf / 65536 + 1));
    const size_t bs = (nf + nblk - 1) / nblk;
    std::vector<uint64_t> blen(nblk + 1, 0), bcnt(nblk + 1, 0);
    parallel_ranges(nblk, 1, nt, [&](size_t b0, size_t b1) {
        for (size_t b = b0; b < b1; b++) {
            uint64_t l = 0, c = 0;
            for (size_t f = b * bs; f < std::min(nf, (b + 1) * bs); f++)
                if (flen[f] >= min_len) {
                    l += flen[f];
                    c++;
                }
            blen[b + 1] = l;
            bcnt[b + 1] = c;
        }
    });
    for (size_t b = 0; b < nblk; b++) {
        blen[b + 1] += blen[b];
        bcnt[b + 1] += bcnt[b];
    }
    const uint64_t start = out.cur_, at = start + blen[nblk];
    const size_t off0 = out.off.size();
    out.off.resize(off0 + bcnt[nblk]);
    out.dataset.resize(out.dataset.size() + bcnt[nblk], dataset);
    parallel_ranges(nblk, 1, nt, [&](size_t b0, size_t b1) {
        for (size_t b = b0; b < b1; b++) {
            uint64_t p = start + blen[b];
            size_t o = off0 + bcnt[b];
            for (size_t f = b * bs; f < std::min(nf, (b + 1) * bs); f++) {
                if (flen[f] >= min_len) {
                    const uint64_t l = flen[f];
                    flen[f] = p;
                    p += l;
                    out.off[o++] = p;
                } else {
                    flen[f] = DROPPED;
                }
            }
        }
    });
    // pieces: output positions, and each chunk's output range
    std::vector<uint64_t> lo(ch.size(), DROPPED), hi(ch.size(), 0);
    parallel_ranges(ch.size(), 1, nt, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; c++) {
            FastaChunk &k = ch[c];
            for (size_t j = 0; j < k.piece.size(); j++) {
                const uint64_t f = flen[k.frag0 + j], len = k.piece[j];
                if (f == DROPPED) {
                    k.piece[j] = DROPPED;
                    continue;
                }
                const uint64_t p = f + (j == 0 ? k.off0 : 0);
                k.piece[j] = p;
                if (lo[c] == DROPPED) lo[c] = p;
                hi[c] = p + len;
            }
        }
    });
    std::vector<uint64_t>().swap(flen);
    // new words start uninitialised: words a chunk fills alone are stored
    // whole; the (new) words at each chunk's two ends are OR-ed into, so
    // they start at zero
    const size_t old_words = out.words.size();
    out.words.resize((at + 15) / 16);
    uint32_t *words = out.words.data();
    for (size_t c = 0; c < ch.size(); c++) {
        if (lo[c] == DROPPED) continue;
        for (const uint64_t w : {lo[c] >> 4, (hi[c] - 1) >> 4})
            if (w >= old_words) words[w] = 0;
    }
    parallel_ranges(ch.size(), 1, nt, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; c++) fasta_pack(buf, ch[c], words);
    });
    out.cur_ = at;
}

}  // namespace

static void parse_buffer(const char *buf, size_t len, uint64_t min_len, uint32_t dataset,
                         HostReads &out, const std::string &what) {
    size_t i = 0;
    while (i < len && (buf[i] == '\n' || buf[i] == '\r' || buf[i] == ' ' || buf[i] == '\t')) i++;
    if (i >= len) return;
    if (buf[i] != '>' && buf[i] != '@') {
        fprintf(stderr, "MC_WARN\tskipped_invalid_fastx_input\tpath\t%s\terror\t%s\n", what.c_str(),
                "Invalid record start");
        return;
    }
    if (buf[i] == '>') {
        const size_t chunk = parse_chunk_bytes(len - i, host_workers(len - i));
        if (len - i > chunk) {
            parse_fasta_parallel(buf + i, len - i, min_len, dataset, out, chunk);
            return;
        }
    }
    Packer pk(out, min_len, dataset);
    const bool fastq = buf[i] == '@';
    while (i < len) {
        // header line
        if (buf[i] != (fastq ? '@' : '>')) {
            // stray line outside a record: skip it
            const char *nl = (const char *)memchr(buf + i, '\n', len - i);
            i = nl ? (size_t)(nl - buf) + 1 : len;
            continue;
        }
        const char *nl = (const char *)memchr(buf + i, '\n', len - i);
        i = nl ? (size_t)(nl - buf) + 1 : len;
        size_t seq_len = 0;
        // sequence lines
        while (i < len && buf[i] != (fastq ? '+' : '>')) {
            const uint8_t c = kCodes.t[(uint8_t)buf[i]];
            if (c < 4) {
                pk.frag.push_back(c);
                seq_len++;
            } else if (c == 5) {
                pk.close();
                seq_len++;
            }
            i++;
        }
        pk.close();
        if (fastq && i < len) {
            // '+' line, then quality characters matching the sequence length
            const char *pl = (const char *)memchr(buf + i, '\n', len - i);
            i = pl ? (size_t)(pl - buf) + 1 : len;
            size_t q = 0;
            while (i < len && q < seq_len) {
                if (buf[i] != '\n' && buf[i] != '\r') q++;
                i++;
            }
            while (i < len && (buf[i] == '\n' || buf[i] == '\r')) i++;
        }
    }
    pk.close();
}

void parse_fastx_text(const char *text, size_t len, uint64_t min_len, uint32_t dataset,
                      HostReads &out) {
    parse_buffer(text, len, min_len, dataset, out, "<memory>");
    out.end_input();
}

void parse_fastx_file(const std::string &path, uint64_t min_len, uint32_t dataset, HostReads &out) {
    int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0)
        throw Error(MCZI_E_IO, "failed to open FASTA/FASTQ input " + path + ": " + strerror(errno));
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        throw Error(MCZI_E_IO, "failed to stat " + path);
    }
    size_t len = (size_t)st.st_size;
    if (len == 0) {
        close(fd);
        out.end_input();
        return;
    }
    void *map = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) throw Error(MCZI_E_IO, "failed to map " + path);
    madvise(map, len, MADV_SEQUENTIAL);
    const char *buf = (const char *)map;
    // gzip / xz / zstd inputs are decompressed whole into host memory first
    const InputCompression zc = input_compression((const uint8_t *)buf, len);
    if (zc != InputCompression::None) {
        std::string text;
        try {
            decompress_input((const uint8_t *)buf, len, zc, path, text);
        } catch (...) {
            munmap(map, len);
            throw;
        }
        munmap(map, len);
        parse_buffer(text.data(), text.size(), min_len, dataset, out, path);
    } else {
        // Chunk boundaries every >= 128 Mi bases (BATCH_BASES, src/lib.rs:28)
        // so the plain launch mode can issue one launch per chunk.
        parse_buffer(buf, len, min_len, dataset, out, path);
        munmap(map, len);
    }
    // split this file's fragments into chunks of >= 128 Mi bases
    const uint64_t batch = 128ull << 20;
    uint64_t first = out.chunk_frag.back();
    uint64_t acc = 0;
    for (uint64_t f = first; f < out.n_frag(); f++) {
        acc += out.off[f + 1] - out.off[f];
        if (acc >= batch && f + 1 < out.n_frag()) {
            out.chunk_frag.push_back(f + 1);
            acc = 0;
        }
    }
    out.end_input();
}

// ---------------------------------------------------------------------------
// "MC {phase} {s:.3}s" (src/lib.rs:994-1000, 1027-1062)
static std::string compact_detail(const std::string &n) {
    static const char *map[][2] = {
        {"minimizer_counting", "minimizer count"},
        {"index_minimizer_presence_counting", "minimizer count"},
        {"dataset_minimizer_counting", "dataset minimizer count"},
        {"kmer_counting", "kmer count"},
        {"index_kmer_counting", "kmer count"},
        {"dataset_kmer_counting", "dataset kmer count"},
        {"filtered_kmer_counting", "filtered kmer count"},
        {"index_kmer_partition_counting", "kmer partition count"},
        {"superkmer_partitioning", "superkmer partition"},
        {"output_streaming", "output"},
        {"kmer_counting_and_output", "kmer count output"},
        {"simplitig_output", "simplitig output"},
        {"ggcat_simplitigs", "ggcat simplitigs"},
        {"sshash_indexing", "sshash index"},
        {"query_subtraction", "query subtraction"},
        {"query_subtraction_no_output", "query scan no output"},
        {"query_subtraction_regular_output", "regular output"},
        {"regular_output", "regular output"},
        {"ggcat_simplitig_output", "ggcat output"},
        {"reform_output", "reform output"},
        {"no_output", "no output"},
        {"fasta_total", "fasta total"},
        {"phase12_total", "phase12 total"},
        {"total", "total"},
    };
    for (auto &e : map)
        if (n == e[0]) return e[1];
    return n;
}

std::string phase_display_name(const char *name) {
    std::string n(name);
    size_t us = n.find('_');
    if (us != std::string::npos && us > 0 && isdigit((unsigned char)n[0])) {
        bool alnum = true;
        for (size_t i = 0; i < us; i++) alnum = alnum && isalnum((unsigned char)n[i]);
        if (alnum) return "phase" + n.substr(0, us) + " " + compact_detail(n.substr(us + 1));
    }
    return compact_detail(n);
}

void log_phase(const char *name, double seconds) {
    fprintf(stderr, "MC %s %.3fs\n", phase_display_name(name).c_str(), seconds);
}

}  // namespace mczi
