// ZI on gfx950: device-resident canonical k-mer index of the index FASTA,
// streaming lookup of every query window, absent set per request (sorted,
// deduplicated), and the serving loop (continuous batching under an HBM
// budget). Reference: src/zi.rs:114-189 (build_dictionary,
// collect_query_kmers_absent_from_index, write_novel_simplitigs).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <set>

#include "device_common.hpp"
#include "zi.hpp"
#include <type_traits>

namespace mczi {

constexpr int ZI_THREADS = 256;
constexpr int ZI_WPT = 8;
constexpr int ZI_TILE = ZI_THREADS * ZI_WPT;
constexpr int ZI_WORDS = (15 + ZI_TILE + 64 + 32) / 16 + 4;  // tile bases + k + extraction slack
constexpr uint32_t NO_REQ = 0xffffffffu;

struct Key2 {
    uint64_t lo, hi;
};

// Canonical MC encoding of the k-mer at base position p (k <= 64).
// Packed-seq codes -> MC bits (code ^ code>>1), lanes reversed for the
// forward encoding (first base most significant); the reverse complement is
// the complemented lanes in stored order (src/zi.rs:224-238).
__device__ __forceinline__ uint64_t rev2_64(uint64_t x) {
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
}
__device__ __forceinline__ uint64_t to_mc(uint64_t x) { return x ^ ((x >> 1) & 0x5555555555555555ull); }

__device__ __forceinline__ uint64_t canon64(const uint32_t *words, uint64_t p, int k) {
    const uint64_t mask = k >= 32 ? ~0ull : ((1ull << (2 * k)) - 1ull);
    const uint64_t x = to_mc(extract32_bases(words, p)) & mask;
    const uint64_t fwd = rev2_64(x) >> (64 - 2 * k);
    const uint64_t rc = (~x) & mask;
    return fwd < rc ? fwd : rc;
}

__device__ __forceinline__ Key2 canon128(const uint32_t *words, uint64_t p, int k) {
    // bases 0..31 in a, 32..k-1 in b (k > 32)
    const int kb = k - 32;
    const uint64_t mb = kb >= 32 ? ~0ull : ((1ull << (2 * kb)) - 1ull);
    const uint64_t a = to_mc(extract32_bases(words, p));
    const uint64_t b = to_mc(extract32_bases(words, p + 32)) & mb;
    // forward: 128-bit lane reversal of (b:a) then >> (128 - 2k)
    uint64_t fhi = rev2_64(a), flo = rev2_64(b);
    const int sh = 128 - 2 * k;  // 2..62
    Key2 f;
    f.lo = (flo >> sh) | (fhi << (64 - sh));
    f.hi = fhi >> sh;
    // reverse complement: complement of (b:a) in stored order
    Key2 r;
    r.lo = ~a;
    r.hi = (~b) & mb;
    const bool fl = f.hi < r.hi || (f.hi == r.hi && f.lo < r.lo);
    return fl ? f : r;
}

__device__ __forceinline__ uint64_t slot_of(uint64_t lo, uint64_t hi) {
    return spread_hash_u64(lo ^ spread_hash_u64(hi + 0x9e3779b97f4a7c15ull));
}

struct ZiTileArgs {
    const uint32_t *words;
    const uint64_t *off;
    const uint32_t *frag_req;  // null for the index build
    const uint64_t *tile_start;
    const uint32_t *tile_frag;
    const uint32_t *tile_len;
    uint64_t n_tiles, n_frag;
    int k;
    unsigned long long *tlo;  // k <= 32: the key of each slot
    ulonglong2 *tw;           // k > 32: (hi, lo) of each slot
    uint64_t mask;
    // lookup outputs
    unsigned long long *seg_cursor;
    unsigned long long *out_lo;
    unsigned long long *out_hi;
    unsigned long long *ctr;  // [0] lookups
};

__device__ int load_fragments(const ZiTileArgs &a, uint64_t ts, int tl, uint64_t f0, int32_t *s_fs) {
    int nf = 0;
    for (int j0 = 0;; j0 += ZI_THREADS) {
        const int j = j0 + (int)threadIdx.x;
        bool in = false;
        if (j <= ZI_TILE + 1 && f0 + (uint64_t)j <= a.n_frag) {
            const int64_t rel = (int64_t)a.off[f0 + j] - (int64_t)ts;
            s_fs[j] = rel < 0 ? -1 : (rel > (1 << 30) ? (1 << 30) : (int32_t)rel);
            in = rel < tl;
        }
        const int c = __syncthreads_count(in);
        nf += c;
        if (c < ZI_THREADS) break;
    }
    __syncthreads();
    return nf;
}

__device__ __forceinline__ int frag_of(const int32_t *s_fs, int nf, int r) {
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_fs[mid] <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// k > 32: a slot is claimed by a CAS of its high word, then its low word
// is written (a concurrent insert of the same key may then see the low word
// still empty and claim a second slot, which lookups tolerate).
__device__ __forceinline__ void insert_wide(ulonglong2 *tw, uint64_t mask, const Key2 &key) {
    uint64_t s = slot_of(key.lo, key.hi) & mask;
    for (;;) {
        unsigned long long *w = reinterpret_cast<unsigned long long *>(tw + s);
        const unsigned long long ch = w[0];
        if (ch == key.hi && w[1] == key.lo) return;
        if (ch == EMPTY_KEY && atomicCAS(w, EMPTY_KEY, key.hi) == EMPTY_KEY) {
            w[1] = key.lo;
            return;
        }
        s = (s + 1) & mask;
    }
}

// Index build: insert the canonical k-mer of every window.
template <bool WIDE>
__global__ __launch_bounds__(ZI_THREADS) void k_zi_build(ZiTileArgs a) {
    __shared__ int32_t s_fs[ZI_TILE + 2];
    for (uint64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        const uint64_t ts = a.tile_start[t];
        const int tl = (int)a.tile_len[t];
        __syncthreads();
        const int nf = load_fragments(a, ts, tl, a.tile_frag[t], s_fs);
        for (int i = 0; i < ZI_WPT; i++) {
            const int r = i * ZI_THREADS + (int)threadIdx.x;
            if (r >= tl) continue;
            const int j = frag_of(s_fs, nf, r);
            if (r + a.k > s_fs[j + 1]) continue;
            if constexpr (!WIDE) {
                const unsigned long long key = canon64(a.words, ts + r, a.k);
                uint64_t s = slot_of(key, 0) & a.mask;
                for (;;) {
                    const unsigned long long cur = a.tlo[s];
                    if (cur == key) break;
                    if (cur == EMPTY_KEY) {
                        const unsigned long long prev = atomicCAS(&a.tlo[s], EMPTY_KEY, key);
                        if (prev == EMPTY_KEY || prev == key) break;
                    }
                    s = (s + 1) & a.mask;
                }
            } else {
                const Key2 key = canon128(a.words, ts + r, a.k);
                insert_wide(a.tw, a.mask, key);
            }
        }
    }
}

// Index build from k-mers already in HBM (the counted, sorted set of an MC
// run): the same slots and probing as k_zi_build, one key per thread.
template <bool WIDE>
__global__ void k_zi_insert(const unsigned long long *__restrict__ lo, const unsigned long long *__restrict__ hi,
                            uint64_t n, unsigned long long *tlo, ulonglong2 *tw, uint64_t mask) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if constexpr (!WIDE) {
            const unsigned long long key = lo[i];
            uint64_t s = slot_of(key, 0) & mask;
            for (;;) {
                const unsigned long long cur = tlo[s];
                if (cur == key) break;
                if (cur == EMPTY_KEY) {
                    const unsigned long long prev = atomicCAS(&tlo[s], EMPTY_KEY, key);
                    if (prev == EMPTY_KEY || prev == key) break;
                }
                s = (s + 1) & mask;
            }
        } else {
            insert_wide(tw, mask, Key2{lo[i], hi[i]});
        }
    }
}

// Lookup: every window of every query fragment is looked up in the index;
// absent canonical k-mers are appended to their request's segment. The
// tile's bases are staged in LDS; each thread first forms the keys of its
// ZI_WPT windows, then probes all of them in rounds, so up to ZI_WPT
// independent table reads per thread are in flight.
template <bool WIDE>
__global__ __launch_bounds__(ZI_THREADS) void k_zi_lookup(ZiTileArgs a) {
    __shared__ int32_t s_fs[ZI_TILE + 2];
    __shared__ uint32_t s_words[ZI_WORDS];
    __shared__ uint32_t s_cnt[ZI_WPT * (ZI_THREADS / 64)];  // absent k-mers per (window i, wave)
    __shared__ uint32_t s_rmin[ZI_THREADS / 64], s_rmax[ZI_THREADS / 64];
    __shared__ unsigned long long s_base;
    unsigned long long my_lookups = 0;
    const int lane = threadIdx.x & 63;
    for (uint64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        const uint64_t ts = a.tile_start[t];
        const int tl = (int)a.tile_len[t];
        const uint64_t f0 = a.tile_frag[t];
        const uint64_t w0 = ts >> 4;
        const int nw = (int)((((ts & 15u) + (uint64_t)tl + 64 + 32) >> 4) + 2);
        __syncthreads();
        for (int i = threadIdx.x; i < ZI_WORDS; i += ZI_THREADS) s_words[i] = i < nw ? a.words[w0 + i] : 0u;
        const int nf = load_fragments(a, ts, tl, f0, s_fs);
        const uint64_t rb = ts & 15u;  // tile start inside s_words
        Key2 key[ZI_WPT];
        uint64_t slot[ZI_WPT];
        uint32_t req[ZI_WPT];
        int fj[ZI_WPT];
        uint32_t pending = 0, absent = 0;
        // The request of every window's fragment, all loads issued before any
        // is used (windows past the tile read the tile's first fragment).
#pragma unroll
        for (int i = 0; i < ZI_WPT; i++) {
            const int r = i * ZI_THREADS + (int)threadIdx.x;
            fj[i] = r < tl ? frag_of(s_fs, nf, r) : 0;
            req[i] = a.frag_req[f0 + (uint64_t)fj[i]];
        }
#pragma unroll
        for (int i = 0; i < ZI_WPT; i++) {
            const int r = i * ZI_THREADS + (int)threadIdx.x;
            key[i] = Key2{0, 0};
            slot[i] = 0;
            if (r >= tl) req[i] = NO_REQ;
            if (r < tl && r + a.k <= s_fs[fj[i] + 1] && req[i] != NO_REQ) {
                my_lookups++;
                if constexpr (!WIDE) {
                    key[i].lo = canon64(s_words, rb + (uint64_t)r, a.k);
                    slot[i] = slot_of(key[i].lo, 0) & a.mask;
                } else {
                    key[i] = canon128(s_words, rb + (uint64_t)r, a.k);
                    slot[i] = slot_of(key[i].lo, key[i].hi) & a.mask;
                }
                pending |= 1u << i;
            }
        }
        // One probe round: the windows' table reads (k > 32: both words of
        // a slot in one 16-byte read), then their checks.
        // ALL (every round for k > 32): every window reads its slot, pending
        // or not -- slot 0 or its last probe, in bounds, the value masked
        // out -- and the checks are bit masks without a branch per window:
        // with a branch per read the compiler moves part of the 16-byte check
        // into it and waits for each read before issuing the next. For k <=
        // 32 only the pending windows read (the reads stay in flight together
        // there, and later rounds have few pending windows; reading all of
        // them took the k = 31 lookup from 3.3e10 to 2.0e10 per second).
        auto probe_round = [&](auto all_tag) {
            constexpr bool ALL = decltype(all_tag)::value;
            using Slot = std::conditional_t<WIDE, uint4, unsigned long long>;
            Slot e[ZI_WPT];
#pragma unroll
            for (int i = 0; i < ZI_WPT; i++) {
                if (!ALL && !(pending & (1u << i))) continue;
                if constexpr (WIDE) e[i] = *reinterpret_cast<const uint4 *>(a.tw + slot[i]);
                else e[i] = a.tlo[slot[i]];
            }
            auto check = [&](int i, bool &empty, bool &found) {
                if constexpr (WIDE) {
                    const unsigned long long hi = ((unsigned long long)e[i].y << 32) | e[i].x;
                    const unsigned long long lo = ((unsigned long long)e[i].w << 32) | e[i].z;
                    empty = hi == EMPTY_KEY;
                    found = hi == key[i].hi && lo == key[i].lo;
                } else {
                    empty = e[i] == EMPTY_KEY;
                    found = e[i] == key[i].lo;
                }
            };
            if constexpr (ALL) {
                uint32_t em = 0, fm = 0;
#pragma unroll
                for (int i = 0; i < ZI_WPT; i++) {
                    bool empty, found;
                    check(i, empty, found);
                    em |= (empty ? 1u : 0u) << i;
                    fm |= (found ? 1u : 0u) << i;
                }
                absent |= em & pending;
                pending &= ~(em | fm);
#pragma unroll
                for (int i = 0; i < ZI_WPT; i++)
                    slot[i] = (pending >> i) & 1u ? (slot[i] + 1) & a.mask : slot[i];
            } else {
#pragma unroll
                for (int i = 0; i < ZI_WPT; i++) {
                    if (!(pending & (1u << i))) continue;
                    bool empty, found;
                    check(i, empty, found);
                    if (empty) {
                        absent |= 1u << i;
                        pending &= ~(1u << i);
                    } else if (found) {
                        pending &= ~(1u << i);
                    } else {
                        slot[i] = (slot[i] + 1) & a.mask;
                    }
                }
            }
        };
        // the first round is written out before the loop, so that its reads
        // (nearly every lookup ends in it) are all in flight together
        if (pending) probe_round(std::integral_constant<bool, WIDE>{});
        while (pending) probe_round(std::integral_constant<bool, WIDE>{});
        // Append the absent k-mers to their request's segment. When they all
        // belong to one request (every tile but those that straddle request
        // boundaries), the block takes one reservation for the tile and each
        // wave writes its k-mers of window i as one contiguous run (order:
        // window i, then wave); otherwise each wave reserves per window i and
        // request. Same-address returning atomics were the lookup's bound.
        const int wid = (int)threadIdx.x >> 6;
        uint32_t rmin = NO_REQ, rmax = 0;
#pragma unroll
        for (int i = 0; i < ZI_WPT; i++)
            if ((absent >> i) & 1u) {
                rmin = min(rmin, req[i]);
                rmax = max(rmax, req[i]);
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            rmin = min(rmin, (uint32_t)__shfl_xor((int)rmin, o, 64));
            rmax = max(rmax, (uint32_t)__shfl_xor((int)rmax, o, 64));
        }
        uint64_t abm[ZI_WPT];
#pragma unroll
        for (int i = 0; i < ZI_WPT; i++) {
            abm[i] = __ballot((absent >> i) & 1u);
            if (lane == 0) s_cnt[i * (ZI_THREADS / 64) + wid] = (uint32_t)__popcll(abm[i]);
        }
        if (lane == 0) {
            s_rmin[wid] = rmin;
            s_rmax[wid] = rmax;
        }
        __syncthreads();
        uint32_t bmin = NO_REQ, bmax = 0;
#pragma unroll
        for (int w = 0; w < ZI_THREADS / 64; w++) {
            bmin = min(bmin, s_rmin[w]);
            bmax = max(bmax, s_rmax[w]);
        }
        if (bmin == NO_REQ) continue;  // nothing absent in this tile (uniform over the block)
        const unsigned long long lt = (1ull << lane) - 1ull;
        if (bmin == bmax) {
            if (threadIdx.x == 0) {
                uint32_t tot = 0;
                for (int c = 0; c < ZI_WPT * (ZI_THREADS / 64); c++) tot += s_cnt[c];
                s_base = atomicAdd(&a.seg_cursor[bmin], (unsigned long long)tot);
            }
            __syncthreads();
            unsigned long long run = s_base;  // first slot of window i's k-mers
#pragma unroll
            for (int i = 0; i < ZI_WPT; i++) {
                unsigned long long before = 0, tot_i = 0;
#pragma unroll
                for (int w = 0; w < ZI_THREADS / 64; w++) {
                    const uint32_t c = s_cnt[i * (ZI_THREADS / 64) + w];
                    before += w < wid ? c : 0u;
                    tot_i += c;
                }
                if ((absent >> i) & 1u) {
                    const unsigned long long pos = run + before + (unsigned long long)__popcll(abm[i] & lt);
                    a.out_lo[pos] = key[i].lo;
                    if constexpr (WIDE) a.out_hi[pos] = key[i].hi;
                }
                run += tot_i;
            }
        } else {
#pragma unroll
            for (int i = 0; i < ZI_WPT; i++) {
                const bool ab = (absent >> i) & 1u;
                uint64_t todo = abm[i];
                while (todo) {
                    const int leader = __ffsll((long long)todo) - 1;
                    const uint32_t rl = __shfl(req[i], leader, 64);
                    const uint64_t same = __ballot(ab && req[i] == rl);
                    unsigned long long base = 0;
                    if (lane == leader) base = atomicAdd(&a.seg_cursor[rl], (unsigned long long)__popcll(same));
                    base = __shfl(base, leader, 64);
                    if (ab && req[i] == rl) {
                        const unsigned long long pos = base + __popcll(same & lt);
                        a.out_lo[pos] = key[i].lo;
                        if constexpr (WIDE) a.out_hi[pos] = key[i].hi;
                    }
                    todo &= ~same;
                }
            }
        }
    }
    __shared__ unsigned long long s_red[ZI_THREADS / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) my_lookups += __shfl_xor(my_lookups, o, 64);
    if (lane == 0) s_red[threadIdx.x >> 6] = my_lookups;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < ZI_THREADS / 64; i++) tot += s_red[i];
        if (tot) atomicAdd(&a.ctr[0], tot);
    }
}

__global__ void k_zi_fill(unsigned long long *p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_zi_copy_u64(const unsigned long long *src, unsigned long long *dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void k_zi_iota(uint32_t *p, uint64_t n, uint64_t base) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)(base + i);
}
__global__ void k_zi_gather_range(const unsigned long long *src, const uint32_t *idx, unsigned long long *dst,
                                  uint64_t b, uint64_t e) {
    for (uint64_t i = b + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

// The k > 32 counterpart of k_finish_runs: (hi, lo) pairs already ordered
// by hi >> shift (shift >= 32, so a run's bounds read the same while its
// elements are rewritten); each run of equal prefixes is ordered by the full
// 128-bit key in LDS, every element moving to its run's start plus its
// stable rank. A run longer than FINISH_MAX sets *flag (the caller then sorts
// the segment in full).
__global__ __launch_bounds__(256) void k_finish_runs_wide(unsigned long long *hi, unsigned long long *lo, uint64_t n,
                                                          int shift, unsigned long long *flag) {
    __shared__ unsigned long long sh[FIN_CHUNK + FINISH_MAX + 1], sl[FIN_CHUNK + FINISH_MAX + 1];
    __shared__ uint64_t s_a, s_b;
    for (uint64_t c0 = (uint64_t)blockIdx.x * FIN_CHUNK; c0 < n; c0 += (uint64_t)gridDim.x * FIN_CHUNK) {
        const uint64_t c1 = c0 + FIN_CHUNK < n ? c0 + FIN_CHUNK : n;
        __syncthreads();  // the previous chunk is written back
        if (threadIdx.x < 64) {
            // [a, b): from the first run starting in [c0, c1) to the end of
            // the run holding c1 - 1 (the first wave tests 64 positions at
            // once on each side)
            const int lane = threadIdx.x;
            uint64_t a = c0;
            if (c0 > 0) {
                const unsigned long long pp = hi[c0 - 1] >> shift;
                const uint64_t i = c0 + (uint64_t)lane;
                const bool other = i >= c1 || (hi[i] >> shift) != pp;
                const unsigned long long m = __ballot(other);
                a = m ? c0 + (uint64_t)__ffsll((long long)m) - 1 : c0 + FINISH_MAX + 1;
            }
            uint64_t b = a;
            if (a < c1) {
                const unsigned long long pl = hi[c1 - 1] >> shift;
                const uint64_t i = c1 + (uint64_t)lane;
                const bool other = i >= n || (hi[i] >> shift) != pl;
                const unsigned long long m = __ballot(other);
                b = m ? c1 + (uint64_t)__ffsll((long long)m) - 1 : c1 + FINISH_MAX;
                if (!m && lane == 0) atomicOr(flag, 1ull);
            }
            if (lane == 0) {
                if (a > c0 + FINISH_MAX) atomicOr(flag, 1ull);
                s_a = a < c1 ? a : c1;
                s_b = a < c1 ? b : c1;
            }
        }
        __syncthreads();
        const uint64_t a = s_a, b = s_b;
        const int len = (int)(b - a);
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            sh[i] = hi[a + i];
            sl[i] = lo[a + i];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            const unsigned long long kh = sh[i], kl = sl[i], pfx = kh >> shift;
            int rs = i;
            while (rs > 0 && (sh[rs - 1] >> shift) == pfx && i - rs < FINISH_MAX) rs--;
            int rank = 0, e = rs;
            while (e < len && (sh[e] >> shift) == pfx && e - rs <= FINISH_MAX) {
                const unsigned long long eh = sh[e], el = sl[e];
                rank += eh < kh || (eh == kh && (el < kl || (el == kl && e < i)));
                e++;
            }
            if (e - rs > FINISH_MAX) {
                atomicOr(flag, 1ull);
                continue;
            }
            hi[a + rs + rank] = kh;
            lo[a + rs + rank] = kl;
        }
    }
}

// Unique within each request's sorted segment: flag the first of every run
// of equal keys among the segment's filled entries [base, end).
template <bool WIDE>
__global__ void k_zi_flags(const unsigned long long *lo, const unsigned long long *hi,
                           const unsigned long long *seg_base, const unsigned long long *seg_end, uint32_t n_seg,
                           uint32_t *flags, uint64_t cap) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) flags[cap] = 0;
    for (uint32_t sg = blockIdx.y; sg < n_seg; sg += gridDim.y) {
        const uint64_t b = seg_base[sg], e = seg_base[sg + 1], f = seg_end[sg];
        for (uint64_t i = b + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e;
             i += (uint64_t)gridDim.x * blockDim.x) {
            bool fl = false;
            if (i < f) {
                fl = i == b || lo[i] != lo[i - 1];
                if constexpr (WIDE) fl = fl || hi[i] != hi[i - 1];
            }
            flags[i] = fl ? 1u : 0u;
        }
    }
}

template <bool WIDE>
__global__ void k_zi_scatter(const unsigned long long *lo, const unsigned long long *hi, const uint32_t *flags,
                             const uint32_t *pos, uint64_t cap, unsigned long long *olo, unsigned long long *ohi) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        if (!flags[i]) continue;
        olo[pos[i]] = lo[i];
        if constexpr (WIDE) ohi[pos[i]] = hi[i];
    }
}

__global__ void k_zi_seg_counts(const uint32_t *pos, const unsigned long long *seg_base, uint32_t n_seg,
                                unsigned long long *out_start, unsigned long long *out_count) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n_seg; r += gridDim.x * blockDim.x) {
        const uint64_t a = pos[seg_base[r]], b = pos[seg_base[r + 1]];
        out_start[r] = a;
        out_count[r] = b - a;
    }
}

// ---------------------------------------------------------------------------
static unsigned zgrid(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}
static void zfill(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_zi_fill, dim3(zgrid(n)), dim3(256), 0, s, p, n, v);
}

struct HostTiles {
    std::vector<uint64_t> start;
    std::vector<uint32_t> frag, len;
};

static void make_tiles(const std::vector<uint64_t> &off, HostTiles &t) {
    t.start.clear();
    t.frag.clear();
    t.len.clear();
    const uint64_t nb = off.back();
    uint64_t f = 0;
    const uint64_t nfr = off.size() - 1;
    for (uint64_t s = 0; s < nb; s += ZI_TILE) {
        while (f + 1 < nfr && off[f + 1] <= s) f++;
        t.start.push_back(s);
        t.frag.push_back((uint32_t)f);
        t.len.push_back((uint32_t)std::min<uint64_t>(ZI_TILE, nb - s));
    }
}

void zi_build_index(ZiIndex &idx, HostReads &&reads, uint32_t k, uint32_t m, hipStream_t s) {
    reads.finish();
    idx.k = k;
    idx.m = m;
    idx.wide = k > 32;
    uint64_t windows = 0;
    for (size_t i = 0; i + 1 < reads.off.size(); i++) {
        const uint64_t n = reads.off[i + 1] - reads.off[i];
        if (n >= k) windows += n - k + 1;
    }
    require(windows > 0, MCZI_E_INVALID, "index FASTA did not contain any sequence of length >= k");
    require(reads.n_frag() < (1ull << 32), MCZI_E_INVALID, "too many index fragments");
    uint64_t slots = 1024;
    while (slots < windows * 2) slots <<= 1;
    idx.mask = slots - 1;
    idx.n_kmers = windows;
    idx.slots.reset(slots * (idx.wide ? 16 : 8));
    zfill(idx.slots.as<unsigned long long>(), slots * (idx.wide ? 2 : 1), EMPTY_KEY, s);
    DevBuf words(reads.words.size() * 4), off(reads.off.size() * 8);
    MCZI_HIP(hipMemcpyAsync(words.as<void>(), reads.words.data(), reads.words.size() * 4, hipMemcpyHostToDevice, s));
    MCZI_HIP(hipMemcpyAsync(off.as<void>(), reads.off.data(), reads.off.size() * 8, hipMemcpyHostToDevice, s));
    HostTiles t;
    make_tiles(reads.off, t);
    DevBuf ts(t.start.size() * 8 + 8), tf(t.frag.size() * 4 + 4), tl(t.len.size() * 4 + 4);
    MCZI_HIP(hipMemcpyAsync(ts.as<void>(), t.start.data(), t.start.size() * 8, hipMemcpyHostToDevice, s));
    MCZI_HIP(hipMemcpyAsync(tf.as<void>(), t.frag.data(), t.frag.size() * 4, hipMemcpyHostToDevice, s));
    MCZI_HIP(hipMemcpyAsync(tl.as<void>(), t.len.data(), t.len.size() * 4, hipMemcpyHostToDevice, s));
    ZiTileArgs a{};
    a.words = words.as<uint32_t>();
    a.off = off.as<uint64_t>();
    a.tile_start = ts.as<uint64_t>();
    a.tile_frag = tf.as<uint32_t>();
    a.tile_len = tl.as<uint32_t>();
    a.n_tiles = t.start.size();
    a.n_frag = reads.n_frag();
    a.k = (int)k;
    a.tlo = idx.wide ? nullptr : idx.slots.as<unsigned long long>();
    a.tw = idx.wide ? idx.slots.as<ulonglong2>() : nullptr;
    a.mask = idx.mask;
    const unsigned g = (unsigned)std::min<uint64_t>(a.n_tiles, (uint64_t)Device::get().cu_count() * 8);
    if (idx.wide) hipLaunchKernelGGL(k_zi_build<true>, dim3(g), dim3(ZI_THREADS), 0, s, a);
    else hipLaunchKernelGGL(k_zi_build<false>, dim3(g), dim3(ZI_THREADS), 0, s, a);
    MCZI_HIP(hipStreamSynchronize(s));
    MCZI_HIP(hipGetLastError());
}

void zi_build_index_from_keys(ZiIndex &idx, const unsigned long long *lo, const unsigned long long *hi, uint64_t n,
                              uint32_t k, uint32_t m, hipStream_t s) {
    require(k <= 32 || hi != nullptr || n == 0, MCZI_E_INVALID, "k > 32 needs the high k-mer words");
    idx.k = k;
    idx.m = m;
    idx.wide = k > 32;
    uint64_t slots = 1024;
    while (slots < n * 2) slots <<= 1;
    idx.mask = slots - 1;
    idx.n_kmers = n;
    idx.slots.reset(slots * (idx.wide ? 16 : 8));
    zfill(idx.slots.as<unsigned long long>(), slots * (idx.wide ? 2 : 1), EMPTY_KEY, s);
    if (n) {
        const unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1u << 20));
        if (idx.wide)
            hipLaunchKernelGGL(k_zi_insert<true>, dim3(g), dim3(256), 0, s, lo, hi, n, (unsigned long long *)nullptr,
                               idx.slots.as<ulonglong2>(), idx.mask);
        else
            hipLaunchKernelGGL(k_zi_insert<false>, dim3(g), dim3(256), 0, s, lo, hi, n,
                               idx.slots.as<unsigned long long>(), (ulonglong2 *)nullptr, idx.mask);
    }
    MCZI_HIP(hipStreamSynchronize(s));
    MCZI_HIP(hipGetLastError());
}

struct ZiWork {
    // host copies kept alive until the batch's asynchronous copies are done
    std::vector<uint64_t> h_off;
    std::vector<uint32_t> frag_req;
    std::vector<uint64_t> t_start;
    std::vector<uint32_t> t_frag, t_len;
    DevBuf words, off, req, ts, tf, tl, seg_base, seg_cursor, seg_count, seg_start, olo, ohi, slo, shi, ulo, uhi, idx,
        idx2, flags, pos, tmp, ctr;
    uint64_t cap = 0;
    std::vector<uint64_t> h_base, h_count, h_start, h_absent;  // per request
    uint64_t lookups = 0;
    hipEvent_t done = nullptr;
    hipEvent_t l0 = nullptr, l1 = nullptr;  // around the lookup kernel
    bool timed = false;
    double lookup_ms = 0;  // the lookup kernel's duration, known after zi_finish
    ~ZiWork() {
        if (done) (void)hipEventDestroy(done);
        if (l0) (void)hipEventDestroy(l0);
        if (l1) (void)hipEventDestroy(l1);
    }
};

std::unique_ptr<ZiWork> zi_make_work() {
    auto w = std::make_unique<ZiWork>();
    MCZI_HIP(hipEventCreateWithFlags(&w->done, hipEventDisableTiming));
    MCZI_HIP(hipEventCreate(&w->l0));
    MCZI_HIP(hipEventCreate(&w->l1));
    return w;
}
void zi_free_work(ZiWork *w) { delete w; }

uint64_t zi_batch_bytes(uint64_t bases, uint64_t frags, bool wide) {
    const uint64_t tiles = bases / ZI_TILE + 1;
    const uint64_t per_key = wide ? 16 * 3 + 8 : 8 * 3;
    return bases / 4 + 256 + frags * 12 + tiles * 16 + bases * per_key + bases * 8 + (64u << 10);
}

void zi_enqueue(ZiWork &w, const ZiIndex &idx, ZiBatchInput &&in, uint32_t n_req, hipStream_t s) {
    const bool wide = idx.wide;
    w.h_off = std::move(in.off);
    w.frag_req = std::move(in.frag_req);
    const std::vector<uint64_t> &qoff = w.h_off;
    const std::vector<uint32_t> &frag_req = w.frag_req;
    const uint64_t nfr = qoff.size() - 1;
    require(frag_req.size() == nfr, MCZI_E_INVALID, "fragment / request tags differ in length");
    require(nfr < (1ull << 32), MCZI_E_INVALID, "too many query fragments in one batch");
    // per-request segment bases: capacity = the request's window count
    std::vector<uint64_t> seg_cap(n_req, 0);
    for (uint64_t f = 0; f < nfr; f++) {
        const uint32_t r = frag_req[f];
        const uint64_t n = qoff[f + 1] - qoff[f];
        if (r != NO_REQ && n >= idx.k) seg_cap[r] += n - idx.k + 1;
    }
    w.h_base.assign(n_req + 1, 0);
    for (uint32_t r = 0; r < n_req; r++) w.h_base[r + 1] = w.h_base[r] + seg_cap[r];
    const uint64_t cap = std::max<uint64_t>(w.h_base[n_req], 1);
    require(cap < (1ull << 32) - 1, MCZI_E_INVALID, "too many query k-mers in one batch (limit 2^32 - 2)");
    w.cap = cap;
    // words of the batch + slack read past the last base (extraction and
    // tile staging); gap and slack contents never reach a key
    require(in.n_words * 16 >= qoff.back(), MCZI_E_INVALID, "query batch words shorter than its fragments");
    w.words.ensure((in.n_words + 48) * 4);
    w.off.ensure(qoff.size() * 8);
    w.req.ensure(std::max<uint64_t>(nfr, 1) * 4);
    for (const ZiWordSpan &sp : in.spans) {
        require(sp.dst_word + sp.n_words <= in.n_words, MCZI_E_INVALID, "query word span outside the batch");
        if (sp.n_words)
            MCZI_HIP(hipMemcpyAsync(w.words.as<uint32_t>() + sp.dst_word, sp.src, sp.n_words * 4,
                                    hipMemcpyHostToDevice, s));
    }
    MCZI_HIP(hipMemcpyAsync(w.off.as<void>(), qoff.data(), qoff.size() * 8, hipMemcpyHostToDevice, s));
    if (nfr)
        MCZI_HIP(hipMemcpyAsync(w.req.as<void>(), frag_req.data(), nfr * 4, hipMemcpyHostToDevice, s));
    HostTiles t;
    make_tiles(qoff, t);
    w.t_start = t.start;
    w.t_frag = t.frag;
    w.t_len = t.len;
    w.ts.ensure(t.start.size() * 8 + 8);
    w.tf.ensure(t.frag.size() * 4 + 4);
    w.tl.ensure(t.len.size() * 4 + 4);
    if (!t.start.empty()) {
        MCZI_HIP(hipMemcpyAsync(w.ts.as<void>(), w.t_start.data(), t.start.size() * 8, hipMemcpyHostToDevice, s));
        MCZI_HIP(hipMemcpyAsync(w.tf.as<void>(), w.t_frag.data(), t.frag.size() * 4, hipMemcpyHostToDevice, s));
        MCZI_HIP(hipMemcpyAsync(w.tl.as<void>(), w.t_len.data(), t.len.size() * 4, hipMemcpyHostToDevice, s));
    }
    w.seg_base.ensure((n_req + 1) * 8);
    w.seg_cursor.ensure((n_req + 1) * 8);
    w.seg_count.ensure((n_req + 1) * 8);
    w.seg_start.ensure((n_req + 1) * 8);
    MCZI_HIP(hipMemcpyAsync(w.seg_cursor.as<void>(), w.h_base.data(), (n_req + 1) * 8, hipMemcpyHostToDevice, s));
    w.olo.ensure(cap * 8);
    if (wide) w.ohi.ensure(cap * 8);
    w.ctr.ensure(64);
    zfill(w.ctr.as<unsigned long long>(), 8, 0, s);

    ZiTileArgs a{};
    a.words = w.words.as<uint32_t>();
    a.off = w.off.as<uint64_t>();
    a.frag_req = w.req.as<uint32_t>();
    a.tile_start = w.ts.as<uint64_t>();
    a.tile_frag = w.tf.as<uint32_t>();
    a.tile_len = w.tl.as<uint32_t>();
    a.n_tiles = t.start.size();
    a.n_frag = nfr;
    a.k = (int)idx.k;
    a.tlo = wide ? nullptr : idx.slots.as<unsigned long long>();
    a.tw = wide ? idx.slots.as<ulonglong2>() : nullptr;
    a.mask = idx.mask;
    a.seg_cursor = w.seg_cursor.as<unsigned long long>();
    a.out_lo = w.olo.as<unsigned long long>();
    a.out_hi = wide ? w.ohi.as<unsigned long long>() : nullptr;
    a.ctr = w.ctr.as<unsigned long long>();
    w.timed = a.n_tiles != 0;
    if (a.n_tiles) {
        const unsigned g = (unsigned)std::min<uint64_t>(a.n_tiles, (uint64_t)Device::get().cu_count() * 8);
        MCZI_HIP(hipEventRecord(w.l0, s));
        if (wide) hipLaunchKernelGGL(k_zi_lookup<true>, dim3(g), dim3(ZI_THREADS), 0, s, a);
        else hipLaunchKernelGGL(k_zi_lookup<false>, dim3(g), dim3(ZI_THREADS), 0, s, a);
        MCZI_HIP(hipEventRecord(w.l1, s));
    }
    MCZI_HIP(hipEventRecord(w.done, s));
}

// Sort the filled part of every request's segment (olo/ohi at h_base[r],
// n[r] entries) into a tight layout (slo/uhi at tb[r]). k <= 32: radix
// passes over the top 24 key bits, then the finishing pass orders each run of
// equal prefixes in LDS (a run too long to finish makes the segment take the
// full sort); small segments take the full sort. k > 32: LSD over (lo, hi)
// carrying the source positions.
static void zi_sort_segments(ZiWork &w, const ZiIndex &idx, const std::vector<uint64_t> &n,
                             const std::vector<uint64_t> &tb, hipStream_t s) {
    const uint32_t n_req = (uint32_t)n.size();
    const uint64_t N = tb[n_req];
    uint64_t max_n = 1;
    for (uint64_t x : n) max_n = std::max(max_n, x);
    unsigned long long *olo = w.olo.as<unsigned long long>(), *slo = w.slo.as<unsigned long long>();
    if (!idx.wide) {
        const unsigned kbits = std::min<unsigned>(64, 2 * idx.k);
        constexpr unsigned PREFIX_BITS = 24;
        constexpr uint64_t SPLIT_MIN = 1u << 16;
        size_t t_split = 0, t_full = 0;
        MCZI_HIP(rocprim::radix_sort_keys(nullptr, t_split, olo, slo, (size_t)max_n, kbits - std::min(kbits, PREFIX_BITS),
                                          kbits, s));
        MCZI_HIP(rocprim::radix_sort_keys(nullptr, t_full, olo, slo, (size_t)max_n, 0, kbits, s));
        w.tmp.ensure(std::max(t_split, t_full) + 16);
        unsigned long long *flag = w.ctr.as<unsigned long long>() + 7;
        zfill(flag, 1, 0, s);
        std::vector<uint32_t> split;
        for (uint32_t r = 0; r < n_req; r++) {
            if (n[r] == 0) continue;
            const bool sp = kbits >= 56 && n[r] >= SPLIT_MIN && (n[r] >> PREFIX_BITS) < FINISH_MAX;
            size_t t = sp ? t_split : t_full;
            MCZI_HIP(rocprim::radix_sort_keys(w.tmp.as<void>(), t, olo + w.h_base[r], slo + tb[r], (size_t)n[r],
                                              sp ? kbits - PREFIX_BITS : 0u, kbits, s));
            if (sp) {
                const unsigned g = (unsigned)std::min<uint64_t>((n[r] + FIN_CHUNK - 1) / FIN_CHUNK, 1u << 20);
                hipLaunchKernelGGL(k_finish_runs<false>, dim3(g), dim3(256), 0, s, slo + tb[r], (uint32_t *)nullptr,
                                   n[r], (int)(kbits - PREFIX_BITS), flag);
                split.push_back(r);
            }
        }
        if (!split.empty()) {
            unsigned long long h = 0;
            MCZI_HIP(hipMemcpyAsync(&h, flag, 8, hipMemcpyDeviceToHost, s));
            MCZI_HIP(hipStreamSynchronize(s));
            for (uint32_t r : h ? split : std::vector<uint32_t>()) {
                size_t t = t_full;
                MCZI_HIP(rocprim::radix_sort_keys(w.tmp.as<void>(), t, olo + w.h_base[r], slo + tb[r],
                                                  (size_t)n[r], 0, kbits, s));
            }
        }
        return;
    }
    const unsigned hbits = 2 * idx.k - 64;
    unsigned long long *ohi = w.ohi.as<unsigned long long>(), *shi = w.shi.as<unsigned long long>(),
                       *uhi = w.uhi.as<unsigned long long>();
    uint32_t *i1 = w.idx.as<uint32_t>(), *i2 = w.idx2.as<uint32_t>();
    // k >= 60 (the high word holds at least 56 key bits): large segments
    // sort (hi, position) by the top 24 key bits only, gather lo, and order
    // the short runs of equal prefixes by the full key in LDS; otherwise (and
    // when a run is too long) LSD over lo, then hi, carrying positions.
    constexpr unsigned PREFIX_BITS = 24;
    constexpr uint64_t SPLIT_MIN = 1u << 16;
    const bool can_split = hbits >= PREFIX_BITS + 32;
    size_t t1 = 0, t2 = 0, tp = 0;
    MCZI_HIP(rocprim::radix_sort_pairs(nullptr, t1, olo, slo, i1, i2, (size_t)max_n, 0, 64, s));
    MCZI_HIP(rocprim::radix_sort_pairs(nullptr, t2, shi, uhi, i2, i1, (size_t)max_n, 0, hbits, s));
    if (can_split)
        MCZI_HIP(rocprim::radix_sort_pairs(nullptr, tp, ohi, uhi, i1, i2, (size_t)max_n, hbits - PREFIX_BITS, hbits,
                                           s));
    w.tmp.ensure(std::max(std::max(t1, t2), tp) + 16);
    unsigned long long *flag = w.ctr.as<unsigned long long>() + 7;
    auto full_sort = [&](uint32_t r) {
        const uint64_t b0 = tb[r], e0 = tb[r] + n[r];
        size_t ts1 = t1, ts2 = t2;
        hipLaunchKernelGGL(k_zi_iota, dim3(zgrid(n[r])), dim3(256), 0, s, i1 + b0, n[r], w.h_base[r]);
        MCZI_HIP(rocprim::radix_sort_pairs(w.tmp.as<void>(), ts1, olo + w.h_base[r], slo + b0, i1 + b0, i2 + b0,
                                           (size_t)n[r], 0, 64, s));
        hipLaunchKernelGGL(k_zi_gather_range, dim3(zgrid(n[r])), dim3(256), 0, s, ohi, i2, shi, b0, e0);
        MCZI_HIP(rocprim::radix_sort_pairs(w.tmp.as<void>(), ts2, shi + b0, uhi + b0, i2 + b0, i1 + b0, (size_t)n[r],
                                           0, hbits, s));
        hipLaunchKernelGGL(k_zi_gather_range, dim3(zgrid(n[r])), dim3(256), 0, s, olo, i1, slo, b0, e0);
    };
    std::vector<uint32_t> split;
    if (can_split) zfill(flag, 1, 0, s);
    for (uint32_t r = 0; r < n_req; r++) {
        if (n[r] == 0) continue;
        if (!(can_split && n[r] >= SPLIT_MIN && (n[r] >> PREFIX_BITS) < FINISH_MAX)) {
            full_sort(r);
            continue;
        }
        const uint64_t b0 = tb[r], e0 = tb[r] + n[r];
        size_t ts = tp;
        hipLaunchKernelGGL(k_zi_iota, dim3(zgrid(n[r])), dim3(256), 0, s, i1 + b0, n[r], w.h_base[r]);
        MCZI_HIP(rocprim::radix_sort_pairs(w.tmp.as<void>(), ts, ohi + w.h_base[r], uhi + b0, i1 + b0, i2 + b0,
                                           (size_t)n[r], hbits - PREFIX_BITS, hbits, s));
        hipLaunchKernelGGL(k_zi_gather_range, dim3(zgrid(n[r])), dim3(256), 0, s, olo, i2, slo, b0, e0);
        const unsigned g = (unsigned)std::min<uint64_t>((n[r] + FIN_CHUNK - 1) / FIN_CHUNK, 1u << 20);
        hipLaunchKernelGGL(k_finish_runs_wide, dim3(g), dim3(256), 0, s, uhi + b0, slo + b0, n[r],
                           (int)(hbits - PREFIX_BITS), flag);
        split.push_back(r);
    }
    if (!split.empty()) {
        unsigned long long h = 0;
        MCZI_HIP(hipMemcpyAsync(&h, flag, 8, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipStreamSynchronize(s));
        if (h)
            for (uint32_t r : split) full_sort(r);
    }
    (void)N;
}

// Wait for the lookup, then (with the filled size of every segment known)
// sort and deduplicate each request's absent k-mers.
double zi_finish(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s) {
    const auto tw = std::chrono::steady_clock::now();
    MCZI_HIP(hipEventSynchronize(w.done));
    const double wait_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count();
    MCZI_HIP(hipGetLastError());
    w.lookup_ms = 0;
    if (w.timed) {
        float f = 0;
        MCZI_HIP(hipEventElapsedTime(&f, w.l0, w.l1));
        w.lookup_ms = f;
    }
    const bool wide = idx.wide;
    std::vector<uint64_t> h_cursor(n_req + 1, 0);
    unsigned long long lk = 0;
    MCZI_HIP(hipMemcpyAsync(h_cursor.data(), w.seg_cursor.as<void>(), n_req * 8, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipMemcpyAsync(&lk, w.ctr.as<void>(), 8, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipStreamSynchronize(s));
    w.lookups = lk;
    std::vector<uint64_t> n(n_req, 0), tb(n_req + 1, 0);
    for (uint32_t r = 0; r < n_req; r++) {
        n[r] = h_cursor[r] - w.h_base[r];
        require(n[r] <= w.h_base[r + 1] - w.h_base[r], MCZI_E_DEVICE, "absent k-mers exceed their request's windows");
        tb[r + 1] = tb[r] + n[r];
    }
    const uint64_t N = tb[n_req];
    w.h_count.assign(n_req + 1, 0);
    w.h_start.assign(n_req + 1, 0);
    w.h_absent = n;
    if (N) {
        w.slo.ensure(N * 8);
        w.ulo.ensure(N * 8);
        w.flags.ensure((N + 1) * 4);
        w.pos.ensure((N + 1) * 4);
        if (wide) {
            w.shi.ensure(N * 8);
            w.uhi.ensure(N * 8);
            w.idx.ensure(N * 4);
            w.idx2.ensure(N * 4);
        }
        zi_sort_segments(w, idx, n, tb, s);
        // unique per request: flags -> exclusive scan -> scatter
        MCZI_HIP(hipMemcpyAsync(w.seg_base.as<void>(), tb.data(), (n_req + 1) * 8, hipMemcpyHostToDevice, s));
        const unsigned long long *sb = w.seg_base.as<unsigned long long>();
        const unsigned long long *khi = wide ? w.uhi.as<unsigned long long>() : nullptr;
        dim3 gg(zgrid(N), std::min<uint32_t>(std::max<uint32_t>(n_req, 1), 64));
        if (wide)
            hipLaunchKernelGGL(k_zi_flags<true>, gg, dim3(256), 0, s, w.slo.as<unsigned long long>(), khi, sb, sb + 1,
                               n_req, w.flags.as<uint32_t>(), N);
        else
            hipLaunchKernelGGL(k_zi_flags<false>, gg, dim3(256), 0, s, w.slo.as<unsigned long long>(), khi, sb, sb + 1,
                               n_req, w.flags.as<uint32_t>(), N);
        size_t t3 = 0;
        MCZI_HIP(rocprim::exclusive_scan(nullptr, t3, w.flags.as<uint32_t>(), w.pos.as<uint32_t>(), 0u, (size_t)N + 1,
                                         rocprim::plus<uint32_t>(), s));
        w.tmp.ensure(t3 + 16);
        MCZI_HIP(rocprim::exclusive_scan(w.tmp.as<void>(), t3, w.flags.as<uint32_t>(), w.pos.as<uint32_t>(), 0u,
                                         (size_t)N + 1, rocprim::plus<uint32_t>(), s));
        if (wide)
            hipLaunchKernelGGL(k_zi_scatter<true>, dim3(zgrid(N)), dim3(256), 0, s, w.slo.as<unsigned long long>(),
                               khi, w.flags.as<uint32_t>(), w.pos.as<uint32_t>(), N, w.ulo.as<unsigned long long>(),
                               w.ohi.as<unsigned long long>());
        else
            hipLaunchKernelGGL(k_zi_scatter<false>, dim3(zgrid(N)), dim3(256), 0, s, w.slo.as<unsigned long long>(),
                               khi, w.flags.as<uint32_t>(), w.pos.as<uint32_t>(), N, w.ulo.as<unsigned long long>(),
                               (unsigned long long *)nullptr);
        hipLaunchKernelGGL(k_zi_seg_counts, dim3(1), dim3(256), 0, s, w.pos.as<uint32_t>(), sb, n_req,
                           w.seg_start.as<unsigned long long>(), w.seg_count.as<unsigned long long>());
        MCZI_HIP(hipMemcpyAsync(w.h_count.data(), w.seg_count.as<void>(), n_req * 8, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipMemcpyAsync(w.h_start.data(), w.seg_start.as<void>(), n_req * 8, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipStreamSynchronize(s));
    }
    return wait_ms;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void zi_copy_out(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s, ZiQueryResult &out) {
    const bool wide = idx.wide;
    out.absent.clear();
    out.absent.resize(n_req);
    out.windows.assign(n_req, 0);
    out.absent_windows.assign(n_req, 0);
    for (uint32_t r = 0; r < n_req; r++) {
        out.windows[r] = w.h_base[r + 1] - w.h_base[r];
        out.absent_windows[r] = w.h_absent[r];
        const uint64_t c = w.h_count[r];
        ZiKeys &kk = out.absent[r];
        kk.n = c;
        const double ta = now_ms();
        kk.lo = (uint64_t *)Device::get().host_alloc(std::max<uint64_t>(c, 1) * 8, true);
        if (wide) kk.hi = (uint64_t *)Device::get().host_alloc(std::max<uint64_t>(c, 1) * 8, true);
        require(kk.lo && (!wide || kk.hi), MCZI_E_NOMEM, "out of host memory for ZI results");
        out.alloc_ms += now_ms() - ta;
        if (c)
            MCZI_HIP(hipMemcpyAsync(kk.lo, w.ulo.as<unsigned long long>() + w.h_start[r], c * 8,
                                    hipMemcpyDeviceToHost, s));
        if (wide && c)
            MCZI_HIP(hipMemcpyAsync(kk.hi, w.ohi.as<unsigned long long>() + w.h_start[r], c * 8,
                                    hipMemcpyDeviceToHost, s));
    }
    MCZI_HIP(hipStreamSynchronize(s));
}

void zi_collect(ZiWork &w, const ZiIndex &idx, uint32_t n_req, hipStream_t s, ZiQueryResult &out) {
    const double wait_ms = zi_finish(w, idx, n_req, s);
    zi_copy_out(w, idx, n_req, s, out);
    out.wait_ms = wait_ms;
}

// ---------------------------------------------------------------------------
// Serving loop

struct ZiRequest {
    uint64_t id = 0;
    HostReads reads;        // fragment offsets (the words live in `pin`)
    uint32_t *pin = nullptr;  // packed words in pinned host memory (ordinary memory past the pin limit)
    ~ZiRequest() {
        if (pin) Device::get().host_release(pin);
    }
    double alloc_ms = 0;  // pinned allocation time of the last pin_words
    // The request's words copied out of the caller's memory at submit, into
    // pinned memory (plain DMA at admission) while the host lets the process
    // pin more, else into ordinary memory (the runtime stages that upload).
    void pin_words(const uint32_t *src, uint64_t n_words) {
        const double t0 = now_ms();
        pin = (uint32_t *)Device::get().host_alloc(std::max<uint64_t>(n_words, 1) * 4, true);
        alloc_ms = now_ms() - t0;
        const size_t grain = 1u << 20;  // words per worker chunk (4 MiB)
        const double t1 = now_ms();
        parallel_ranges(n_words, grain, host_workers(n_words), [&](size_t b, size_t e) {
            memcpy(pin + b, src + b, (e - b) * 4);
        });
        copy_ms = now_ms() - t1;
    }
    double copy_ms = 0;  // time of the copy into the request's words
    bool make_fasta = false;
    uint64_t bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    int err = 0;
    std::string err_msg;
    ZiKeys absent;
    std::string fasta;
    uint64_t windows = 0;
};


}  // namespace mczi

using namespace mczi;

struct mczi_zi_server {
    ZiIndex idx;
    uint64_t budget = 0, used = 0;
    uint32_t max_inflight = 2;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<ZiRequest *> queue;
    std::map<uint64_t, std::unique_ptr<ZiRequest>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    struct Slot {
        hipStream_t stream = nullptr;
        std::unique_ptr<ZiWork> work;
        std::vector<ZiRequest *> batch;
        uint64_t bytes = 0;
        bool busy = false;
        bool enqueued = false;  // lookup enqueued (or failed: err)
        bool finished = false;  // absent sets sorted and counted (or failed: err)
        int err = 0;
        std::string err_msg;
        double wait_ms = 0, sort_ms = 0;
        uint64_t seq = 0;
    };
    std::vector<Slot> slots;
    // Batches move through three in-order stages, each on its own thread:
    // admit (upload + lookup enqueued), finish (sort + unique once the
    // lookup is done), done (absent sets copied out, requests completed).
    // A batch's copy-out overlaps the next batch's sort on the device.
    uint64_t seq_issue = 0, seq_fin = 0, seq_done = 0;
    std::thread admit_thread, finish_thread, done_thread;
    uint64_t completed = 0, lookups = 0;
    double busy_ms = 0;
    double lookup_kernel_ms = 0;  // summed duration of the lookup kernels
    const bool trace = getenv("MCZI_ZI_TRACE") != nullptr;  // per-batch host timings on stderr
    std::string fatal;

    void admit_loop();
    void finish_loop();
    void done_loop();
};

void mczi_zi_server::admit_loop() {
    std::unique_lock<std::mutex> lk(mu);
    const uint64_t max_batch_bases = 256ull << 20;
    while (!stop) {
        int free_slot = -1;
        for (size_t i = 0; i < slots.size(); i++)
            if (!slots[i].busy) {
                free_slot = (int)i;
                break;
            }
        if (queue.empty() || free_slot < 0) {
            cv.wait(lk);
            continue;
        }
        // admit from the head while memory and the batch size allow
        std::vector<ZiRequest *> batch;
        uint64_t bytes = 0, bases = 0, frags = 0;
        while (!queue.empty()) {
            ZiRequest *r = queue.front();
            const uint64_t nb = r->reads.n_bases(), nf = r->reads.n_frag();
            const uint64_t need = zi_batch_bytes(bases + nb + 16 * (batch.size() + 1), frags + nf + batch.size() + 1,
                                                 idx.wide);
            if (zi_batch_bytes(nb + 16, nf + 1, idx.wide) > budget) {
                // can never fit
                queue.pop_front();
                r->err = MCZI_E_NOMEM;
                r->err_msg = "request needs more device memory than the server's HBM budget";
                r->done = true;
                r->t_done = now_ms();
                cv.notify_all();
                continue;
            }
            if (!batch.empty() && (used + need > budget || bases + nb > max_batch_bases)) break;
            if (batch.empty() && used + need > budget) break;
            batch.push_back(r);
            queue.pop_front();
            bytes = need;
            bases += nb;
            frags += nf;
        }
        if (batch.empty()) {
            cv.wait(lk);
            continue;
        }
        Slot &sl = slots[free_slot];
        sl.busy = true;
        sl.enqueued = false;
        sl.finished = false;
        sl.err = 0;
        sl.err_msg.clear();
        sl.batch = batch;
        sl.bytes = bytes;
        sl.seq = seq_issue++;
        used += bytes;
        const double t = now_ms();
        for (auto *r : batch) r->t_admit = t;
        lk.unlock();
        const double ta0 = now_ms();
        // lay the requests out at word boundaries (a gap before a request is a
        // fragment of no request); each request's pinned words are copied
        // straight to its place in the device buffer
        ZiBatchInput in;
        for (uint32_t ri = 0; ri < batch.size(); ri++) {
            const ZiRequest &rq = *batch[ri];
            uint64_t base = in.off.back();
            if (base % 16) {
                base += 16 - base % 16;
                in.off.push_back(base);
                in.frag_req.push_back(NO_REQ);
            }
            if (rq.reads.off[0] > 0) {  // bases before the first fragment belong to none
                in.off.push_back(base + rq.reads.off[0]);
                in.frag_req.push_back(NO_REQ);
            }
            for (uint64_t f = 0; f < rq.reads.n_frag(); f++) {
                in.off.push_back(base + rq.reads.off[f + 1]);
                in.frag_req.push_back(ri);
            }
            const uint64_t nw = (rq.reads.n_bases() + 15) / 16;
            in.spans.push_back(ZiWordSpan{rq.pin, nw, base / 16});
            in.n_words = base / 16 + nw;
        }
        int err = 0;
        std::string msg;
        const double ta1 = now_ms();
        try {
            if (!sl.work) sl.work = zi_make_work();
            zi_enqueue(*sl.work, idx, std::move(in), (uint32_t)batch.size(), sl.stream);
        } catch (const Error &e) {
            err = e.code;
            msg = e.what();
        }
        if (trace)
            fprintf(stderr, "ZI_TRACE t=%.2f admit slot=%d reqs=%zu bases=%llu concat_ms=%.2f enqueue_ms=%.2f\n", now_ms(), free_slot,
                    batch.size(), (unsigned long long)bases, ta1 - ta0, now_ms() - ta1);
        lk.lock();
        sl.err = err;  // a failed batch still passes the later stages in order
        sl.err_msg = msg;
        sl.enqueued = true;
        cv.notify_all();
    }
}

void mczi_zi_server::finish_loop() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        Slot *sl = nullptr;
        for (auto &s : slots)
            if (s.busy && s.enqueued && !s.finished && s.seq == seq_fin) sl = &s;
        if (!sl) {
            if (stop) return;
            cv.wait(lk);
            continue;
        }
        const uint32_t n_req = (uint32_t)sl->batch.size();
        const bool failed = sl->err != 0;
        lk.unlock();
        int err = 0;
        std::string msg;
        double wait = 0;
        const double t0 = now_ms();
        if (!failed) {
            try {
                wait = zi_finish(*sl->work, idx, n_req, sl->stream);
            } catch (const Error &e) {
                err = e.code;
                msg = e.what();
            }
        }
        if (!failed && !err)
            for (ZiRequest *r : sl->batch) {  // the words reached the device: their pinned block is free
                Device::get().host_release(r->pin);
                r->pin = nullptr;
            }
        lk.lock();
        if (err) {
            sl->err = err;
            sl->err_msg = msg;
        }
        sl->wait_ms = wait;
        sl->sort_ms = now_ms() - t0 - wait;
        sl->finished = true;
        seq_fin++;
        cv.notify_all();
    }
}

void mczi_zi_server::done_loop() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        // next issued batch in order
        Slot *sl = nullptr;
        for (auto &s : slots)
            if (s.busy && s.finished && s.seq == seq_done) sl = &s;
        if (!sl) {
            if (stop) return;
            cv.wait(lk);
            continue;
        }
        std::vector<ZiRequest *> batch = sl->batch;
        int err = sl->err;
        std::string msg = sl->err_msg;
        const double wait_ms = sl->wait_ms, sort_ms = sl->sort_ms;
        lk.unlock();
        ZiQueryResult res;
        const double t0 = now_ms();
        if (!err) {
            try {
                zi_copy_out(*sl->work, idx, (uint32_t)batch.size(), sl->stream, res);
            } catch (const Error &e) {
                err = e.code;
                msg = e.what();
            }
        }
        const double t1 = now_ms();
        for (size_t i = 0; i < batch.size() && !err; i++) {
            batch[i]->absent = std::move(res.absent[i]);
            batch[i]->windows = res.windows[i];
            if (batch[i]->make_fasta) {
                const auto v = batch[i]->absent.counted();
                batch[i]->fasta = simplitig_fasta_text(v.data(), v.size(), idx.k);
            }
        }
        if (trace) {
            uint64_t nb = 0;
            for (size_t i = 0; i < batch.size() && !err; i++) nb += batch[i]->absent.n * (idx.wide ? 16 : 8);
            fprintf(stderr,
                    "ZI_TRACE t=%.2f done reqs=%zu wait_ms=%.2f sort_ms=%.2f copy_ms=%.2f alloc_ms=%.2f post_ms=%.2f "
                    "copy_bytes=%llu pinned_gib=%.1f pooled_gib=%.1f\n",
                    now_ms(), batch.size(), wait_ms, sort_ms, t1 - t0, res.alloc_ms, now_ms() - t1,
                    (unsigned long long)nb, Device::get().host_slab_bytes() / double(1ull << 30),
                    Device::get().pooled_bytes() / double(1ull << 30));
        }
        lk.lock();
        busy_ms += t1 - t0 + sort_ms;
        lookups += sl->work && !err ? sl->work->lookups : 0;
        lookup_kernel_ms += sl->work && !err ? sl->work->lookup_ms : 0;
        const double t = now_ms();
        for (auto *r : batch) {
            if (err) {
                r->err = err;
                r->err_msg = msg;
            }
            r->done = true;
            r->t_done = t;
            completed++;
        }
        used -= sl->bytes;
        sl->busy = false;
        sl->enqueued = false;
        sl->finished = false;
        sl->batch.clear();
        seq_done++;
        cv.notify_all();
    }
}

namespace {
thread_local std::string z_err;

template <typename F>
int zguard(const char *ctx, F &&f) {
    try {
        f();
        return MCZI_OK;
    } catch (const Error &e) {
        z_err = std::string(ctx) + ": " + e.what();
        return e.code;
    } catch (const std::exception &e) {
        z_err = std::string(ctx) + ": " + e.what();
        return MCZI_E_INVALID;
    }
}

void check_k(uint32_t k, uint32_t m) {
    // validate_k (src/zi.rs:103-107) and the minimizer check (src/zi.rs:70-74)
    require(k >= 3 && k <= 63, MCZI_E_INVALID, "k must be in [3, 63] for sshash-rs");
    require(k % 2 == 1, MCZI_E_INVALID, "k must be odd for sshash-rs");
    require(m > 0 && m < k, MCZI_E_INVALID, "minimizer size must be > 0 and < k");
}

uint32_t default_m(uint32_t k) {
    // default_minimizer_size (src/zi.rs:109-112)
    uint32_t m = std::min<uint32_t>(19, k - 2);
    return m % 2 == 1 ? m : m - 1;
}

}  // namespace

ZiScan mczi::zi_subtract_files(const ZiIndex &idx, const char *const *qp, size_t n, hipStream_t s) {
    HostReads q;
    for (size_t i = 0; i < n; i++) parse_fastx_file(qp[i], idx.k, 0, q);
    q.finish();
    ZiBatchInput in;
    in.off = q.off;
    in.frag_req.assign(q.n_frag(), 0);
    in.n_words = (q.n_bases() + 15) / 16;
    in.spans.push_back(ZiWordSpan{q.words.data(), in.n_words, 0});
    auto w = zi_make_work();
    zi_enqueue(*w, idx, std::move(in), 1, s);
    ZiQueryResult res;
    zi_collect(*w, idx, 1, s, res);
    ZiScan out;
    out.windows = res.windows[0];
    out.absent_windows = res.absent_windows[0];
    out.absent = std::move(res.absent[0]);
    return out;
}

namespace {

// Device bytes one query pass holds per fragment of L bases: its packed
// words and offset, and the absent-k-mer segment slots and sort buffers of
// its windows.
static uint64_t zi_pass_bytes(uint64_t L, uint32_t k) {
    const uint64_t per_window = k > 32 ? 40 : 24;
    return L / 4 + 16 + (L >= k ? (L - k + 1) * per_window : 0);
}

// Query fragments [f0, f1) of q as one lookup pass; returns the sorted
// unique absent k-mers.
static std::vector<mczi_counted_kmer> zi_pass(const ZiIndex &idx, const HostReads &q, size_t f0, size_t f1,
                                              hipStream_t s) {
    // the pass starts at the word holding its first base; the bases of that
    // word before it form a leading fragment of no request (shorter than k)
    const uint64_t w0 = q.off[f0] / 16, lead = q.off[f0] - w0 * 16;
    ZiBatchInput in;
    in.off.assign(1, 0);
    if (lead) in.off.push_back(lead);
    for (size_t i = f0 + 1; i <= f1; i++) in.off.push_back(q.off[i] - w0 * 16);
    in.frag_req.assign(in.off.size() - 1, 0);
    if (lead) in.frag_req[0] = NO_REQ;
    in.n_words = (in.off.back() + 15) / 16;
    in.spans.push_back(ZiWordSpan{q.words.data() + w0, in.n_words, 0});
    auto w = zi_make_work();
    zi_enqueue(*w, idx, std::move(in), 1, s);
    ZiQueryResult res;
    zi_collect(*w, idx, 1, s, res);
    return res.absent[0].counted();
}

// The query fragments are looked up in passes of at most ram_bytes of
// device memory each (all at once when ram_bytes is 0); the sorted absent
// sets of the passes are merged.
std::vector<mczi_counted_kmer> zi_run(const char *index_path, const char *const *qp, size_t n, uint32_t k,
                                      uint64_t ram_bytes = 0) {
    Device &d = Device::get();
    ZiIndex idx;
    HostReads ir;
    parse_fastx_file(index_path, 1, 0, ir);
    zi_build_index(idx, std::move(ir), k, default_m(k), d.stream());
    HostReads q;
    for (size_t i = 0; i < n; i++) parse_fastx_file(qp[i], k, 0, q);
    q.finish();
    const size_t nf = q.n_frag();
    if (const char *e = getenv("MCZI_ZI_PASS_BYTES"))  // pass size override (tests)
        ram_bytes = strtoull(e, nullptr, 10);
    std::vector<mczi_counted_kmer> acc, merged;
    size_t f0 = 0;
    do {
        size_t f1 = f0;
        uint64_t est = 0;
        while (f1 < nf) {
            const uint64_t e = zi_pass_bytes(q.off[f1 + 1] - q.off[f1], k);
            if (ram_bytes && f1 > f0 && est + e > ram_bytes) break;
            est += e;
            f1++;
        }
        std::vector<mczi_counted_kmer> v = zi_pass(idx, q, f0, f1, d.stream());
        if (f0 == 0) {
            acc = std::move(v);
        } else {
            auto less = [](const mczi_counted_kmer &a, const mczi_counted_kmer &b) {
                return a.hi != b.hi ? a.hi < b.hi : a.lo < b.lo;
            };
            auto same = [](const mczi_counted_kmer &a, const mczi_counted_kmer &b) {
                return a.hi == b.hi && a.lo == b.lo;
            };
            merged.resize(acc.size() + v.size());
            merged.erase(std::merge(acc.begin(), acc.end(), v.begin(), v.end(), merged.begin(), less), merged.end());
            merged.erase(std::unique(merged.begin(), merged.end(), same), merged.end());
            acc.swap(merged);
        }
        f0 = f1;
    } while (f0 < nf);
    return acc;
}
}  // namespace

// The generic error accessor lives in capi.cpp; ZI errors are mirrored there.
extern "C" void mczi_set_error_from_zi(const char *msg);

extern "C" {

int mczi_zi_subtract(const char *index_path, const char *const *query_paths, size_t n, uint32_t k, uint32_t m,
                     const char *out_path, uint32_t threads, uint32_t ram_limit_gib) {
    int rc = zguard("ZI subtraction failed", [&] {
        require(index_path && out_path && (n == 0 || query_paths), MCZI_E_INVALID, "null argument");
        if (m == 0) m = default_m(k);
        check_k(k, m);
        require(n > 0, MCZI_E_INVALID, "at least one query FASTA is required");
        if (threads) g_host_threads.store(threads, std::memory_order_relaxed);  // src/zi.rs -t
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        auto t0 = std::chrono::steady_clock::now();
        auto v = zi_run(index_path, query_paths, n, k, (uint64_t)ram_limit_gib << 30);
        auto t1 = std::chrono::steady_clock::now();
        write_simplitig_fasta(out_path, v.data(), v.size(), k, 0, nullptr);
        auto t2 = std::chrono::steady_clock::now();
        fprintf(stderr, "ZI_PHASE\t2_query_subtraction\t%.6f\n", std::chrono::duration<double>(t1 - t0).count());
        fprintf(stderr, "ZI_PHASE\t3_simplitig_output\t%.6f\n", std::chrono::duration<double>(t2 - t1).count());
        fprintf(stderr, "ZI_PHASE\ttotal\t%.6f\n", std::chrono::duration<double>(t2 - t0).count());
    });
    if (rc) mczi_set_error_from_zi(z_err.c_str());
    return rc;
}

int mczi_zi_absent(const char *index_path, const char *const *query_paths, size_t n, uint32_t k,
                   mczi_counted_kmer **out, size_t *n_out) {
    int rc = zguard("ZI subtraction failed", [&] {
        require(index_path && out && n_out && (n == 0 || query_paths), MCZI_E_INVALID, "null argument");
        check_k(k, default_m(k));
        require(n > 0, MCZI_E_INVALID, "at least one query FASTA is required");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        auto v = zi_run(index_path, query_paths, n, k);
        *n_out = v.size();
        *out = (mczi_counted_kmer *)malloc(std::max<size_t>(v.size(), 1) * sizeof(mczi_counted_kmer));
        require(*out != nullptr, MCZI_E_NOMEM, "out of host memory");
        if (!v.empty()) memcpy(*out, v.data(), v.size() * sizeof(mczi_counted_kmer));
    });
    if (rc) mczi_set_error_from_zi(z_err.c_str());
    return rc;
}

static int server_create(HostReads &&ir, uint32_t k, uint32_t m, uint64_t budget, uint32_t max_inflight,
                         mczi_zi_server **out) {
    int rc = zguard("failed to create ZI server", [&] {
        require(out != nullptr, MCZI_E_INVALID, "null output pointer");
        if (m == 0) m = default_m(k);
        check_k(k, m);
        auto s = std::make_unique<mczi_zi_server>();
        Device &d = Device::get();
        {
            std::lock_guard<std::mutex> lk(d.call_mutex());
            zi_build_index(s->idx, std::move(ir), k, m, d.stream());
        }
        s->budget = budget ? budget : (8ull << 30);
        // pinned staging for request words and absent sets, taken once here:
        // pinning on the request path costs about 29 ms per GiB and stalls
        // the copies into other pinned blocks (MCZI_ZI_HOST_ARENA_GIB; default
        // 16 GiB, at most an eighth of the host's memory)
        {
            const char *e = getenv("MCZI_ZI_HOST_ARENA_GIB");
            const uint64_t host = (uint64_t)sysconf(_SC_PHYS_PAGES) * (uint64_t)sysconf(_SC_PAGESIZE);
            const uint64_t bytes = e ? strtoull(e, nullptr, 10) << 30 : std::min<uint64_t>(16ull << 30, host / 8);
            d.host_reserve(bytes);
        }
        s->max_inflight = std::max<uint32_t>(1, std::min<uint32_t>(max_inflight ? max_inflight : 2, 8));
        s->slots.resize(s->max_inflight);
        for (auto &sl : s->slots) MCZI_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        mczi_zi_server *raw = s.get();
        s->admit_thread = std::thread([raw] { raw->admit_loop(); });
        s->finish_thread = std::thread([raw] { raw->finish_loop(); });
        s->done_thread = std::thread([raw] { raw->done_loop(); });
        *out = s.release();
    });
    if (rc) mczi_set_error_from_zi(z_err.c_str());
    return rc;
}

int mczi_zi_server_create(const char *index_path, uint32_t k, uint32_t m, uint64_t hbm_budget_bytes,
                          uint32_t max_inflight, mczi_zi_server **out) {
    HostReads ir;
    int rc = zguard("failed to read ZI index", [&] {
        require(index_path != nullptr, MCZI_E_INVALID, "null index path");
        parse_fastx_file(index_path, 1, 0, ir);
    });
    if (rc) {
        mczi_set_error_from_zi(z_err.c_str());
        return rc;
    }
    return server_create(std::move(ir), k, m, hbm_budget_bytes, max_inflight, out);
}

int mczi_zi_server_create_from_text(const char *text, size_t len, uint32_t k, uint32_t m, uint64_t hbm_budget_bytes,
                                    uint32_t max_inflight, mczi_zi_server **out) {
    HostReads ir;
    int rc = zguard("failed to read ZI index", [&] {
        require(text != nullptr, MCZI_E_INVALID, "null index text");
        parse_fastx_text(text, len, 1, 0, ir);
    });
    if (rc) {
        mczi_set_error_from_zi(z_err.c_str());
        return rc;
    }
    return server_create(std::move(ir), k, m, hbm_budget_bytes, max_inflight, out);
}

int mczi_zi_server_submit(mczi_zi_server *s, const char *fastx, size_t len, int make_fasta, uint64_t *request_id) {
    int rc = zguard("failed to submit ZI request", [&] {
        require(s && request_id && (fastx || len == 0), MCZI_E_INVALID, "null argument");
        auto r = std::make_unique<ZiRequest>();
        r->t_submit = now_ms();
        parse_fastx_text(fastx, len, s->idx.k, 0, r->reads);
        r->reads.finish();
        r->pin_words(r->reads.words.data(), (r->reads.n_bases() + 15) / 16);
        decltype(r->reads.words)().swap(r->reads.words);
        r->make_fasta = make_fasta != 0;
        std::lock_guard<std::mutex> lk(s->mu);
        require(s->fatal.empty(), MCZI_E_DEVICE, s->fatal);
        r->id = s->next_id++;
        *request_id = r->id;
        if (s->trace)
            fprintf(stderr, "ZI_TRACE t=%.2f submit id=%llu bases=%llu submit_ms=%.2f pin_alloc_ms=%.2f pin_copy_ms=%.2f\n",
                    now_ms(), (unsigned long long)r->id, (unsigned long long)r->reads.n_bases(),
                    now_ms() - r->t_submit, r->alloc_ms, r->copy_ms);
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        s->cv.notify_all();
    });
    if (rc) mczi_set_error_from_zi(z_err.c_str());
    return rc;
}

int mczi_zi_server_submit_packed(mczi_zi_server *s, const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                                 int make_fasta, uint64_t *request_id) {
    int rc = zguard("failed to submit ZI request", [&] {
        require(s && request_id && frag_off && (words || n_frag == 0), MCZI_E_INVALID, "null argument");
        auto r = std::make_unique<ZiRequest>();
        r->t_submit = now_ms();
        HostReads &h = r->reads;
        for (size_t i = 0; i < n_frag; i++)
            require(frag_off[i + 1] >= frag_off[i], MCZI_E_INVALID, "fragment offsets must not decrease");
        const uint64_t nb = frag_off[n_frag];
        r->pin_words(words, (nb + 15) / 16);
        h.cur_ = nb;
        // empty fragments are dropped: the lookup's tile walk assumes at most
        // one fragment start per base position
        const std::vector<uint64_t> keep = nonempty_fragments(frag_off, n_frag);
        if (!keep.empty() || (n_frag && nb == frag_off[0])) {
            h.off.assign(1, frag_off[0]);
            for (uint64_t i : keep) h.off.push_back(frag_off[i + 1]);
        } else {
            h.off.assign(frag_off, frag_off + n_frag + 1);
        }
        r->make_fasta = make_fasta != 0;
        std::lock_guard<std::mutex> lk(s->mu);
        require(s->fatal.empty(), MCZI_E_DEVICE, s->fatal);
        r->id = s->next_id++;
        *request_id = r->id;
        if (s->trace)
            fprintf(stderr, "ZI_TRACE t=%.2f submit id=%llu bases=%llu submit_ms=%.2f pin_alloc_ms=%.2f pin_copy_ms=%.2f\n",
                    now_ms(), (unsigned long long)r->id, (unsigned long long)r->reads.n_bases(),
                    now_ms() - r->t_submit, r->alloc_ms, r->copy_ms);
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        s->cv.notify_all();
    });
    if (rc) mczi_set_error_from_zi(z_err.c_str());
    return rc;
}

int mczi_zi_server_poll(mczi_zi_server *s, uint64_t request_id, int wait, mczi_zi_result **result) {
    int code = MCZI_OK;
    int rc = zguard("failed to poll ZI request", [&] {
        require(s && result, MCZI_E_INVALID, "null argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(request_id);
        require(it != s->reqs.end(), MCZI_E_NOTFOUND, "unknown request id " + std::to_string(request_id));
        ZiRequest *r = it->second.get();
        if (!r->done) {
            if (!wait) {
                code = MCZI_E_PENDING;
                return;
            }
            s->cv.wait(lk, [&] { return r->done; });
        }
        std::unique_ptr<ZiRequest> own = std::move(it->second);
        s->reqs.erase(it);
        lk.unlock();
        if (own->err) throw Error(own->err, own->err_msg);
        auto *res = (mczi_zi_result *)calloc(1, sizeof(mczi_zi_result));
        require(res != nullptr, MCZI_E_NOMEM, "out of host memory");
        res->request_id = own->id;
        res->n_windows = own->windows;
        res->n_absent = own->absent.n;
        res->absent_lo = own->absent.lo;
        res->absent_hi = own->absent.hi;
        own->absent.lo = own->absent.hi = nullptr;
        own->absent.n = 0;
        res->fasta = (char *)malloc(own->fasta.size() + 1);
        memcpy(res->fasta, own->fasta.data(), own->fasta.size());
        res->fasta[own->fasta.size()] = 0;
        res->fasta_len = own->fasta.size();
        res->queue_ms = own->t_admit - own->t_submit;
        res->latency_ms = own->t_done - own->t_submit;
        *result = res;
    });
    if (rc) {
        mczi_set_error_from_zi(z_err.c_str());
        return rc;
    }
    if (code == MCZI_E_PENDING) mczi_set_error_from_zi("request is still running");
    return code;
}

void mczi_zi_result_free(mczi_zi_result *r) {
    if (!r) return;
    if (r->absent_lo || r->absent_hi) {
        Device &d = Device::get();
        d.host_release(r->absent_lo);
        d.host_release(r->absent_hi);
    }
    free(r->fasta);
    free(r);
}

int mczi_zi_server_stats(mczi_zi_server *s, uint64_t *index_kmers, uint64_t *completed, uint64_t *lookups,
                         double *busy_ms) {
    if (!s) return MCZI_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (index_kmers) *index_kmers = s->idx.n_kmers;
    if (completed) *completed = s->completed;
    if (lookups) *lookups = s->lookups;
    if (busy_ms) *busy_ms = s->busy_ms;
    return MCZI_OK;
}

int mczi_zi_server_lookup_stats(mczi_zi_server *s, uint64_t *lookups, double *lookup_kernel_ms) {
    if (!s) return MCZI_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (lookups) *lookups = s->lookups;
    if (lookup_kernel_ms) *lookup_kernel_ms = s->lookup_kernel_ms;
    return MCZI_OK;
}

void mczi_zi_server_destroy(mczi_zi_server *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // let in-flight batches finish
        s->cv.wait(lk, [&] {
            for (auto &sl : s->slots)
                if (sl.busy) return false;
            return true;
        });
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->admit_thread.joinable()) s->admit_thread.join();
    if (s->finish_thread.joinable()) s->finish_thread.join();
    if (s->done_thread.joinable()) s->done_thread.join();
    for (auto &sl : s->slots) {
        sl.work.reset();
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
}

}  // extern "C"
