// MC counting path on gfx950: minimizer scan, pass-1 minimizer counts,
// super-k-mer partitioning into HBM buckets, LDS-staged exact counting with
// HBM spill, abundance / dataset-presence filters, output ordering.
//
// Reference path (src/lib.rs): for_each_minimizer_run_on_seq 3694-3752 ->
// add_minimizer_counts_on_seq 3462-3481 (pass 1) -> add_filtered_superkmers_packed
// 3483-3523 / add_all_superkmers_packed 3525-3560 (pass 2) ->
// process_superkmer_partition 2400-2425 / ..._to_kff_u8_bucketed 2512-2570
// (count + filter). fof: 637-703, 4253-4370, 2184-2247.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>

#include "device_common.hpp"
#include "internal.hpp"

namespace mczi {

// ---------------------------------------------------------------------------
// Scan kernel geometry
constexpr int SCAN_THREADS = 256;
constexpr int WPT = 8;                        // windows (and hash positions) per thread
constexpr int SPAN = SCAN_THREADS * WPT;      // hash / prefix positions per tile
static_assert(WPT == 8, "the padded LDS indexing assumes 8 positions per thread");
static_assert(SCAN_THREADS * WPT <= 2048, "piece starts are packed in 11 bits");
constexpr int MAXK = 64;
constexpr int TILE = SPAN - 2 * MAXK;         // windows per tile: tl + k - 1 < SPAN
constexpr int MAXW = 64;                      // w = k - m + 1 <= 64
constexpr int WORDS_SPAN = (15 + SPAN + 160) / 16 + 4;
// LDS arrays read and written with a stride of WPT elements per thread get
// one pad element every 8, so the 64 lanes of a wave hit distinct banks.
#define PADDED(n) ((n) + (n) / 8 + 1)
__device__ __forceinline__ int P8(int i) { return i + (i >> 3); }
constexpr unsigned long long STAGE_CHUNK = 8192;  // staging slots a scan block reserves at a time
static_assert(STAGE_CHUNK >= 2 * TILE, "an abandoned staging chunk must be at least half used");
// 128-bit super-k-mer record (k <= 32): bases in MC's 2-bit code, base j at
// bits 2j: bases 0..31 in x, bases 32..43 in y bits 0..23; y bits 24..29 hold
// the base count, 30..34 the dataset (fof), 35..63 the internal partition q.
// A staging slot no record filled holds y = ~0 (partition REC_PAD_Q).
constexpr int REC_MAX_BASES = 44;
constexpr uint32_t REC_PAD_Q = 0x1FFFFFFFu;
__host__ __device__ __forceinline__ uint32_t rec_q(const ulonglong2 &r) { return (uint32_t)(r.y >> 35); }
__host__ __device__ __forceinline__ int rec_len(const ulonglong2 &r) { return (int)((r.y >> 24) & 63u); }
__host__ __device__ __forceinline__ uint32_t rec_ds(const ulonglong2 &r) { return (uint32_t)(r.y >> 30) & 31u; }

// 256-bit record for 33 <= k <= 63: bases 0..95 in w[0..2], bases 96..103 in
// w[3] bits 0..15, base count at w[3] bits 16..23, partition at bits 32..63
// (REC_PAD_QW in an unfilled slot).
constexpr int REC_MAX_BASES_W = 104;
constexpr uint32_t REC_PAD_QW = 0xFFFFFFFFu;
struct Rec32 {
    unsigned long long w[4];
};
__host__ __device__ __forceinline__ uint32_t rec_q(const Rec32 &r) { return (uint32_t)(r.w[3] >> 32); }
__host__ __device__ __forceinline__ int rec_len(const Rec32 &r) { return (int)((r.w[3] >> 16) & 255u); }

// Internal partitions per batch: at most 2^22 (two 11-bit digits).
constexpr uint32_t MAX_INTERNAL_PARTITIONS = 1u << 22;

// Count kernel geometry
constexpr int CNT_THREADS = 256;
constexpr int CNT_LOG2_SLOTS = 11;
constexpr int CNT_SLOTS = 1 << CNT_LOG2_SLOTS;  // LDS table slots (u64 key + u32 value)
constexpr int CNT_MAXPROBE = 32;
constexpr int OVF_MAXPROBE = 1 << 12;  // at most half-full tables keep probe runs far shorter

enum ScanMode {
    STAGE = 0,    // super-k-mer records + partition ids (or only their count)
    WINDOWS = 1,  // selected hash of every window
    RUNS = 2      // pass 1: (run key, windows or dataset) of every piece
};

struct ScanArgs {
    const uint32_t *words;
    const uint64_t *off;
    const uint32_t *dataset;
    const uint64_t *tile_start;
    const uint32_t *tile_frag;
    const uint32_t *tile_len;
    uint64_t tile_begin, tile_end;
    uint64_t n_frag;
    int k, m, w;
    uint32_t x;
    uint32_t sat;
    uint32_t p_int;
    uint32_t p_mask;  // p_int - 1 when p_int is a power of two, else 0
    uint32_t wrec;
    uint32_t ds_lo;   // fof: first dataset of the group of 32 being staged
    // partition-group passes: stage only the pieces whose key hash h has
    // (h * grp_n) >> 32 == grp_i (grp_n <= 1: every piece)
    uint32_t grp_n, grp_i;
    uint32_t *win_out;
    unsigned long long *stats;  // [0] pieces [4] records [5] .skm bytes of the staged runs
    // pass 1 (RUNS): one entry per piece, key << 32 | (windows, or dataset with fof)
    unsigned long long *run_out;
    uint64_t run_cap;
    unsigned long long *run_ctr;
    // pass-1 filter of the two-pass mode: set of the keys that pass (key + 1
    // per slot, 0 = empty), or null
    const unsigned long long *mtab;
    uint64_t mtab_mask;
    // optional, per tile: the first and last staged piece when the run they
    // belong to may continue across the tile edge (key << 32 | windows, bit
    // 31 = the tile's only piece; ~0 = none)
    unsigned long long *tile_first, *tile_last;
    // staging of super-k-mer records and their internal partition
    ulonglong2 *stage_rec;
    Rec32 *stage_rec_w;  // k > 32
    uint32_t *stage_q;   // the partition of each staging slot (padding: the pad partition), or null
    uint64_t stage_cap;
    unsigned long long *stage_ctr;  // record slots reserved so far (the batch total at the end)
    unsigned long long *err;
};

__device__ __forceinline__ uint32_t lds_code(const uint32_t *sw, int i) {
    return (sw[i >> 4] >> ((i & 15) * 2)) & 3u;
}

__device__ __forceinline__ uint64_t high_bit_mask(int nbases) {
    // bit 2j+1 set for every base j < nbases (the T/G bit of a packed code)
    const uint64_t all = 0xAAAAAAAAAAAAAAAAull;
    return nbases >= 32 ? all : (all & ((1ull << (2 * nbases)) - 1ull));
}

// Inclusive scans over the 64 lanes of a wave with DPP row shifts (within
// 16-lane rows) and row broadcasts (lane 15 -> next row, lane 31 -> rows 2-3).
__device__ __forceinline__ int wave_incl_add(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}
__device__ __forceinline__ uint32_t wave_incl_xor(uint32_t x) {
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}
// (x >= 0 for every lane)
__device__ __forceinline__ int wave_incl_max(int x) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
    return x;
}

// Exclusive block scan of one int per thread (SCAN_THREADS threads).
__device__ __forceinline__ int block_exclusive_scan(int v, int *total, int *s_wsum) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int x = wave_incl_add(v);
    if (lane == 63) s_wsum[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_THREADS / 64; i++) {
        int s = s_wsum[i];
        pre += (i < wid) ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// Block-wide sum of one value per thread; thread 0 adds it to *dst.
__device__ __forceinline__ void block_sum_atomic(unsigned long long v, unsigned long long *dst,
                                                 unsigned long long *s_red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += s_red[i];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}


// 8 bases starting at base index `pos` of the LDS word stream (base j at
// bits 2j of the result).
__device__ __forceinline__ uint32_t extract16_bases(const uint32_t *sw, int pos) {
    const int w = pos >> 4;
    const uint32_t sh = (uint32_t)(pos & 15) * 2u;
    const uint64_t v = (uint64_t)sw[w] | ((uint64_t)sw[w + 1] << 32);
    return (uint32_t)(v >> sh) & 0xffffu;
}

// 32 bases starting at base `pos` of the LDS word stream (base j at bits 2j):
// two 32-bit funnel shifts over three words, no branch on the bit offset. The
// word after the last one read must exist (the word stream is padded).
__device__ __forceinline__ uint64_t lds_bases32(const uint32_t *sw, int pos) {
    const int i = pos >> 4;
    const uint32_t sh = (uint32_t)(pos & 15) * 2u;
    const uint32_t w0 = sw[i], w1 = sw[i + 1], w2 = sw[i + 2];
    return ((uint64_t)__builtin_amdgcn_alignbit(w2, w1, sh) << 32) | __builtin_amdgcn_alignbit(w1, w0, sh);
}

// Bytes of a super-k-mer of L bases in a partition file (append_record_len
// and append_packed_dna_range, src/lib.rs:3170-3216).
__device__ __forceinline__ uint64_t skm_bytes(uint32_t L) { return (L < 255u ? 1u : 5u) + (L + 3u) / 4u; }

// Pass-1 filter: whether key is in the set of keys above the threshold.
__device__ __forceinline__ bool mtab_has(const ScanArgs &a, uint32_t key) {
    uint64_t i = spread_hash_u64(key) & a.mtab_mask;
    for (;;) {
        const unsigned long long v = a.mtab[i];
        if (v == 0) return false;
        if (v == (unsigned long long)key + 1) return true;
        i = (i + 1) & a.mtab_mask;
    }
}

// Staging slots [b, e) that no record fills (the rest of an abandoned chunk)
// are marked with the padding partition, which partitioning skips.
template <bool WIDE>
__device__ __forceinline__ void stage_pad(const ScanArgs &a, uint64_t b, uint64_t e) {
    for (uint64_t i = b + threadIdx.x; i < e && i < a.stage_cap; i += SCAN_THREADS) {
        if constexpr (WIDE) a.stage_rec_w[i] = Rec32{{0ull, 0ull, 0ull, ~0ull}};
        else a.stage_rec[i] = make_ulonglong2(0ull, ~0ull);
        if (a.stage_q) a.stage_q[i] = WIDE ? REC_PAD_QW : REC_PAD_Q;
    }
}

// One tile = up to TILE consecutive base positions of the segmented stream,
// clipped to its input chunk. For every window (k-mer start) p of the tile:
// canonical ntHash of each m-mer, #TG strand vote over the k bases,
// leftmost / rightmost minimum of the upper 16 hash bits over the w m-mers
// (sm/sliding_min.rs:279-340, sm/canonical.rs:11-29, sm/minimizers.rs:213-216),
// then runs of equal selected hash (sm/collect.rs:38-75) cut at fragment and
// tile boundaries ("pieces"). STAGE writes the pieces as super-k-mer records,
// WINDOWS the selected hash of every window. GW selects the window-minimum
// scheme (w >= WPT: Gil-Werman over the thread's windows; else a direct scan
// of the w hashes). STAGE serves the counting path, which has k <= 32.
// WC != 0 fixes the window length w at compile time (the reference default
// k = 31, m = 21 has w = 11), which unrolls the minimum over the middle
// hashes and makes every s_hash offset a constant.
// A piece of a tile as staging needs it: windows [s, e), its internal
// partition q and dataset within the group, and whether it is staged.
struct PieceInfo {
    int s, e;
    uint32_t q, ds;
    bool staged;
};

// Piece pi of the tile (pst: starts | fragment << 11; v: selected hash per
// window): its window range clipped to its fragment's last window, whether it
// is staged (fof: the current group of 32 datasets; two-pass: a key above the
// threshold, add_filtered_superkmers_packed, src/lib.rs:3483-3523; partition
// groups: the current group), its partition and its record / byte counts.
__device__ __forceinline__ PieceInfo piece_info(const ScanArgs &a, int pi, int np, int tl, int k, uint64_t f0,
                                                uint64_t t, const uint32_t *pst, const int32_t *s_fs,
                                                const uint8_t *s_ds,
                                                const uint32_t *s_v, const uint8_t *s_flag, int &my_nrec,
                                                unsigned long long &my_bytes) {
    PieceInfo P{0, 0, 0u, 0u, false};
    const uint32_t pw = pst[pi];
    P.s = (int)(pw & 2047u);
    const int jf = (int)(pw >> 11);
    int e = (pi + 1 < np) ? (int)(pst[pi + 1] & 2047u) : tl;
    const int vend = s_fs[jf + 1] - k + 1;
    if (vend < e) e = vend;
    P.e = e;
    if (a.dataset != nullptr) {
        P.ds = jf < SCAN_THREADS ? (uint32_t)s_ds[jf] : a.dataset[f0 + jf] - a.ds_lo;
        if (P.ds >= 32u) return P;
    }
    const uint32_t key = spread_hash_u32(s_v[P8(P.s)]);
    if (a.mtab != nullptr && !mtab_has(a, key)) return P;
    const uint32_t hq = spread_hash_u32(key);
    if (a.grp_n > 1 && (uint32_t)(((uint64_t)hq * a.grp_n) >> 32) != a.grp_i) return P;
    P.staged = true;
    P.q = a.p_mask ? (hq & a.p_mask) : hq % a.p_int;
    my_nrec += e - P.s <= (int)a.wrec ? 1 : (e - P.s + (int)a.wrec - 1) / (int)a.wrec;
    my_bytes += skm_bytes((uint32_t)(e - P.s + k - 1));
    if (a.tile_first != nullptr) {
        // a run cut by the tile edge: the first piece starts at window 0
        // inside a fragment that began earlier, the last one ends at the tile
        // end inside a fragment that goes on
        const unsigned long long ent =
            ((unsigned long long)key << 32) | (uint32_t)(e - P.s) | (np == 1 ? 0x80000000ull : 0ull);
        if (pi == 0 && P.s == 0 && !(s_flag[P8(0)] & 2)) a.tile_first[t] = ent;
        if (pi == np - 1 && e == tl && vend > tl) a.tile_last[t] = ent;
    }
    return P;
}

// The records of a staged piece from slot pos on (bases [start, end + k - 1)
// of each chunk of at most wrec windows, src/lib.rs:3483-3560); returns the
// next slot. Bases are stored in MC's 2-bit code (A0 C1 G2 T3,
// src/lib.rs:3192-3197), base j of the record at bits 2j.
template <bool WIDE>
__device__ __forceinline__ uint64_t write_piece(const ScanArgs &a, const PieceInfo &P, uint64_t pos,
                                                const uint32_t *s_words, int rb0, int k) {
    for (int cs = P.s; cs < P.e; cs += (int)a.wrec) {
        const int ce = min(cs + (int)a.wrec, P.e);
        const uint32_t L = (uint32_t)(ce - cs + k - 1);
        if constexpr (WIDE) {
            Rec32 rw;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                uint64_t x = lds_bases32(s_words, rb0 + cs + 32 * c);
                const int nbc = (int)L - 32 * c;  // bases of this word
                if (nbc <= 0) x = 0;
                else if (nbc < 32) x &= (1ull << (2 * nbc)) - 1ull;
                rw.w[c] = x ^ ((x >> 1) & 0x5555555555555555ull);  // packed code -> MC code
            }
            rw.w[3] = (rw.w[3] & 0xffffull) | ((uint64_t)L << 16) | ((uint64_t)P.q << 32);
            if (pos < a.stage_cap) {
                a.stage_rec_w[pos] = rw;
                if (a.stage_q) a.stage_q[pos] = P.q;
            } else {
                atomicOr(a.err, 2ull);
            }
        } else {
            uint64_t lo64 = lds_bases32(s_words, rb0 + cs);
            uint64_t hi64 = 0;
            if (L < 32) lo64 &= (1ull << (2 * L)) - 1ull;
            if (L > 32) {
                hi64 = lds_bases32(s_words, rb0 + cs + 32);
                hi64 &= (1ull << (2 * (L - 32))) - 1ull;
            }
            lo64 ^= (lo64 >> 1) & 0x5555555555555555ull;  // packed code -> MC code
            hi64 ^= (hi64 >> 1) & 0x5555555555555555ull;
            hi64 |= ((uint64_t)L << 24) | ((uint64_t)P.ds << 30) | ((uint64_t)P.q << 35);
            if (pos < a.stage_cap) {
                a.stage_rec[pos] = make_ulonglong2(lo64, hi64);
                if (a.stage_q) a.stage_q[pos] = P.q;
            } else {
                atomicOr(a.err, 2ull);
            }
        }
        pos++;
    }
    return pos;
}

template <int MODE, bool GW, bool WIDE = false, int WC = 0>
__global__ __launch_bounds__(SCAN_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_scan(ScanArgs a) {
    __shared__ uint32_t s_words[WORDS_SPAN];
    __shared__ uint32_t s_hash[PADDED(SPAN)];
    // prefix XORs live only until the hashes are written; the selected
    // hashes and piece starts reuse their space
    __shared__ union {
        struct {
            uint32_t sa[PADDED(SPAN + MAXW)];
            uint32_t sb[PADDED(SPAN + MAXW)];
        } pre;
        struct {
            uint32_t v[PADDED(SPAN)];
            uint32_t pst[TILE];
        } run;
    } s_u;
    __shared__ uint32_t s_wxor[2][SCAN_THREADS / 64];  // per-wave XOR totals of the two prefix streams
    __shared__ uint8_t s_flag[PADDED(SPAN)];
    __shared__ int32_t s_fs[TILE + 2];
    __shared__ uint8_t s_ds[SCAN_THREADS];  // fof: dataset - ds_lo of the tile's first fragments (255: other group)
    __shared__ int s_wsum[SCAN_THREADS / 64];
    __shared__ unsigned long long s_sbase, s_hole_b, s_hole_e;
    // ntHash terms by (tid % 4, position in the thread, base code):
    // (rotr(F[c], a), rotl(F[c ^ 2], a)) with a = 7 (8 tid + t) mod 32
    __shared__ uint2 s_seed[4 * WPT * 4];

    const int tid = threadIdx.x;
    if (tid < 4 * WPT * 4) {
        const uint32_t c = (uint32_t)tid & 3u, tt = ((uint32_t)tid >> 2) % WPT, row = (uint32_t)tid / (WPT * 4);
        const uint32_t amt = (7u * (8u * row + tt)) & 31u;
        s_seed[tid] = make_uint2(__builtin_rotateright32(nt_seed(c), amt), __builtin_rotateleft32(nt_seed(c ^ 2u), amt));
    }
    const int k = a.k, m = a.m, w = WC ? WC : a.w;
    unsigned long long my_pieces = 0, my_records = 0, my_bytes = 0;
    unsigned long long chunk_cur = 0, chunk_end = 0;  // thread 0: the block's current staging chunk
    // each thread owns positions r0 .. r0 + WPT - 1 of a tile; P8(r0 + c),
    // the padded LDS index, is 9 * tid + c + floor(c / 8) since r0 = 8 * tid
    const int r0 = tid * WPT;
    auto px = [&](int c) { return 9 * tid + c + (c >> 3); };

    // A tile's schedule entry is loaded one tile ahead, and its words and
    // first fragment offsets while the tile before it is scanned, so a tile
    // starts with no global round trip: each thread holds one word
    // (WORDS_SPAN <= SCAN_THREADS) and one offset.
    static_assert(WORDS_SPAN <= SCAN_THREADS, "one staged word per thread");
    // Only the fragment offsets a tile can need are loaded: its fragments
    // lie in [tile_frag[t], tile_frag[t + 1]] (the next tile starts at or
    // after its end), so entries 0 .. tile_frag[t + 1] - tile_frag[t] + 1;
    // the rest read as "beyond the tile".
    uint64_t t = a.tile_begin + blockIdx.x;
    uint64_t ts = 0, f0 = 0, f1 = 0;
    int tl = 0;
    auto next_frag = [&](uint64_t t_) -> uint64_t { return t_ + 1 < a.tile_end ? a.tile_frag[t_ + 1] : a.n_frag; };
    if (t < a.tile_end) {
        ts = a.tile_start[t];
        tl = (int)a.tile_len[t];
        f0 = a.tile_frag[t];
        f1 = next_frag(t);
    }
    uint32_t pword = 0;
    uint64_t poff = 0;
    // the dataset of fragment f0 + tid of the next tile (ds_lo + 255: none);
    // it is mapped into the staged group of 32 when the tile starts
    uint32_t pds = a.ds_lo + 255u;
    // (the loads are issued before any of their values is used, so that all
    // stay in flight while the current tile is scanned)
    auto load_first = [&](uint64_t ts_, int tl_, uint64_t f0_, uint64_t f1_) {
        const uint64_t j = f0_ + (uint64_t)tid;
        poff = (j <= f1_ + 1 && j <= a.n_frag) ? a.off[j] : (~0ull >> 1);
        pds = a.ds_lo + 255u;
        if (MODE == STAGE && a.dataset != nullptr && j <= f1_ && j < a.n_frag) pds = a.dataset[j];
        const int nw = (((int)(ts_ & 15u) + tl_ + 160) >> 4) + 3;
        pword = tid < nw ? a.words[(ts_ >> 4) + tid] : 0u;
    };
    if (t < a.tile_end) load_first(ts, tl, f0, f1);
    uint64_t ts_n = 0, f0_n = 0, f1_n = 0;  // the next tile's entry
    int tl_n = 0;
    for (; t < a.tile_end; t += gridDim.x, ts = ts_n, tl = tl_n, f0 = f0_n, f1 = f1_n) {
        const uint64_t tn = t + gridDim.x;
        // The next tile's entry, loaded unconditionally (clamped to the last
        // tile; unused past it), so that no wait on these loads opens the
        // tile: they are first used once the fragment bounds are in LDS.
        const uint64_t tc = min(tn, a.tile_end - 1);
        ts_n = a.tile_start[tc];
        tl_n = (int)a.tile_len[tc];
        f0_n = a.tile_frag[tc];
        const uint32_t f1_next = a.tile_frag[min(tn + 1, a.tile_end - 1)];
        const int rb0 = (int)(ts & 15u);
        __syncthreads();  // previous tile fully consumed
        if (tid < WORDS_SPAN) s_words[tid] = pword;
        if (MODE == STAGE) {
            const uint32_t d = pds - a.ds_lo;
            s_ds[tid] = (uint8_t)(d < 32u ? d : 255u);
        }
        // Fragment boundaries relative to ts: s_fs[j] = off[f0 + j] - ts (clamped),
        // for j = 0 .. nf, nf = number of fragments with a base in [ts, ts + tl).
        int nf = 0;
        for (int j0 = 0;; j0 += SCAN_THREADS) {
            const int j = j0 + tid;
            bool in = false;
            if (j <= TILE + 1 && f0 + (uint64_t)j <= a.n_frag) {
                const int64_t rel = (int64_t)(j0 == 0 ? poff : a.off[f0 + j]) - (int64_t)ts;
                s_fs[j] = rel < 0 ? -1 : (rel > (1 << 30) ? (1 << 30) : (int32_t)rel);
                in = rel < tl;
            }
            const int c = __syncthreads_count(in);
            nf += c;
            if (c < SCAN_THREADS) break;
        }
        __syncthreads();
        // (after the block's last tile this prefetches the clamped entry, unused)
        f1_n = tn + 1 < a.tile_end ? (uint64_t)f1_next : a.n_frag;
        load_first(ts_n, tl_n, f0_n, f1_n);

        // m-mer hashes for relative positions i in [0, tl + w - 1). ntHash is
        // XOR-linear, so with the exclusive prefix XORs over the tile's bases
        //   SA_p = XOR_{p' < p} rotr(F[c_p'], 7p'),  SB_p = XOR_{p' < p} rotl(F[c_p' ^ 2], 7p')
        // (rotations mod 32) the canonical hash NtHasher<true> rolls is
        //   fw(i) = rotl(SA_{i+m} ^ SA_i, 7(i+m-1)),  rc(i) = rotr(SB_{i+m} ^ SB_i, 7i),
        //   h(i) = fw(i) + rc(i)
        // (vendor/simd-minimizers-2.3.1/src/minimizers.rs:258-291): one term
        // per base and one combination per position, no per-thread m-base
        // warm-up.
        {
            const uint32_t bw = extract16_bases(s_words, rb0 + r0);
            // the terms of position r0 + t come from the thread's row of the
            // seed table (the rotation 7 (r0 + t) depends on tid % 4 and t)
            const uint2 *seed_row = s_seed + (tid & 3) * (WPT * 4);
            uint32_t ta[WPT], tb[WPT], xa = 0, xb = 0;
#pragma unroll
            for (int t = 0; t < WPT; t++) {
                ta[t] = xa;
                tb[t] = xb;
                const uint2 fg = seed_row[t * 4 + ((bw >> (2 * t)) & 3u)];
                xa ^= fg.x;
                xb ^= fg.y;
            }
            // exclusive block XOR scan of (xa, xb): DPP within the wave, then
            // the totals of the waves before this one
            const int lane = tid & 63, wid = tid >> 6;
            const uint32_t ia = wave_incl_xor(xa), ib = wave_incl_xor(xb);
            if (lane == 63) {
                s_wxor[0][wid] = ia;
                s_wxor[1][wid] = ib;
            }
            __syncthreads();
            uint32_t pa = ia ^ xa, pb = ib ^ xb;
#pragma unroll
            for (int q = 0; q < SCAN_THREADS / 64; q++)
                if (q < wid) {
                    pa ^= s_wxor[0][q];
                    pb ^= s_wxor[1][q];
                }
#pragma unroll
            for (int t = 0; t < WPT; t++) {
                s_u.pre.sa[px(t)] = pa ^ ta[t];
                s_u.pre.sb[px(t)] = pb ^ tb[t];
            }
            __syncthreads();
            uint32_t ai = (uint32_t)(7 * r0) & 31u;
            const uint32_t am = (uint32_t)(7 * (m - 1)) & 31u;
#pragma unroll
            for (int t = 0; t < WPT; t++) {
                // positions past nh - 1 get values that no window reads (the
                // arrays are padded for them)
                const uint32_t A = s_u.pre.sa[px(t + m)] ^ pa ^ ta[t];
                const uint32_t Bv = s_u.pre.sb[px(t + m)] ^ pb ^ tb[t];
                s_hash[px(t)] = __builtin_rotateleft32(A, (ai + am) & 31u) + __builtin_rotateright32(Bv, ai);
                ai = (ai + 7u) & 31u;
            }
        }
        __syncthreads();

        // selected hash per window. Keys compare the upper 16 hash bits, then
        // the position: (h & 0xffff0000) | t gives the leftmost minimum,
        // (h & 0xffff0000) | (0xffff - t) the rightmost (sm/sliding_min.rs:
        // 279-340). With w >= WPT the thread's WPT windows share the middle
        // hashes [WPT-1, w-1]; suffix minima on the left and prefix minima on
        // the right complete each window (van Herk / Gil-Werman).
        int jfirst = 0;     // fragment holding the thread's first window
        uint32_t smask = 0;  // piece starts among the thread's windows (bit 0 provisional)
        uint32_t jd = 0;     // 4 bits per window: its fragment minus jfirst (at most WPT)
        if (r0 < tl) {
            int lo = 0, hi = nf - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_fs[mid] <= r0) lo = mid;
                else hi = mid - 1;
            }
            jfirst = lo;
            int j = lo;
            int fcur = s_fs[j], fnext = s_fs[j + 1];  // bounds of the fragment holding r
            const int rb = rb0 + r0;
            // Strand vote: the number of T/G bases (bit 2b+1 of each code) in
            // window i's bases [i, i + k). Window 0 counts them; window i
            // then drops base i - 1 (from H0) and adds base i - 1 + k (from
            // E, the bases from k on).
            constexpr bool KBIG = MODE == WINDOWS || WIDE;  // k may exceed 32
            const uint64_t H0 = lds_bases32(s_words, rb) & 0xAAAAAAAAAAAAAAAAull;
            const uint32_t E = (uint32_t)lds_bases32(s_words, rb + k);
            uint32_t tg = (uint32_t)__popcll(H0 & high_bit_mask(k < 32 ? k : 32));
            if (KBIG && k > 32)
                tg += (uint32_t)__popcll(lds_bases32(s_words, rb + 32) & high_bit_mask(k - 32));
            // suffix minima over [i, w-1] (the shared middle folded in); prefix
            // minima over [w, w+i-1] are kept running in the window loop
            uint32_t sufL[WPT], sufR[WPT], preL = ~0u, preR = ~0u;
            uint32_t pv = 0;
            bool pvalid = false;
            if constexpr (GW) {
                uint32_t midL = ~0u, midR = ~0u;
                for (int t = WPT - 1; t <= w - 1; t++) {
                    const uint32_t h = s_hash[px(t)] & 0xffff0000u;
                    midL = min(midL, h | (uint32_t)t);
                    midR = min(midR, h | (0xffffu - (uint32_t)t));
                }
                sufL[WPT - 1] = midL;
                sufR[WPT - 1] = midR;
#pragma unroll
                for (int i = WPT - 2; i >= 0; i--) {
                    const uint32_t h = s_hash[px(i)] & 0xffff0000u;
                    sufL[i] = min(sufL[i + 1], h | (uint32_t)i);
                    sufR[i] = min(sufR[i + 1], h | (0xffffu - (uint32_t)i));
                }
            }
#pragma unroll
            for (int i = 0; i < WPT; i++) {
                const int r = r0 + i;
                if (GW && i > 0) {
                    const int t = w + i - 1;
                    const uint32_t h = s_hash[px(t)] & 0xffff0000u;
                    preL = min(preL, h | (uint32_t)t);
                    preR = min(preR, h | (0xffffu - (uint32_t)t));
                }
                // windows past the tile end are computed as invalid (no break,
                // so the loop has no nested exec masks; the LDS arrays are
                // padded for them)
                if (i > 0) tg = tg + ((E >> (2 * i - 1)) & 1u) - (((uint32_t)H0 >> (2 * i - 1)) & 1u);
                // fragments are never empty (batches drop empty ones), so one
                // window step crosses at most one fragment start
                if (j + 1 < nf && fnext <= r) {
                    j++;
                    fcur = fnext;
                    fnext = s_fs[j + 1];
                }
                jd |= (uint32_t)(j - jfirst) << (4 * i);
                const bool fstart = fcur == r;
                const bool valid = r + k <= fnext && r < tl;
                uint32_t v = 0;
                if constexpr (GW) {
                    // read for every window (the positions are in the tile's
                    // hashes either way); kept only for valid ones
                    const int lp = (int)(min(sufL[i], preL) & 0xffffu);
                    const int rp = (int)(0xffffu - (min(sufR[i], preR) & 0xffffu));
                    v = s_hash[px((2u * tg > (uint32_t)k) ? lp : rp)];
                    v = valid ? v : 0u;
                } else if (valid) {
                    int lp, rp;
                    {
                        uint32_t best = s_hash[px(i)] >> 16;
                        lp = rp = i;
                        for (int ii = 1; ii < w; ii++) {
                            const uint32_t hv = s_hash[px(i + ii)] >> 16;
                            if (hv < best) {
                                best = hv;
                                lp = rp = i + ii;
                            } else if (hv == best) {
                                rp = i + ii;
                            }
                        }
                    }
                    v = s_hash[px((2u * tg > (uint32_t)k) ? lp : rp)];
                }
                s_u.run.v[px(i)] = v;
                s_flag[px(i)] = (uint8_t)((valid ? 1 : 0) | (fstart ? 2 : 0));
                // a piece starts at a valid window whose predecessor is
                // invalid, in another fragment or selects another hash (the
                // first window's predecessor is the previous thread's last)
                if (valid && (i == 0 || fstart || !pvalid || pv != v)) smask |= 1u << i;
                pv = v;
                pvalid = valid;
            }
        }
        __syncthreads();

        if constexpr (MODE == WINDOWS) {
            for (int r = tid; r < tl; r += SCAN_THREADS)
                if (s_flag[P8(r)] & 1) a.win_out[ts + (uint64_t)r] = s_u.run.v[P8(r)];
            continue;
        }

        // piece starts -> compacted list of (start | fragment << 11), the
        // fragment index inside the tile walked forward from the thread's
        // first fragment
        if ((smask & 1u) && r0 > 0 && !(s_flag[px(0)] & 2) && (s_flag[px(-1)] & 1) &&
            s_u.run.v[px(-1)] == s_u.run.v[px(0)])
            smask &= ~1u;
        const int cnt = __popc(smask);
        int np = 0;
        int o = block_exclusive_scan(cnt, &np, s_wsum);
        for (uint32_t sm = smask; sm; sm &= sm - 1) {
            const int i = __builtin_ctz(sm);
            const uint32_t j = (uint32_t)jfirst + ((jd >> (4 * i)) & 15u);
            s_u.run.pst[o++] = (uint32_t)(r0 + i) | (j << 11);
        }
        __syncthreads();

        // Pieces -> super-k-mer records of at most wrec windows (bases
        // [start, end + k - 1), src/lib.rs:3483-3560), staged contiguously per
        // tile together with their partition: key =
        // spread_hash_u32(minimizer hash) (visit_direct_hash_runs), partition
        // = spread_hash_u32(key) % P (minimizer_shard, src/lib.rs:2927-2929).
        // Bases are stored in MC's 2-bit code (A0 C1 G2 T3, src/lib.rs:
        // 3192-3197), base j of the record at bits 2j.
        if constexpr (MODE == RUNS) {
            // pass 1 (add_minimizer_counts_on_seq, src/lib.rs:3462-3481, and
            // the dataset presence of src/lib.rs:4291-4341): every piece's
            // key with its window count (or its dataset), one reservation
            // per tile
            const int mine = np > tid ? (np - tid + SCAN_THREADS - 1) / SCAN_THREADS : 0;
            int tot_r = 0;
            const int off_r = block_exclusive_scan(mine, &tot_r, s_wsum);
            if (tid == 0) s_sbase = tot_r ? atomicAdd(a.run_ctr, (unsigned long long)tot_r) : 0ull;
            __syncthreads();
            uint64_t pos = s_sbase + (uint64_t)off_r;
            for (int pi = tid; pi < np; pi += SCAN_THREADS, pos++) {
                const uint32_t pw = s_u.run.pst[pi];
                const int s = (int)(pw & 2047u);
                int e = (pi + 1 < np) ? (int)(s_u.run.pst[pi + 1] & 2047u) : tl;
                const int vend = s_fs[(pw >> 11) + 1] - k + 1;
                if (vend < e) e = vend;
                my_pieces++;
                const uint32_t key = spread_hash_u32(s_u.run.v[P8(s)]);
                const uint32_t val = a.dataset != nullptr ? a.dataset[f0 + (pw >> 11)] : (uint32_t)(e - s);
                if (pos < a.run_cap) a.run_out[pos] = ((unsigned long long)key << 32) | val;
            }
            continue;
        }
        // Pieces are staged in rounds of 2 * SCAN_THREADS (one round for
        // almost every tile): each thread takes two pieces of the round, keeps
        // their data in registers between counting and writing their records,
        // and the round's records get consecutive staging slots.
        my_pieces += (uint64_t)(np > tid ? (np - tid + SCAN_THREADS - 1) / SCAN_THREADS : 0);
        for (int rb = 0; rb < np; rb += 2 * SCAN_THREADS) {
            int my_nrec = 0;
            PieceInfo P0{0, 0, 0u, 0u, false}, P1{0, 0, 0u, 0u, false};
            if (rb + tid < np)
                P0 = piece_info(a, rb + tid, np, tl, k, f0, t, s_u.run.pst, s_fs, s_ds, s_u.run.v, s_flag, my_nrec,
                                my_bytes);
            if (rb + tid + SCAN_THREADS < np)
                P1 = piece_info(a, rb + tid + SCAN_THREADS, np, tl, k, f0, t, s_u.run.pst, s_fs, s_ds, s_u.run.v, s_flag,
                                my_nrec, my_bytes);
            int tot = 0;
            const int off = block_exclusive_scan(my_nrec, &tot, s_wsum);
            // Staging slots: each block takes chunks of STAGE_CHUNK slots (or
            // the round's records, if more) from the batch-wide counter and
            // hands out consecutive slots to its rounds. A round that does not
            // fit in the rest of the chunk starts a new one; the rest is marked
            // as padding, which partitioning skips. Slots at or beyond
            // stage_cap are counted but not written: the host then sizes the
            // buffer from the counts and scans again. The order of records in
            // the buffer is free: partitioning and counting do not depend on it.
            if (tid == 0) {
                s_hole_b = s_hole_e = 0;
                unsigned long long base = 0;
                if (tot) {
                    if (chunk_cur + (unsigned long long)tot > chunk_end) {
                        s_hole_b = chunk_cur;
                        s_hole_e = chunk_end;
                        const unsigned long long sz = (unsigned long long)tot > STAGE_CHUNK ? tot : STAGE_CHUNK;
                        chunk_cur = atomicAdd(a.stage_ctr, sz);
                        chunk_end = chunk_cur + sz;
                    }
                    base = chunk_cur;
                    chunk_cur += (unsigned long long)tot;
                }
                s_sbase = base;
                my_records += (unsigned long long)tot;
            }
            __syncthreads();
            stage_pad<WIDE>(a, s_hole_b, s_hole_e);
            uint64_t pos = s_sbase + (uint64_t)off;
            if (P0.staged) pos = write_piece<WIDE>(a, P0, pos, s_words, rb0, k);
            if (P1.staged) pos = write_piece<WIDE>(a, P1, pos, s_words, rb0, k);
            if (rb + 2 * SCAN_THREADS < np) __syncthreads();  // s_sbase / s_hole of the next round
        }
    }
    if constexpr (MODE == STAGE) {
        __syncthreads();
        if (tid == 0) {
            s_hole_b = chunk_cur;
            s_hole_e = chunk_end;
        }
        __syncthreads();
        stage_pad<WIDE>(a, s_hole_b, s_hole_e);
    }
    __shared__ unsigned long long s_red[SCAN_THREADS / 64];
    block_sum_atomic(my_pieces, &a.stats[0], s_red);
    if constexpr (MODE == STAGE) {
        block_sum_atomic(my_records, &a.stats[4], s_red);
        block_sum_atomic(my_bytes, &a.stats[5], s_red);
    }
}

// ---------------------------------------------------------------------------
// Partition-file bytes of runs a tile edge cut into pieces (the scan counted
// each piece as a super-k-mer; the reference writes one per run,
// add_packed_range, src/lib.rs:4893-4917). Each chain of pieces joined across tile
// edges is walked by the thread of its first edge; the correction (bytes of
// the joined run minus those of its pieces) is added to *corr as a two's
// complement 64-bit value.
__global__ void k_tile_join(const unsigned long long *first, const unsigned long long *last, uint64_t n_tiles,
                            uint32_t k, unsigned long long *corr) {
    constexpr unsigned long long NONE = ~0ull, SINGLE = 0x80000000ull;
    unsigned long long my_corr = 0;
    auto joins = [&](uint64_t t) {  // the run of tile t - 1's last piece goes on in tile t
        const unsigned long long L = last[t - 1], F = first[t];
        return L != NONE && F != NONE && (L >> 32) == (F >> 32);
    };
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; t < n_tiles;
         t += (uint64_t)gridDim.x * blockDim.x) {
        if (!joins(t)) continue;
        // not the first edge of its chain if tile t - 1 is one piece joined from t - 2
        if (t >= 2 && (first[t - 1] & SINGLE) && first[t - 1] != NONE && joins(t - 1)) continue;
        uint64_t total = last[t - 1] & 0x7fffffffull;
        uint64_t parts = (total + k - 1 < 255 ? 1 : 5) + (total + k - 1 + 3) / 4;
        uint64_t j = t;
        for (;;) {
            const uint64_t w = first[j] & 0x7fffffffull;
            total += w;
            parts += (w + k - 1 < 255 ? 1 : 5) + (w + k - 1 + 3) / 4;
            if ((first[j] & SINGLE) && last[j] != NONE && j + 1 < n_tiles && joins(j + 1)) {
                j++;
                continue;
            }
            break;
        }
        const uint64_t L = total + k - 1;
        my_corr += (unsigned long long)((long long)((L < 255 ? 1 : 5) + (L + 3) / 4) - (long long)parts);
    }
    __shared__ unsigned long long s_red[4];
    block_sum_atomic(my_corr, corr, s_red);  // one atomic per block (two's complement sum)
}

// ---------------------------------------------------------------------------
// Pass 1 from the run entries of the RUNS scan, sorted by key (fof: by key,
// then dataset). A key's entries are contiguous. It passes when its windows
// sum above x (the saturating count min(x + 1, sum of min(run, 255)) > x,
// src/lib.rs:3477-3478, 4211-4236, 4625-4629; pieces of one run add up to the
// run) or, with fof, when it occurs in more than x datasets (the minimizers
// DatasetMinimizerTable retains, src/lib.rs:4291-4341, 659-664). Counts go to
// ctr[0] (distinct keys) and ctr[1] (passing keys); with a set, the passing
// keys are inserted (key + 1 per slot, linear probing).
__global__ void k_mtab_scan(const unsigned long long *__restrict__ e, uint64_t n, int fof, uint64_t x,
                            unsigned long long *mtab, uint64_t mask, unsigned long long *ctr) {
    __shared__ unsigned long long s_red[4];
    unsigned long long my_d = 0, my_p = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t key = (uint32_t)(e[i] >> 32);
        if (i > 0 && (uint32_t)(e[i - 1] >> 32) == key) continue;
        my_d++;
        uint64_t acc = 0;
        bool pass = false;
        for (uint64_t j = i; j < n && (uint32_t)(e[j] >> 32) == key; j++) {
            if (fof) acc += (j == i || e[j] != e[j - 1]) ? 1u : 0u;  // distinct datasets
            else acc += (uint32_t)e[j];                              // windows
            if (acc > x) {
                pass = true;
                break;
            }
        }
        if (!pass) continue;
        my_p++;
        if (mtab) {
            uint64_t sl = spread_hash_u64(key) & mask;
            const unsigned long long v = (unsigned long long)key + 1;
            while (atomicCAS(&mtab[sl], 0ull, v) != 0ull) sl = (sl + 1) & mask;
        }
    }
    // block sums (256 threads)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_d += __shfl_xor(my_d, o, 64);
        my_p += __shfl_xor(my_p, o, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_red[wid] = my_d;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 4; w++) t += s_red[w];
        if (t) atomicAdd(&ctr[0], t);
    }
    __syncthreads();
    if (lane == 0) s_red[wid] = my_p;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 4; w++) t += s_red[w];
        if (t) atomicAdd(&ctr[1], t);
    }
}

// ---------------------------------------------------------------------------
// Partitioning (src/lib.rs:3483-3560 and 4821-4925 append each super-k-mer to
// the file of its partition): the staged records move into one contiguous
// bucket per internal partition q in HBM, in two counting passes over the
// digits of q (which each record carries).
// - Coarse pass, digit q >> fbits: every block counts the digits of a fixed
//   range of staging slots (k_part_count), the counts are scanned per digit
//   over the blocks (k_part_colscan) and over the digits (k_part_bucketscan),
//   and every block moves its records to its slots (k_part_coarse).
// - Fine pass, one coarse bucket per block (k_part_fine): it counts the fine
//   digits q & (2^fbits - 1), writes the [start, end) of each of its
//   partitions, and moves the records.
// A move goes through an LDS tile of PT_THREADS * ITEMS records ordered by
// digit (local ranks from LDS atomics, starts from a block scan), so the
// records of one digit leave the tile as one run of consecutive stores.
// Padding slots are dropped by the coarse pass. Record order inside a
// partition is free: counting does not depend on it.
constexpr int PT_THREADS = 512;
constexpr int PT_MAX_DIGITS = 2048;

template <class R> struct PartRec;
template <> struct PartRec<ulonglong2> {
    static constexpr int ITEMS = 16;
    static constexpr uint32_t PAD = REC_PAD_Q;
};
template <> struct PartRec<Rec32> {
    static constexpr int ITEMS = 8;
    static constexpr uint32_t PAD = REC_PAD_QW;
};

struct PartArgs {
    const unsigned long long *n_slots;  // staged slots, padding included (device counter)
    uint64_t cap;                       // slots of each record buffer
    uint32_t p_int;
    int fbits;                          // fine digit bits
    uint32_t nc, nf;                    // coarse buckets, fine digits
    uint32_t nblk;                      // blocks of the coarse pass
    uint32_t *hist;                     // [nc][nblk] coarse digit counts, then their prefix over blocks
    unsigned long long *cbase;          // [nc + 1] bucket starts, then the total
    unsigned long long *pstart, *pend;  // [p_int] partition bounds
    unsigned long long *err;
    // the partition word of each record, 4 bytes apart: of the staging slots
    // (read by the coarse count) and of the coarse pass's output (written by
    // the coarse pass, read by the fine count), so that counting reads 4
    // bytes per record instead of the record's line
    const uint32_t *stage_q;
    uint32_t *coarse_q;
};

// The tile-aligned range of staging slots of coarse-pass block blockIdx.x.
__device__ __forceinline__ void pt_range(const PartArgs &a, uint64_t tile_n, uint64_t &b, uint64_t &e) {
    const uint64_t n = min((uint64_t)*a.n_slots, a.cap);
    const uint64_t tiles = (n + tile_n - 1) / tile_n;
    const uint64_t per = (tiles + a.nblk - 1) / a.nblk * tile_n;
    b = min(n, (uint64_t)blockIdx.x * per);
    e = min(n, b + per);
}

// Exclusive scan of s_cnt[0, nd) into s_start (nd <= PT_MAX_DIGITS); returns
// the total. Each thread sums PT_MAX_DIGITS / PT_THREADS consecutive digits.
__device__ __forceinline__ uint32_t pt_digit_scan(const uint32_t *s_cnt, uint32_t *s_start, uint32_t nd,
                                                  uint32_t *s_wsum) {
    constexpr int PER = PT_MAX_DIGITS / PT_THREADS;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint32_t d = (uint32_t)(tid * PER + i);
        v[i] = d < nd ? s_cnt[d] : 0u;
        sum += v[i];
    }
    const uint32_t incl = (uint32_t)wave_incl_add((int)sum);
    if (lane == 63) s_wsum[wid] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < PT_THREADS / 64; w++) {
        const uint32_t x = s_wsum[w];
        pre += w < wid ? x : 0u;
        tot += x;
    }
    uint32_t run = pre + incl - sum;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint32_t d = (uint32_t)(tid * PER + i);
        if (d < nd) s_start[d] = run;
        run += v[i];
    }
    __syncthreads();
    return tot;
}

// Moves the records of one tile (rec[], slots [tb, te)) to dst: digit
// (q >> shift) & dmask, the next free slot of each digit in s_cur (advanced
// here). Padding records are skipped. s_cnt is zero on entry and on return.
template <class R>
__device__ __forceinline__ void pt_move_tile(const R (&rec)[PartRec<R>::ITEMS], R *__restrict__ dst,
                                             uint32_t *__restrict__ dst_q, uint64_t tb, uint64_t te, int shift,
                                             uint32_t dmask, uint32_t nd, uint64_t cap, R *s_stage, uint32_t *s_cnt,
                                             uint32_t *s_start, uint32_t *s_cur, uint32_t *s_wsum,
                                             unsigned long long *err) {
    constexpr int IT = PartRec<R>::ITEMS;
    const int tid = threadIdx.x;
    int dig[IT];
    uint32_t rank[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const uint64_t i = tb + (uint64_t)it * PT_THREADS + tid;
        dig[it] = -1;
        if (i < te) {
            const uint32_t q = rec_q(rec[it]);
            if (q != PartRec<R>::PAD) dig[it] = (int)((q >> shift) & dmask);
        }
    }
#pragma unroll
    for (int it = 0; it < IT; it++)
        if (dig[it] >= 0) rank[it] = atomicAdd(&s_cnt[dig[it]], 1u);
    __syncthreads();
    const uint32_t total = pt_digit_scan(s_cnt, s_start, nd, s_wsum);
#pragma unroll
    for (int it = 0; it < IT; it++)
        if (dig[it] >= 0) s_stage[s_start[dig[it]] + rank[it]] = rec[it];
    __syncthreads();
    // The thread's slots j = it * PT_THREADS + tid of the ordered tile, in
    // fixed phases (all record reads, then all digit offsets, then the
    // stores) so that each phase's LDS reads are in flight together rather
    // than one dependent chain per record.
    R out[IT];
    uint32_t oq[IT];
    uint64_t pos[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const uint32_t j = (uint32_t)(it * PT_THREADS + tid);
        if (j < total) out[it] = s_stage[j];
    }
    bool over = false;
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const uint32_t j = (uint32_t)(it * PT_THREADS + tid);
        pos[it] = ~0ull;
        if (j < total) {
            oq[it] = rec_q(out[it]);
            const uint32_t d = (oq[it] >> shift) & dmask;
            pos[it] = (uint64_t)s_cur[d] + (j - s_start[d]);
            over |= pos[it] >= cap;
        }
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        if (pos[it] < cap) {
            dst[pos[it]] = out[it];
            if (dst_q) dst_q[pos[it]] = oq[it];
        }
    }
    if (over) atomicOr(err, 16ull);
    __syncthreads();
    for (uint32_t d = tid; d < nd; d += PT_THREADS) {
        s_cur[d] += s_cnt[d];
        s_cnt[d] = 0;
    }
    __syncthreads();
}

// Moves slots [b, e) tile by tile; the next tile's records are loaded into
// registers before the current tile is ranked and stored, so the loads are in
// flight during the LDS phases (block barriers wait on LDS only).
template <class R>
__device__ __forceinline__ void pt_move_range(const R *__restrict__ src, R *__restrict__ dst,
                                              uint32_t *__restrict__ dst_q, uint64_t b, uint64_t e, int shift,
                                              uint32_t dmask, uint32_t nd, uint64_t cap, R *s_stage, uint32_t *s_cnt,
                                              uint32_t *s_start, uint32_t *s_cur, uint32_t *s_wsum,
                                              unsigned long long *err) {
    constexpr int IT = PartRec<R>::ITEMS;
    constexpr uint64_t TN = (uint64_t)PT_THREADS * IT;
    if (b >= e) return;
    // slots past e load slot e - 1 (pt_move_tile ignores them): unconditional
    // loads keep the prefetch's wait counts exact
    R cur[IT], nxt[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) cur[it] = src[min(b + (uint64_t)it * PT_THREADS + threadIdx.x, e - 1)];
    for (uint64_t t = b; t < e; t += TN) {
        const uint64_t tn = t + TN;
#pragma unroll
        for (int it = 0; it < IT; it++) nxt[it] = src[min(tn + (uint64_t)it * PT_THREADS + threadIdx.x, e - 1)];
        pt_move_tile<R>(cur, dst, dst_q, t, min(e, tn), shift, dmask, nd, cap, s_stage, s_cnt, s_start, s_cur,
                        s_wsum, err);
#pragma unroll
        for (int it = 0; it < IT; it++) cur[it] = nxt[it];
    }
}

// Digit counts of slots [b, e) into s_h (LDS atomics), padding skipped, from
// the records' partition words qarr; the next tile's words are loaded before
// the current tile is counted (unconditionally, slots past e reading slot
// e - 1 and counted as padding, so that the loads stay in flight).
template <class R>
__device__ __forceinline__ void pt_count_range(const uint32_t *__restrict__ qarr, uint64_t b, uint64_t e, int shift,
                                               uint32_t dmask, uint32_t *s_h) {
    constexpr int IT = PartRec<R>::ITEMS;
    constexpr uint64_t TN = (uint64_t)PT_THREADS * IT;
    if (b >= e) return;
    uint32_t cur[IT], nxt[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) cur[it] = qarr[min(b + (uint64_t)it * PT_THREADS + threadIdx.x, e - 1)];
    for (uint64_t t = b; t < e; t += TN) {
        const uint64_t tn = t + TN;
#pragma unroll
        for (int it = 0; it < IT; it++) nxt[it] = qarr[min(tn + (uint64_t)it * PT_THREADS + threadIdx.x, e - 1)];
#pragma unroll
        for (int it = 0; it < IT; it++)
            if (t + (uint64_t)it * PT_THREADS + threadIdx.x < e && cur[it] != PartRec<R>::PAD)
                atomicAdd(&s_h[(cur[it] >> shift) & dmask], 1u);
#pragma unroll
        for (int it = 0; it < IT; it++) cur[it] = nxt[it];
    }
}

// Coarse digit counts per coarse-pass block range: PT_COUNT_SPLIT blocks share
// each range (so the count runs at several blocks per CU) and add their
// counts to the zeroed hist.
constexpr int PT_COUNT_SPLIT = 4;
template <class R>
__global__ __launch_bounds__(PT_THREADS) void k_part_count(PartArgs a) {
    __shared__ uint32_t s_h[PT_MAX_DIGITS];
    constexpr uint64_t TN = (uint64_t)PT_THREADS * PartRec<R>::ITEMS;
    for (uint32_t d = threadIdx.x; d < a.nc; d += PT_THREADS) s_h[d] = 0;
    __syncthreads();
    const uint32_t g = blockIdx.x / PT_COUNT_SPLIT, part = blockIdx.x % PT_COUNT_SPLIT;
    const uint64_t n = min((uint64_t)*a.n_slots, a.cap);
    const uint64_t tiles = (n + TN - 1) / TN;
    const uint64_t per = (tiles + a.nblk - 1) / a.nblk * TN;
    const uint64_t gb = min(n, (uint64_t)g * per), ge = min(n, gb + per);
    const uint64_t sub = (per / TN + PT_COUNT_SPLIT - 1) / PT_COUNT_SPLIT * TN;
    const uint64_t b = min(ge, gb + part * sub), e = min(ge, b + sub);
    pt_count_range<R>(a.stage_q, b, e, a.fbits, 0xffffffffu, s_h);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < a.nc; d += PT_THREADS)
        if (s_h[d]) atomicAdd(&a.hist[(uint64_t)d * a.nblk + g], s_h[d]);
}

// Per coarse digit (one block each): exclusive prefix of its counts over the
// blocks, in place; the digit's total into cbase.
__global__ __launch_bounds__(256) void k_part_colscan(PartArgs a) {
    __shared__ uint32_t s_w[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    uint32_t *col = a.hist + (uint64_t)blockIdx.x * a.nblk;
    const uint32_t per = (a.nblk + 255) / 256;
    const uint32_t b = min(a.nblk, (uint32_t)tid * per), e = min(a.nblk, b + per);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; i++) sum += col[i];
    const uint32_t incl = (uint32_t)wave_incl_add((int)sum);
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int w = 0; w < 4; w++) {
        pre += w < wid ? s_w[w] : 0u;
        tot += s_w[w];
    }
    uint32_t run = pre + incl - sum;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t v = col[i];
        col[i] = run;
        run += v;
    }
    if (tid == 0) a.cbase[blockIdx.x] = tot;
}

// Bucket starts: exclusive prefix of the coarse digit totals (one block).
__global__ __launch_bounds__(1024) void k_part_bucketscan(PartArgs a) {
    __shared__ unsigned long long s_w[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t per = (a.nc + 1023) / 1024;
    const uint32_t b = min(a.nc, (uint32_t)tid * per), e = min(a.nc, b + per);
    unsigned long long sum = 0;
    for (uint32_t i = b; i < e; i++) sum += a.cbase[i];
    unsigned long long incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    unsigned long long pre = 0, tot = 0;
    for (int w = 0; w < 16; w++) {
        pre += w < wid ? s_w[w] : 0ull;
        tot += s_w[w];
    }
    __syncthreads();
    unsigned long long run = pre + incl - sum;
    for (uint32_t i = b; i < e; i++) {
        const unsigned long long v = a.cbase[i];
        a.cbase[i] = run;
        run += v;
    }
    if (tid == 0) a.cbase[a.nc] = tot;
}

template <class R>
__global__ __launch_bounds__(PT_THREADS) void k_part_coarse(const R *__restrict__ src, R *__restrict__ dst,
                                                            PartArgs a) {
    __shared__ R s_stage[PT_THREADS * PartRec<R>::ITEMS];
    __shared__ uint32_t s_cnt[PT_MAX_DIGITS], s_start[PT_MAX_DIGITS], s_cur[PT_MAX_DIGITS];
    __shared__ uint32_t s_wsum[PT_THREADS / 64];
    constexpr uint64_t TN = (uint64_t)PT_THREADS * PartRec<R>::ITEMS;
    for (uint32_t d = threadIdx.x; d < a.nc; d += PT_THREADS) {
        s_cnt[d] = 0;
        s_cur[d] = (uint32_t)(a.cbase[d] + a.hist[(uint64_t)d * a.nblk + blockIdx.x]);
    }
    __syncthreads();
    uint64_t b, e;
    pt_range(a, TN, b, e);
    pt_move_range<R>(src, dst, a.coarse_q, b, e, a.fbits, 0xffffffffu, a.nc, a.cap, s_stage, s_cnt, s_start, s_cur,
                     s_wsum, a.err);
}

template <class R>
__global__ __launch_bounds__(PT_THREADS) void k_part_fine(const R *__restrict__ src, R *__restrict__ dst,
                                                          PartArgs a) {
    __shared__ R s_stage[PT_THREADS * PartRec<R>::ITEMS];
    __shared__ uint32_t s_cnt[PT_MAX_DIGITS], s_start[PT_MAX_DIGITS], s_cur[PT_MAX_DIGITS];
    __shared__ uint32_t s_wsum[PT_THREADS / 64];
    const uint32_t bkt = blockIdx.x, fmask = a.nf - 1;
    const uint64_t bs = a.cbase[bkt], be = a.cbase[bkt + 1];
    for (uint32_t d = threadIdx.x; d < a.nf; d += PT_THREADS) s_cnt[d] = 0;
    __syncthreads();
    pt_count_range<R>(a.coarse_q, bs, be, 0, fmask, s_cnt);
    __syncthreads();
    pt_digit_scan(s_cnt, s_start, a.nf, s_wsum);
    for (uint32_t f = threadIdx.x; f < a.nf; f += PT_THREADS) {
        const uint32_t q = (bkt << a.fbits) | f;
        const uint64_t st = bs + s_start[f];
        if (q < a.p_int) {
            a.pstart[q] = st;
            a.pend[q] = st + s_cnt[f];
        }
        s_cur[f] = (uint32_t)st;
        s_cnt[f] = 0;
    }
    __syncthreads();
    pt_move_range<R>(src, dst, nullptr, bs, be, 0, fmask, a.nf, a.cap, s_stage, s_cnt, s_start, s_cur, s_wsum,
                     a.err);
}

// ---------------------------------------------------------------------------
struct CountArgs {
    const ulonglong2 *records;
    const Rec32 *records_w;            // k > 32
    unsigned long long *ovf_hi;        // k > 32: high words of spilled k-mers
    unsigned long long *out_hi;        // k > 32: high words of emitted k-mers
    const unsigned long long *base;    // partition start
    const unsigned long long *cursor;  // partition end
    uint32_t p_int;
    int k;
    uint32_t x;
    unsigned long long *ovf_keys;
    uint32_t *ovf_vals;
    uint32_t *ovf_part;
    uint64_t ovf_mask;
    unsigned long long *out_kmer;
    uint32_t *out_count;
    uint32_t *out_part;
    uint64_t out_cap;
    // k-mers whose 32-bit count wrapped: one (lo, hi) entry per wrap, each
    // worth 2^32 occurrences (ctr[7] entries, at most CARRY_CAP)
    unsigned long long *carry_lo, *carry_hi;
    // [1] output slots reserved (padding included) [2] distinct (LDS)
    // [3] distinct (HBM) [4] error bits [6] padding slots [7] carry entries
    unsigned long long *ctr;
};

// Counts are 32-bit in the tables; the reference's counts are u64
// (add_count saturates a u64, src/lib.rs:3003-3010). A k-mer seen 2^32 times
// in one batch wraps its 32-bit count: the wrapping add appends the k-mer to
// the carry list, the emit keeps it whatever its low count, and the host adds
// 2^32 per entry to its count on download. Error bit 32: list full.
constexpr uint32_t CARRY_CAP = 4096;

__device__ __forceinline__ void carry_add(const CountArgs &a, uint64_t lo, uint64_t hi) {
    const unsigned long long i = atomicAdd(&a.ctr[7], 1ull);
    if (i < CARRY_CAP) {
        a.carry_lo[i] = lo;
        a.carry_hi[i] = hi;
    } else {
        atomicOr(&a.ctr[4], 32ull);
    }
}

// Whether (lo, hi) has a carry entry (a linear pass over the short list).
__device__ __forceinline__ bool has_carry(const CountArgs &a, uint64_t lo, uint64_t hi) {
    const unsigned long long n =
        min((unsigned long long)CARRY_CAP, __hip_atomic_load(&a.ctr[7], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    for (unsigned long long i = 0; i < n; i++)
        if (a.carry_lo[i] == lo && a.carry_hi[i] == hi) return true;
    return false;
}

// Adds one occurrence (MODE 0) or ORs a dataset bit (MODE 1); true when a
// count wrapped past 2^32 - 1.
template <int MODE>
__device__ __forceinline__ bool val_update(uint32_t *p, uint32_t v) {
    if constexpr (MODE == 0) return atomicAdd(p, v) > 0xFFFFFFFFu - v;
    atomicOr(p, v);
    return false;
}

// The spill table takes new keys only while at most an eighth of it is used
// (one distinct spilled k-mer per 128 windows); past that every spill fails
// fast with error bit 1 and the host counts again with smaller partitions
// (and a larger table), instead of probing a crowded table. Tables that
// overflow that much are far too full for their partitions: a normal count
// spills a few hundred k-mers.
__device__ __forceinline__ bool spill_table_full(const CountArgs &a) {
    return __hip_atomic_load(&a.ctr[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > (a.ovf_mask + 1) / 8;
}

// (kept out of line: the spill path is rare, and inlined into every probe
// site it cost the count kernel registers and instruction cache)
template <int MODE>
__device__ void overflow_insert(const CountArgs &a, unsigned long long key, uint32_t val,
                                                          uint32_t q) {
    if (spill_table_full(a)) {
        atomicOr(&a.ctr[4], 1ull);
        return;
    }
    uint64_t slot = spread_hash_u64(key ^ 0x9e3779b97f4a7c15ull) & a.ovf_mask;
    for (int p = 0; p < OVF_MAXPROBE; p++) {
        const unsigned long long cur =
            __hip_atomic_load(&a.ovf_keys[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == key) {
            if (val_update<MODE>(&a.ovf_vals[slot], val)) carry_add(a, key, 0);
            return;
        }
        if (cur == EMPTY_KEY) {
            const unsigned long long prev = atomicCAS(&a.ovf_keys[slot], EMPTY_KEY, key);
            if (prev == EMPTY_KEY) {
                a.ovf_part[slot] = q;
                atomicAdd(&a.ctr[3], 1ull);
                val_update<MODE>(&a.ovf_vals[slot], val);
                return;
            }
            if (prev == key) {
                if (val_update<MODE>(&a.ovf_vals[slot], val)) carry_add(a, key, 0);
                return;
            }
        }
        slot = (slot + 1) & a.ovf_mask;
    }
    atomicOr(&a.ctr[4], 1ull);  // spill table full: host retries with more partitions
}

// Wave-aggregated append: returns the output slot of this lane (if pred).
__device__ __forceinline__ unsigned long long wave_append(bool pred, unsigned long long *cursor) {
    const uint64_t mask = __ballot(pred);
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        if (lane == leader) base = atomicAdd(cursor, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader, 64);
    }
    return base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
}

// Stream compaction for kernels of SCAN_THREADS threads: the indexes of
// [0, n) that keep() accepts get consecutive output slots, one atomic
// reservation per SCAN_THREADS * CMP_ITEMS indexes (a wave-level append
// would take one same-address atomic per 64 indexes, which serialize);
// inside a reservation, slots follow the thread, then the round.
constexpr int CMP_ITEMS = 16;
template <class Keep, class Emit>
__device__ __forceinline__ void block_compact(uint64_t n, unsigned long long *cursor, Keep keep, Emit emit) {
    __shared__ int s_w[SCAN_THREADS / 64];
    __shared__ unsigned long long s_base;
    const uint64_t step = (uint64_t)SCAN_THREADS * CMP_ITEMS;
    for (uint64_t b0 = (uint64_t)blockIdx.x * step; b0 < n; b0 += (uint64_t)gridDim.x * step) {
        uint32_t m = 0;
        int cnt = 0;
#pragma unroll
        for (int it = 0; it < CMP_ITEMS; it++) {
            const uint64_t i = b0 + (uint64_t)it * SCAN_THREADS + threadIdx.x;
            if (i < n && keep(i)) {
                m |= 1u << it;
                cnt++;
            }
        }
        int tot = 0;
        const int off = block_exclusive_scan(cnt, &tot, s_w);
        if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
        __syncthreads();
        unsigned long long pos = s_base + (unsigned long long)off;
#pragma unroll
        for (int it = 0; it < CMP_ITEMS; it++)
            if ((m >> it) & 1u) emit(b0 + (uint64_t)it * SCAN_THREADS + threadIdx.x, pos++);
        __syncthreads();  // s_base is rewritten next round
    }
}

// Output slots are reserved in chunks of OUT_CHUNK per block: one device-scope
// atomic per chunk instead of one per partition. The unused rest of an
// abandoned chunk, and of each block's last one, holds padding entries (key
// ~0, partition OUT_PAD) that sort after every k-mer; ctr[6] counts them.
constexpr unsigned long long OUT_CHUNK = 1024;
constexpr uint32_t OUT_PAD = 0xFFFFFFFFu;

template <bool WIDE>
__device__ __forceinline__ void out_pad(const CountArgs &a, unsigned long long b, unsigned long long e) {
    for (unsigned long long i = b + threadIdx.x; i < e && i < a.out_cap; i += blockDim.x) {
        a.out_kmer[i] = EMPTY_KEY;
        if constexpr (WIDE) a.out_hi[i] = EMPTY_KEY;
        a.out_count[i] = 0;
        a.out_part[i] = OUT_PAD;
    }
}

// Thread 0 keeps the block's current chunk [o_cur, o_end) in registers. For
// tot survivors of a partition: *s_obase gets their first slot; the padded
// rest of an abandoned chunk goes to s_pad[0..1] (written by every thread
// after the caller's barrier).
__device__ __forceinline__ void out_reserve(const CountArgs &a, unsigned long long tot, unsigned long long &o_cur,
                                            unsigned long long &o_end, unsigned long long &my_pad,
                                            unsigned long long *s_obase, unsigned long long *s_pad) {
    if (threadIdx.x != 0) return;
    s_pad[0] = s_pad[1] = 0;
    if (!tot) return;
    if (o_cur + tot > o_end) {
        s_pad[0] = o_cur;
        s_pad[1] = o_end;
        my_pad += o_end - o_cur;
        const unsigned long long sz = tot > OUT_CHUNK ? tot : OUT_CHUNK;
        o_cur = atomicAdd(&a.ctr[1], sz);
        o_end = o_cur + sz;
    }
    *s_obase = o_cur;
    o_cur += tot;
}

// The end of a count kernel: the rest of the block's last chunk is padded and
// the padding total added to ctr[6].
template <bool WIDE>
__device__ __forceinline__ void out_finish(const CountArgs &a, unsigned long long o_cur, unsigned long long o_end,
                                           unsigned long long my_pad, unsigned long long *s_pad) {
    __syncthreads();
    if (threadIdx.x == 0) {
        s_pad[0] = o_cur;
        s_pad[1] = o_end;
        my_pad += o_end - o_cur;
        if (my_pad) atomicAdd(&a.ctr[6], my_pad);
    }
    __syncthreads();
    out_pad<WIDE>(a, s_pad[0], s_pad[1]);
}

// LDS table slot of a k-mer: one multiplicative mix of the low half and the
// rotated high half (the spill table uses spread_hash_u64, an independent
// function).
template <int LS = CNT_LOG2_SLOTS>
__device__ __forceinline__ uint32_t lds_slot(uint64_t key) {
    const uint32_t h = ((uint32_t)key ^ __builtin_rotateleft32((uint32_t)(key >> 32), 16)) * 0x9E3779B1u;
    return h >> (32 - LS);
}

// Adds val to key's entry of the LDS table (claiming an empty slot), or to the
// HBM spill table when 32 probes find neither. The home slot is read first:
// lanes that read one slot share the read, and a present key needs no CAS.
// WRAP: the add returns the old count to catch a wrap past 2^32 - 1 (only a
// partition with that many k-mer instances can wrap one).
template <int MODE, int LS, bool WRAP>
__device__ __forceinline__ void lds_insert(const CountArgs &a, unsigned long long *s_keys, uint32_t *s_vals,
                                           unsigned long long key, uint32_t val, uint32_t q, int *s_wrap) {
    constexpr int SLOTS = 1 << LS;
    uint32_t slot = lds_slot<LS>(key);
    auto add = [&](uint32_t sl) {
        if constexpr (MODE == 0 && WRAP) {
            if (atomicAdd(&s_vals[sl], val) > 0xFFFFFFFFu - val) {
                carry_add(a, key, 0);
                *s_wrap = 1;
            }
        } else if constexpr (MODE == 0) {
            __hip_atomic_fetch_add(&s_vals[sl], val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            __hip_atomic_fetch_or(&s_vals[sl], val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    };
    {
        const unsigned long long cur = __hip_atomic_load(&s_keys[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        unsigned long long prev = cur;
        if (cur == EMPTY_KEY) prev = atomicCAS(&s_keys[slot], EMPTY_KEY, key);
        if (prev == key || prev == EMPTY_KEY) {
            add(slot);
            return;
        }
    }
    slot = (slot + 1) & (SLOTS - 1);
    for (int p = 1; p < CNT_MAXPROBE; p++) {
        const unsigned long long prev = atomicCAS(&s_keys[slot], EMPTY_KEY, key);
        if (prev == EMPTY_KEY || prev == key) {
            add(slot);
            return;
        }
        slot = (slot + 1) & (SLOTS - 1);
    }
    overflow_insert<MODE>(a, key, val, q);
}

// Records of a partition are counted in rounds of up to CNT_ROUND records.
// At high coverage most records of a partition are copies of a few
// super-k-mers, so a round first merges equal records (same bases and length)
// in an LDS table: the first copy represents them and carries their number
// (MODE 0) or the OR of their dataset bits (MODE 1). Only representatives are
// expanded into k-mers, sorted by k-mer count (longest first) so that the 64
// records a wave takes have about the same length and the lanes roll them in
// step.
constexpr int CNT_THREADS_K = 512;            // threads of a k <= 32 count block
constexpr int CNT_ROUND = 2 * CNT_THREADS_K;
constexpr int CNT_NBUCKET = 64;               // k-mers per record <= REC_MAX_BASES
constexpr int CNT_LOG2_DSLOTS = 11;           // record-merge table: 2 slots per record
constexpr int CNT_DSLOTS = 1 << CNT_LOG2_DSLOTS;
static_assert(REC_MAX_BASES < CNT_NBUCKET, "one bucket per k-mer count");
static_assert(CNT_DSLOTS >= 2 * CNT_ROUND, "the merge table is at most half full");
constexpr unsigned long long REC_KEY_Y = 0x3FFFFFFFull;  // bases 32..43 and the base count

__device__ __forceinline__ uint32_t rec_merge_slot(uint64_t x, uint64_t ykey) {
    const uint32_t h = ((uint32_t)x * 0x9E3779B1u) ^ ((uint32_t)(x >> 32) * 0x85EBCA77u) ^
                       ((uint32_t)ykey * 0xC2B2AE3Du);
    return h >> (32 - CNT_LOG2_DSLOTS);
}

// Per partition: all k-mers of its super-k-mer records go into an LDS
// open-addressing table (2^LS slots, linear probing, insert-only); a k-mer
// whose probe window is full goes to the HBM spill table instead, so each
// k-mer lives in exactly one of the two. Then count > x (MODE 0: counts,
// src/lib.rs:2417 / 2502) or popcount(mask) > x (MODE 1: dataset presence,
// src/lib.rs:2603) is emitted with one output reservation per partition.
// The reference's pass-1 minimizer filter (two-pass 3462-3523, fof
// 4291-4341) is not needed here: every occurrence of a k-mer K has the same
// minimizer (the scheme is canonical), so that minimizer's window sum
// (dataset set) is >= c(K) (covers K's datasets), and the filter can only
// drop k-mers that fail "> x" anyway; the emitted set and counts are equal.
// Each lane rolls the canonical k-mers of one record as
// add_packed_superkmer_kmer_counts does (src/lib.rs:3589-3614): forward
// = (forward << 2 | b) & mask, reverse complement = rc >> 2 | (3 - b) << 2(k-1),
// key = min of the two; only the first k-mer of a record is built whole.
// The next round's records (of this partition or the block's next one) are
// loaded while the current round is counted.
template <int MODE, int LS = CNT_LOG2_SLOTS>
// At most 80 VGPRs (six waves per SIMD): with 2^11 or 2^10 table slots three
// or four blocks then fit a CU (LDS 50 / 38 KiB per block); 2^12 slots keep
// two blocks (74 KiB).
__global__ __launch_bounds__(CNT_THREADS_K) __attribute__((amdgpu_waves_per_eu(6))) void k_count(CountArgs a) {
    constexpr int SLOTS = 1 << LS;  // LDS table slots (the tuner picks 2^10 .. 2^12)
    constexpr int PER_T = CNT_ROUND / CNT_THREADS_K;
    __shared__ unsigned long long s_keys[SLOTS];
    __shared__ uint32_t s_vals[SLOTS];
    // the round's records, 12 bytes each: bases 0..31, then bases 32..43 and
    // the base count (what merging compares and rolling reads; the partition
    // is the block's, the dataset travels in s_agg)
    __shared__ uint32_t s_raw[CNT_ROUND * 3];
    auto raw_put = [&](int i, const ulonglong2 &r) {
        s_raw[3 * i] = (uint32_t)r.x;
        s_raw[3 * i + 1] = (uint32_t)(r.x >> 32);
        s_raw[3 * i + 2] = (uint32_t)r.y;
    };
    auto raw_get = [&](int i) {
        return make_ulonglong2((uint64_t)s_raw[3 * i] | ((uint64_t)s_raw[3 * i + 1] << 32), (uint64_t)s_raw[3 * i + 2]);
    };
    __shared__ uint32_t s_dtab[CNT_DSLOTS];     // merge table: record index + 1, 0 = empty
    __shared__ uint32_t s_agg[CNT_ROUND];       // per representative: copies, or OR of dataset bits
    __shared__ uint16_t s_order[CNT_ROUND];     // representatives, most k-mers first
    __shared__ uint32_t s_bucket[CNT_NBUCKET];  // representatives per k-mer count (descending)
    __shared__ int s_wsum[CNT_THREADS_K / 64];
    __shared__ unsigned long long s_red[CNT_THREADS_K / 64];
    __shared__ unsigned long long s_obase, s_pad[2];
    __shared__ int s_wrap;  // a count wrapped in the current partition
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int k = a.k;
    unsigned long long o_cur = 0, o_end = 0, my_pad = 0;  // thread 0: the block's output chunk
    const uint64_t kmask = k >= 32 ? ~0ull : ((1ull << (2 * k)) - 1ull);
    const int rshift = 64 - 2 * k;
    const int rcsh = 2 * k - 2;
    if (tid == 0) s_wrap = 0;
    unsigned long long my_distinct = 0;
    // the tables start empty; each round and each emit leave them empty again
    for (int i = tid; i < SLOTS; i += CNT_THREADS_K) {
        s_keys[i] = EMPTY_KEY;
        s_vals[i] = 0;
    }
    for (int i = tid; i < CNT_DSLOTS; i += CNT_THREADS_K) s_dtab[i] = 0;
    for (int i = tid; i < CNT_ROUND; i += CNT_THREADS_K) s_agg[i] = 0;
    if (tid < CNT_NBUCKET) s_bucket[tid] = 0;
    __syncthreads();

    const uint32_t G = gridDim.x;
    uint32_t q = blockIdx.x;
    uint64_t rb = 0, re = 0, rbn = 0, ren = 0;  // this partition's and the next one's records
    if (q < a.p_int) {
        rb = a.base[q];
        re = a.cursor[q];
    }
    if (q + G < a.p_int) {
        rbn = a.base[q + G];
        ren = a.cursor[q + G];
    }
    ulonglong2 rr[PER_T];
    uint64_t lb = 0, le = 0;  // the records held in rr: [lb, le)
    auto load_round = [&](uint64_t b, uint64_t e) {
        lb = b;
        le = e;
#pragma unroll
        for (int i = 0; i < PER_T; i++) {
            const uint64_t ri = b + (uint64_t)(i * CNT_THREADS_K + tid);
            if (ri < e) rr[i] = a.records[ri];
        }
    };
    load_round(rb, min(rb + CNT_ROUND, re));
    for (; q < a.p_int; q += G) {
        // at most REC_MAX_BASES k-mer instances per record: fewer than 2^32
        // in the partition and no count can wrap
        const bool may_wrap = MODE == 0 && (re - rb) * (uint64_t)REC_MAX_BASES >= (1ull << 32);
        for (uint64_t cb = rb; cb < re; cb += CNT_ROUND) {
            const uint64_t ce = min(cb + CNT_ROUND, re);
            if (lb != cb || le != ce) load_round(cb, ce);  // not prefetched (after an empty partition)
            const int nr = (int)(ce - cb);
#pragma unroll
            for (int i = 0; i < PER_T; i++) {
                const int li = i * CNT_THREADS_K + tid;
                if (li < nr) raw_put(li, rr[i]);
            }
            __syncthreads();
            // merge equal records; representatives are bucketed by k-mer count
            int bk[PER_T];
            uint32_t rk[PER_T];
#pragma unroll
            for (int i = 0; i < PER_T; i++) {
                const int li = i * CNT_THREADS_K + tid;
                bk[i] = -1;
                rk[i] = 0;
                if (li < nr) {
                    const uint64_t x = rr[i].x, ykey = rr[i].y & REC_KEY_Y;
                    uint32_t s = rec_merge_slot(x, ykey);
                    int rep = li;
                    for (int p = 0; p < CNT_DSLOTS; p++) {
                        const uint32_t prev = atomicCAS(&s_dtab[s], 0u, (uint32_t)li + 1u);
                        if (prev == 0) break;
                        const ulonglong2 o = raw_get((int)prev - 1);
                        if (o.x == x && (o.y & REC_KEY_Y) == ykey) {
                            rep = (int)prev - 1;
                            break;
                        }
                        s = (s + 1) & (CNT_DSLOTS - 1);
                    }
                    if constexpr (MODE == 0) atomicAdd(&s_agg[rep], 1u);
                    else atomicOr(&s_agg[rep], 1u << (uint32_t)((rr[i].y >> 30) & 31u));
                    if (rep == li) {
                        const int n = max(rec_len(rr[i]) - k + 1, 0);
                        bk[i] = CNT_NBUCKET - 1 - n;
                        rk[i] = atomicAdd(&s_bucket[bk[i]], 1u);
                    }
                }
            }
            // the next round's records load while this one is counted
            {
                uint64_t nb = rbn, ne = min(rbn + CNT_ROUND, ren);
                if (ce < re) {
                    nb = ce;
                    ne = min(ce + CNT_ROUND, re);
                }
                load_round(nb, ne);
            }
            __syncthreads();
            // bucket starts (every wave scans the 64 counts), then the
            // representatives' order
            int nrep;
            {
                const uint32_t c = s_bucket[lane];
                const int incl = wave_incl_add((int)c);
                nrep = __builtin_amdgcn_readlane(incl, 63);
                const int st = incl - (int)c;
#pragma unroll
                for (int i = 0; i < PER_T; i++) {
                    const int at = __shfl(st, bk[i] < 0 ? 0 : bk[i], 64);
                    if (bk[i] >= 0) s_order[at + (int)rk[i]] = (uint16_t)(i * CNT_THREADS_K + tid);
                }
            }
            for (int i = tid; i < CNT_DSLOTS; i += CNT_THREADS_K) s_dtab[i] = 0;
            __syncthreads();
            if (tid < CNT_NBUCKET) s_bucket[tid] = 0;  // next round (ordered by its barriers)
            // group g of CNT_THREADS_K representatives, in runs of 256: lane
            // l of wave w takes the (4 l + w % 4)-th of run w / 4, so that
            // records with equal k-mers at equal steps (copies of one locus
            // with an error or cut, of equal length) spread over four waves
            // while each wave's records stay close in length
            for (int g = 0; g < nrep; g += CNT_THREADS_K) {
                const int li = g + (wid >> 2) * 256 + 4 * lane + (wid & 3);
                ulonglong2 r = make_ulonglong2(0ull, 0ull);
                int n = 0;
                uint32_t val = 0;
                if (li < nrep) {
                    const int idx = s_order[li];
                    r = raw_get(idx);
                    val = s_agg[idx];
                    s_agg[idx] = 0;
                    n = max(rec_len(r) - k + 1, 0);
                }
                // lane 0 holds the wave's longest record
                const int nmax = __builtin_amdgcn_readfirstlane(n);
                const uint64_t x = r.x, y = r.y & 0xFFFFFFull;  // bases 0..31, 32..43
                const uint64_t w0 = x & kmask;
                uint64_t fwd = __builtin_bitreverse64(w0);
                fwd = (((fwd >> 1) & 0x5555555555555555ull) | ((fwd & 0x5555555555555555ull) << 1)) >> rshift;
                uint64_t rc = (~w0) & kmask;
                // bases k, k + 1, ... still to come (those past 43 read as 0)
                uint64_t la = k >= 32 ? y : ((x >> (2 * k)) | (y << (64 - 2 * k)));
                for (int j = 0; j < nmax; j++) {
                    if (j > 0) {
                        // k <= 11: bases k + 32 on are past the first 32 of la
                        if (j == 33) la = y >> (2 * k);
                        const uint32_t b = (uint32_t)la & 3u;
                        la >>= 2;
                        fwd = ((fwd << 2) | b) & kmask;
                        rc = (rc >> 2) | ((uint64_t)(b ^ 3u) << rcsh);
                    }
                    if (j < n) {
                        const unsigned long long key = fwd < rc ? fwd : rc;
                        if (may_wrap) lds_insert<MODE, LS, true>(a, s_keys, s_vals, key, val, q, &s_wrap);
                        else lds_insert<MODE, LS, false>(a, s_keys, s_vals, key, val, q, &s_wrap);
                    }
                }
            }
            __syncthreads();
        }
        // emit: count survivors, one reservation per partition, then write
        constexpr int PER = SLOTS / CNT_THREADS_K;
        uint32_t emask = 0;
        int cnt = 0;
        const bool wrapped = MODE == 0 && s_wrap;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int sl = i * CNT_THREADS_K + tid;
            const unsigned long long key = s_keys[sl];
            if (key == EMPTY_KEY) continue;
            my_distinct++;
            const uint32_t c = (MODE == 0) ? s_vals[sl] : (uint32_t)__popc(s_vals[sl]);
            if (c > a.x || (wrapped && has_carry(a, key, 0))) {
                emask |= 1u << i;
                cnt++;
            }
        }
        const int x = wave_incl_add(cnt);
        if (lane == 63) s_wsum[wid] = x;
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < CNT_THREADS_K / 64; w++) {
            pre += (w < wid) ? s_wsum[w] : 0;
            tot += s_wsum[w];
        }
        out_reserve(a, (unsigned long long)tot, o_cur, o_end, my_pad, &s_obase, s_pad);
        __syncthreads();
        out_pad<false>(a, s_pad[0], s_pad[1]);
        unsigned long long pos = s_obase + (unsigned long long)(pre + x - cnt);
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int sl = i * CNT_THREADS_K + tid;
            if (emask & (1u << i)) {
                if (pos < a.out_cap) {
                    a.out_kmer[pos] = s_keys[sl];
                    a.out_count[pos] = (MODE == 0) ? s_vals[sl] : (uint32_t)__popc(s_vals[sl]);
                    a.out_part[pos] = q;
                }
                pos++;
            }
            s_keys[sl] = EMPTY_KEY;  // each thread resets the slots it emitted from
            s_vals[sl] = 0;
        }
        if (tid == 0) s_wrap = 0;
        __syncthreads();
        // the next partition; the bounds of the one after it load now
        rb = rbn;
        re = ren;
        rbn = ren = 0;
        if ((uint64_t)q + 2ull * G < a.p_int) {
            rbn = a.base[q + 2 * G];
            ren = a.cursor[q + 2 * G];
        }
    }
    out_finish<false>(a, o_cur, o_end, my_pad, s_pad);
    block_sum_atomic(my_distinct, &a.ctr[2], s_red);
}

template <int MODE>
__global__ void k_overflow_emit(CountArgs a) {
    const uint64_t n = a.ovf_mask + 1;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = i0 + threadIdx.x;
        unsigned long long key = EMPTY_KEY;
        uint32_t c = 0;
        if (i < n) {
            key = a.ovf_keys[i];
            if (key != EMPTY_KEY) c = (MODE == 0) ? a.ovf_vals[i] : (uint32_t)__popc(a.ovf_vals[i]);
        }
        const bool emit = key != EMPTY_KEY && (c > a.x || (MODE == 0 && a.ctr[7] && has_carry(a, key, 0)));
        const unsigned long long pos = wave_append(emit, &a.ctr[1]);
        if (emit && pos < a.out_cap) {
            a.out_kmer[pos] = key;
            a.out_count[pos] = c;
            a.out_part[pos] = a.ovf_part[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Counting for 33 <= k <= 63: canonical k-mers are 128-bit (hi < 2^62), the
// u128 k-mers of src/lib.rs:3589-3614. A table slot (LDS or spill) is claimed
// by a CAS of its high word to (hi | KEY_BUSY); the claimer then writes the
// low word and releases the high word. A reader that meets a busy slot looks
// again in the next round; the number of rounds is bounded and reports error
// bit 8 if the bound is ever reached.
constexpr unsigned long long KEY_BUSY = 1ull << 63;
constexpr int SPIN_LIMIT = 1 << 20;

__device__ __forceinline__ uint64_t rev_groups64(uint64_t x) {
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
}

// Canonical MC encoding of the k MC-coded bases in (b:a): a holds bases
// 0..31, b bases 32..k-1 (masked by mb), base j at bits 2j. Forward = 2-bit
// groups reversed over 2k bits; reverse complement = complement in order.
__device__ __forceinline__ void canon128_mc(uint64_t a, uint64_t b, int k, uint64_t mb, uint64_t &clo,
                                            uint64_t &chi) {
    const uint64_t fh = rev_groups64(a), fl = rev_groups64(b);
    const int sh = 128 - 2 * k;  // 2..62
    const uint64_t f_lo = (fl >> sh) | (fh << (64 - sh));
    const uint64_t f_hi = fh >> sh;
    const uint64_t r_lo = ~a, r_hi = (~b) & mb;
    const bool fwd = f_hi < r_hi || (f_hi == r_hi && f_lo < r_lo);
    clo = fwd ? f_lo : r_lo;
    chi = fwd ? f_hi : r_hi;
}

// One insertion round per loop iteration for every active lane; a lane that
// claims a slot releases it in the same round, so a lane that saw the claim
// (busy) finds the key published when it looks again next round. No lane
// leaves the loop early (the exit is a wave-wide vote), which keeps the
// claimer's release ahead of the waiters' next look.
__device__ void overflow_insert_w(const CountArgs &a, uint64_t klo, uint64_t khi, uint32_t q) {
    uint64_t slot = spread_hash_u64(klo ^ spread_hash_u64(khi + 0x9e3779b97f4a7c15ull)) & a.ovf_mask;
    bool done = spill_table_full(a);
    if (done) atomicOr(&a.ctr[4], 1ull);
    int probes = 0, waits = 0;
    while (true) {
        if (!done) {
            unsigned long long h =
                __hip_atomic_load(&a.ovf_hi[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            bool claimed = false;
            if (h == EMPTY_KEY) {
                const unsigned long long prev = atomicCAS(&a.ovf_hi[slot], EMPTY_KEY, khi | KEY_BUSY);
                if (prev == EMPTY_KEY) claimed = true;
                else h = prev;
            }
            if (claimed) {
                a.ovf_part[slot] = q;
                __hip_atomic_store(&a.ovf_keys[slot], klo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&a.ovf_hi[slot], khi, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                atomicAdd(&a.ctr[3], 1ull);
                atomicAdd(&a.ovf_vals[slot], 1u);
                done = true;
            } else if (h & KEY_BUSY) {
                if (++waits > SPIN_LIMIT) {
                    atomicOr(&a.ctr[4], 8ull);
                    done = true;
                }
            } else if (h == khi &&
                       __hip_atomic_load(&a.ovf_keys[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == klo) {
                if (atomicAdd(&a.ovf_vals[slot], 1u) == 0xFFFFFFFFu) carry_add(a, klo, khi);
                done = true;
            } else {
                slot = (slot + 1) & a.ovf_mask;
                if (++probes >= OVF_MAXPROBE) {
                    atomicOr(&a.ctr[4], 1ull);  // spill table full: host retries with more partitions
                    done = true;
                }
            }
        }
        if (__all(done)) break;
    }
}

// Same structure as k_count<0> (counts, emit count > x), over 256-bit
// records and 128-bit keys.
__global__ __launch_bounds__(CNT_THREADS) void k_count_wide(CountArgs a) {
    __shared__ unsigned long long s_klo[CNT_SLOTS];
    __shared__ unsigned long long s_khi[CNT_SLOTS];
    __shared__ uint32_t s_vals[CNT_SLOTS];
    __shared__ Rec32 s_rec[CNT_THREADS / 64][64];
    __shared__ uint8_t s_mark[CNT_THREADS / 64][64];
    __shared__ int s_wsum[CNT_THREADS / 64];
    __shared__ unsigned long long s_red[CNT_THREADS / 64];
    __shared__ unsigned long long s_obase, s_pad[2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int k = a.k;
    unsigned long long o_cur = 0, o_end = 0, my_pad = 0;  // thread 0: the block's output chunk
    const uint64_t mb = (k - 32) >= 32 ? ~0ull : ((1ull << (2 * (k - 32))) - 1ull);
    unsigned long long my_distinct = 0;
    __shared__ int s_wrap;  // a count wrapped in the current partition
    for (uint32_t q = blockIdx.x; q < a.p_int; q += gridDim.x) {
        for (int i = tid; i < CNT_SLOTS; i += CNT_THREADS) {
            s_klo[i] = EMPTY_KEY;
            s_khi[i] = EMPTY_KEY;
            s_vals[i] = 0;
        }
        if (tid == 0) s_wrap = 0;
        __syncthreads();
        const uint64_t rb = a.base[q], re = a.cursor[q];
        for (uint64_t wb = rb + (uint64_t)wid * 64; wb < re; wb += CNT_THREADS) {
            const uint64_t ri = wb + lane;
            Rec32 r{{0, 0, 0, 0}};
            int n = 0;
            if (ri < re) {
                r = a.records_w[ri];
                n = rec_len(r) - k + 1;
                if (n < 0) n = 0;
            }
            s_rec[wid][lane] = r;
            const int incl = wave_incl_add(n);
            const int total = __builtin_amdgcn_readlane(incl, 63);
            const int my_excl = incl - n;
            int carry = 0;
            for (int j0 = 0; j0 < total; j0 += 64) {
                const int j = j0 + lane;
                s_mark[wid][lane] = 0;
                if (n > 0 && my_excl >= j0 && my_excl < j0 + 64) s_mark[wid][my_excl - j0] = (uint8_t)(lane + 1);
                __builtin_amdgcn_wave_barrier();
                int mk = wave_incl_max((int)s_mark[wid][lane]);
                mk = max(mk, carry);
                carry = __builtin_amdgcn_readlane(mk, 63);
                __builtin_amdgcn_wave_barrier();
                const int o = (mk - 1) & 63;
                const int excl = __shfl(my_excl, o, 64);
                if (j >= total) continue;
                const int off = j - excl;  // k-mer start inside the record (<= 104 - k)
                const Rec32 ro = s_rec[wid][o];
                const uint64_t W3 = ro.w[3] & 0xffffull;  // bases 96..103
                const int bo = 2 * off, wi = bo >> 6, sh = bo & 63;
                const uint64_t x0 = wi == 0 ? ro.w[0] : (wi == 1 ? ro.w[1] : ro.w[2]);
                const uint64_t x1 = wi == 0 ? ro.w[1] : (wi == 1 ? ro.w[2] : W3);
                const uint64_t x2 = wi == 0 ? ro.w[2] : (wi == 1 ? W3 : 0ull);
                const uint64_t fa = sh ? ((x0 >> sh) | (x1 << (64 - sh))) : x0;
                const uint64_t fb = (sh ? ((x1 >> sh) | (x2 << (64 - sh))) : x1) & mb;
                uint64_t klo, khi;
                canon128_mc(fa, fb, k, mb, klo, khi);
                uint32_t slot = lds_slot(klo ^ (khi * 0x9E3779B97F4A7C15ull));
                bool done = false, spill = false;
                int probes = 0, waits = 0;
                while (true) {  // same round structure as overflow_insert_w
                    if (!done) {
                        unsigned long long h =
                            __hip_atomic_load(&s_khi[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        bool claimed = false;
                        if (h == EMPTY_KEY) {
                            const unsigned long long prev = atomicCAS(&s_khi[slot], EMPTY_KEY, khi | KEY_BUSY);
                            if (prev == EMPTY_KEY) claimed = true;
                            else h = prev;
                        }
                        if (claimed) {
                            __hip_atomic_store(&s_klo[slot], klo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_store(&s_khi[slot], khi, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                            atomicAdd(&s_vals[slot], 1u);
                            done = true;
                        } else if (h & KEY_BUSY) {
                            if (++waits > SPIN_LIMIT) {
                                atomicOr(&a.ctr[4], 8ull);
                                done = true;
                            }
                        } else if (h == khi && __hip_atomic_load(&s_klo[slot], __ATOMIC_ACQUIRE,
                                                                 __HIP_MEMORY_SCOPE_WORKGROUP) == klo) {
                            if (atomicAdd(&s_vals[slot], 1u) == 0xFFFFFFFFu) {
                                carry_add(a, klo, khi);
                                s_wrap = 1;
                            }
                            done = true;
                        } else {
                            slot = (slot + 1) & (CNT_SLOTS - 1);
                            if (++probes >= CNT_MAXPROBE) {
                                spill = true;
                                done = true;
                            }
                        }
                    }
                    if (__all(done)) break;
                }
                if (spill) overflow_insert_w(a, klo, khi, q);
            }
        }
        __syncthreads();
        constexpr int PER = CNT_SLOTS / CNT_THREADS;
        uint32_t emask = 0;
        int cnt = 0;
        const bool wrapped = s_wrap;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int sl = i * CNT_THREADS + tid;
            if (s_khi[sl] == EMPTY_KEY) continue;
            my_distinct++;
            if (s_vals[sl] > a.x || (wrapped && has_carry(a, s_klo[sl], s_khi[sl]))) {
                emask |= 1u << i;
                cnt++;
            }
        }
        const int x = wave_incl_add(cnt);
        if (lane == 63) s_wsum[wid] = x;
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < CNT_THREADS / 64; w++) {
            pre += (w < wid) ? s_wsum[w] : 0;
            tot += s_wsum[w];
        }
        out_reserve(a, (unsigned long long)tot, o_cur, o_end, my_pad, &s_obase, s_pad);
        __syncthreads();
        out_pad<true>(a, s_pad[0], s_pad[1]);
        unsigned long long pos = s_obase + (unsigned long long)(pre + x - cnt);
#pragma unroll
        for (int i = 0; i < PER; i++) {
            if (!(emask & (1u << i))) continue;
            const int sl = i * CNT_THREADS + tid;
            if (pos < a.out_cap) {
                a.out_kmer[pos] = s_klo[sl];
                a.out_hi[pos] = s_khi[sl];
                a.out_count[pos] = s_vals[sl];
                a.out_part[pos] = q;
            }
            pos++;
        }
        __syncthreads();
    }
    out_finish<true>(a, o_cur, o_end, my_pad, s_pad);
    block_sum_atomic(my_distinct, &a.ctr[2], s_red);
}

__global__ void k_overflow_emit_wide(CountArgs a) {
    const uint64_t n = a.ovf_mask + 1;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = i0 + threadIdx.x;
        unsigned long long hi = EMPTY_KEY;
        uint32_t c = 0;
        if (i < n) {
            hi = a.ovf_hi[i];
            if (hi != EMPTY_KEY) c = a.ovf_vals[i];
        }
        const bool emit = hi != EMPTY_KEY && (c > a.x || (a.ctr[7] && has_carry(a, a.ovf_keys[i], hi)));
        const unsigned long long pos = wave_append(emit, &a.ctr[1]);
        if (emit && pos < a.out_cap) {
            a.out_kmer[pos] = a.ovf_keys[i];
            a.out_hi[pos] = hi;
            a.out_count[pos] = c;
            a.out_part[pos] = a.ovf_part[i];
        }
    }
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// 16-byte stores for large clears (n16 = number of 16-byte units)
__global__ void k_fill_zero16(uint4 *p, uint64_t n16) {
    const uint4 z = make_uint4(0, 0, 0, 0);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = z;
}

// KFF u8 order key: (encoded & 255, encoded >> 8) as one 64-bit key (k <= 32).
__global__ void k_kff_keys(const unsigned long long *kmer, const uint32_t *part, uint64_t n,
                           uint32_t p_user, unsigned long long *key2, uint32_t *pu,
                           uint32_t *idx) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long K = kmer[i];
        key2[i] = ((K & 255ull) << 56) | (K >> 8);
        pu[i] = part[i] == OUT_PAD ? p_user : part[i] % p_user;  // padding sorts last
        idx[i] = (uint32_t)i;
    }
}
__global__ void k_iota_u32(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_gather_u64(const unsigned long long *src, const uint32_t *idx, unsigned long long *dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

__global__ void k_gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t *dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}
__global__ void k_gather_kmer(const unsigned long long *ks, const uint32_t *cs, const uint32_t *idx,
                              unsigned long long *kd, uint32_t *cd, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = idx[i];
        kd[i] = ks[j];
        cd[i] = cs[j];
    }
}

// Output records in the C ABI's layout (lo, hi, count as u64), so the host
// receives them with one copy.
__global__ void k_pack_counted(const unsigned long long *__restrict__ lo,
                               const unsigned long long *__restrict__ hi,
                               const uint32_t *__restrict__ cnt, unsigned long long *__restrict__ out,
                               uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        out[3 * i] = lo[i];
        out[3 * i + 1] = hi ? hi[i] : 0ull;
        out[3 * i + 2] = cnt[i];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct McPlan {
    uint32_t k = 0;
    // tile schedule
    DevBuf tile_start, tile_frag, tile_len;
    uint64_t n_tiles = 0;
    std::vector<uint64_t> chunk_tile;  // tile index where each chunk starts
    uint64_t n_windows = 0;
    // the tiles a scan covers: [tile_lo, tile_hi) (the whole batch, except
    // while one input is counted on its own)
    uint64_t tile_lo = 0, tile_hi = ~0ull;
    // state
    // records: the coarse pass's output; stage_rec: the staging slots, then
    // the records in partition order (fine pass); base / cursor: each
    // partition's [start, end); hist / cbase: coarse digit counts per block
    // and bucket starts
    DevBuf base, cursor, records, ctr, ovf_keys, ovf_vals, ovf_part, stage_rec, hist, cbase;
    DevBuf stage_q, coarse_q;  // partition words of the staging slots and of the coarse output
    DevBuf records_w, stage_rec_w, ovf_hi, out_hi;  // k > 32
    uint64_t stage_cap = 0, n_staged = 0, n_records = 0;
    uint32_t part_blocks = 0;  // blocks of the coarse pass
    uint32_t ds_lo = 0;        // fof: first dataset of the group being counted
    // pass 1: run entries, their sorted copy, the set of passing keys; tile
    // edge entries for the exact partition byte count
    DevBuf runs, runs_sorted, run_tmp, mtab, tile_first, tile_last;
    uint64_t run_cap = 0, mtab_mask = 0;
    bool use_mtab = false, want_tiles = false;
    DevBuf out_kmer, out_count, out_part;
    DevBuf carry;  // wrapped counts: CARRY_CAP low words, then CARRY_CAP high words
    // partition-group passes (records beyond one device pass): grp_n passes,
    // the current one grp_i; outputs gather in acc_* between passes
    uint32_t grp_n = 1, grp_i = 0;
    uint64_t grp_sig = 0;
    double out_scale = 1.0;
    DevBuf acc_kmer, acc_hi, acc_count, acc_part;
    DevBuf sort_tmp, key2, pu, idx, idx2, pu2, k2out;
    uint64_t ovf_slots = 0;
    uint32_t p_int = 0;
    uint64_t out_cap = 0;
    uint64_t rec_sig = 0;
    double target_override = 0;
    // tuned launch geometry: LDS table of 2^cnt_ls slots (k <= 32), scan
    // blocks per CU (0: every resident block); tune_sig names the
    // configuration the tuning cache was last consulted for
    int cnt_ls = CNT_LOG2_SLOTS;
    uint32_t scan_bpc = 0;
    uint64_t tune_sig = 0;
    double cache_target = 0;  // instances per partition from the tuning cache
    double effective_target = 0;  // instances per partition of the last count (after spill retries)
    // graph cache
    hipGraphExec_t graph = nullptr;
    uint64_t graph_sig = 0;
    ~McPlan() {
        if (graph) (void)hipGraphExecDestroy(graph);
    }
};

void free_plan(McPlan *p) { delete p; }

// Blocks of `threads` threads of kernel f that one CU holds at once.
static unsigned resident_blocks(const void *f, unsigned threads) {
    int nb = 0;
    MCZI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, (int)threads, 0));
    return (unsigned)std::max(nb, 1);
}

static unsigned grid_for(uint64_t n, unsigned threads, unsigned max_blocks) {
    uint64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (unsigned)b;
}

static void fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n, v);
}
static void fill_u64(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n, v);
}
static void fill_zero_bytes(void *p, uint64_t bytes, hipStream_t s) {
    // bytes is a multiple of 16 for every caller
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    hipLaunchKernelGGL(k_fill_zero16, dim3(grid_for(n16, 256, 16384)), dim3(256), 0, s,
                       (uint4 *)p, n16);
}

static void prepare_batch(Batch &b, const uint64_t *off);

// The dataset of each input when all its fragments share one.
static void set_input_datasets(Batch &b, const uint32_t *dataset) {
    const size_t ni = b.input_frag.size() - 1;
    b.input_ds.assign(ni, dataset ? Batch::NO_DATASET : 0u);
    if (!dataset) return;
    for (size_t i = 0; i < ni; i++) {
        const uint64_t f0 = b.input_frag[i], f1 = b.input_frag[i + 1];
        if (f1 <= f0) continue;
        const uint32_t d = dataset[f0];
        bool same = true;
        for (uint64_t f = f0 + 1; f < f1 && same; f++) same = dataset[f] == d;
        if (same) b.input_ds[i] = d;
    }
}

std::unique_ptr<Batch> make_batch(HostReads &&r) {
    r.finish();
    auto b = std::make_unique<Batch>();
    Device &d = Device::get();
    b->n_bases = r.n_bases();
    b->n_frag = r.n_frag();
    b->n_words = r.words.size();
    b->chunk_frag = r.chunk_frag;
    b->input_frag = r.input_frag;
    if (b->input_frag.back() != b->n_frag) b->input_frag.push_back(b->n_frag);
    set_input_datasets(*b, r.dataset.empty() ? nullptr : r.dataset.data());
    b->n_datasets = r.n_datasets;
    b->words.reset(r.words.size() * 4);
    b->off.reset(r.off.size() * 8);
    d.upload(b->words.as<void>(), r.words.data(), r.words.size() * 4, d.stream());
    d.upload(b->off.as<void>(), r.off.data(), r.off.size() * 8, d.stream());
    if (!r.dataset.empty()) {
        b->has_dataset = true;
        b->dataset.reset(r.dataset.size() * 4);
        MCZI_HIP(hipMemcpyAsync(b->dataset.as<void>(), r.dataset.data(), r.dataset.size() * 4,
                                hipMemcpyHostToDevice, d.stream()));
    }
    MCZI_HIP(hipStreamSynchronize(d.stream()));
    prepare_batch(*b, r.off.data());
    return b;
}

// Batch from caller-owned packed words (no intermediate host copy of the
// words): uploads through the device's pinned staging, 32 zero words of
// slack after the data.
std::unique_ptr<Batch> make_batch_packed(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                                         const uint32_t *dataset, uint32_t n_datasets,
                                         std::vector<uint64_t> &&chunk_frag, std::vector<uint64_t> &&input_frag) {
    auto b = std::make_unique<Batch>();
    Device &d = Device::get();
    const uint64_t nb = frag_off[n_frag];
    const uint64_t nw = (nb + 15) / 16;
    b->n_bases = nb;
    b->n_frag = n_frag;
    b->n_words = nw + 32;
    b->chunk_frag = std::move(chunk_frag);
    if (b->chunk_frag.back() != n_frag) b->chunk_frag.push_back(n_frag);
    b->input_frag = std::move(input_frag);
    if (b->input_frag.empty()) b->input_frag.push_back(0);
    if (b->input_frag.back() != n_frag) b->input_frag.push_back(n_frag);
    set_input_datasets(*b, dataset);
    b->n_datasets = n_datasets;
    b->words.reset(b->n_words * 4);
    b->off.reset((n_frag + 1) * 8);
    fill_u32(b->words.as<uint32_t>() + nw, 32, 0u, d.stream());
    MCZI_HIP(hipStreamSynchronize(d.stream()));
    d.upload(b->words.as<void>(), words, nw * 4, d.stream());
    d.upload(b->off.as<void>(), frag_off, (n_frag + 1) * 8, d.stream());
    if (dataset) {
        b->has_dataset = true;
        b->dataset.reset(std::max<size_t>(n_frag, 1) * 4);
        d.upload(b->dataset.as<void>(), dataset, n_frag * 4, d.stream());
    }
    prepare_batch(*b, frag_off);
    return b;
}

// Tile schedule: tiles of TILE positions, clipped at chunk ends, each with
// the fragment holding its first position, and the fragment length summary
// that gives the window count for any k. Built once when the batch is
// created, on the host worker threads; neither depends on k.
static void build_tiles(Batch &b, McPlan &p, const uint64_t *off) {
    const uint64_t nf = b.n_frag;
    require(nf < (1ull << 32), MCZI_E_INVALID, "too many fragments in one batch");
    // fragment lengths: one histogram per chunk of fragments, merged
    const size_t grain = 1u << 20;
    std::vector<std::array<uint64_t, 66>> part((nf + grain - 1) / grain);
    parallel_ranges(nf, grain, host_workers(nf), [&](size_t lo, size_t hi) {
        std::array<uint64_t, 66> h{};  // [0..63] lengths below 64, [64] count and [65] bases of longer ones
        for (size_t i = lo; i < hi; i++) {
            const uint64_t n = off[i + 1] - off[i];
            if (n < 64) {
                h[n]++;
            } else {
                h[64]++;
                h[65] += n;
            }
        }
        part[lo / grain] = h;
    });
    std::fill(std::begin(b.len_hist), std::end(b.len_hist), 0ull);
    b.long_cnt = b.long_sum = 0;
    for (const auto &h : part) {
        for (int j = 0; j < 64; j++) b.len_hist[j] += h[j];
        b.long_cnt += h[64];
        b.long_sum += h[65];
    }
    // the same summary per input
    const size_t ni = b.input_frag.size() - 1;
    b.input_hist.assign(ni, std::array<uint64_t, 66>{});
    parallel_ranges(ni, 64, host_workers(nf), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            std::array<uint64_t, 66> &h = b.input_hist[i];
            for (uint64_t f = b.input_frag[i]; f < b.input_frag[i + 1]; f++) {
                const uint64_t n = off[f + 1] - off[f];
                if (n < 64) {
                    h[n]++;
                } else {
                    h[64]++;
                    h[65] += n;
                }
            }
        }
    });
    // tiles: chunk c covers bases [off[chunk_frag[c]], off[chunk_frag[c + 1]]);
    // tile t of a chunk starts at its start + t * TILE, in the last fragment
    // of the chunk whose offset is <= that position
    p.chunk_tile.clear();
    uint64_t n_tiles = 0;
    for (size_t c = 0; c + 1 < b.chunk_frag.size(); c++) {
        p.chunk_tile.push_back(n_tiles);
        const uint64_t cs = off[b.chunk_frag[c]], ce = off[b.chunk_frag[c + 1]];
        n_tiles += (ce - cs + TILE - 1) / TILE;
    }
    p.chunk_tile.push_back(n_tiles);
    std::vector<uint64_t> ts(n_tiles);
    std::vector<uint32_t> tf(n_tiles), tl(n_tiles);
    for (size_t c = 0; c + 1 < b.chunk_frag.size(); c++) {
        const uint64_t f0 = b.chunk_frag[c], fe = b.chunk_frag[c + 1];
        const uint64_t cs = off[f0], ce = off[fe];
        const uint64_t t0 = p.chunk_tile[c], nt_c = p.chunk_tile[c + 1] - t0;
        parallel_ranges(nt_c, 1u << 14, host_workers(nt_c * 16), [&](size_t lo, size_t hi) {
            // fragment of the first tile by binary search, then walked forward
            const uint64_t s0 = cs + lo * TILE;
            uint64_t f = (uint64_t)(std::upper_bound(off + f0, off + fe, s0) - off) - 1;
            for (size_t t = lo; t < hi; t++) {
                const uint64_t s = cs + t * TILE;
                while (f + 1 < fe && off[f + 1] <= s) f++;
                ts[t0 + t] = s;
                tf[t0 + t] = (uint32_t)f;
                tl[t0 + t] = (uint32_t)std::min<uint64_t>(TILE, ce - s);
            }
        });
    }
    p.n_tiles = n_tiles;
    Device &d = Device::get();
    p.tile_start.reset(std::max<size_t>(n_tiles, 1) * 8);
    p.tile_frag.reset(std::max<size_t>(n_tiles, 1) * 4);
    p.tile_len.reset(std::max<size_t>(n_tiles, 1) * 4);
    if (n_tiles) {
        d.upload(p.tile_start.as<void>(), ts.data(), n_tiles * 8, d.stream());
        d.upload(p.tile_frag.as<void>(), tf.data(), n_tiles * 4, d.stream());
        d.upload(p.tile_len.as<void>(), tl.data(), n_tiles * 4, d.stream());
    }
    MCZI_HIP(hipStreamSynchronize(d.stream()));
}

static void prepare_batch(Batch &b, const uint64_t *off) {
    b.plan.reset(new McPlan());
    build_tiles(b, *b.plan, off);
}

// Windows (k-mer starts) of the batch: sum over fragments of len - k + 1.
static uint64_t windows_for(const Batch &b, uint32_t k) {
    require(k >= 1 && k <= 64, MCZI_E_INVALID, "k must be greater than 0 and <= 64");
    uint64_t n = b.long_sum - b.long_cnt * (k - 1);
    for (uint32_t L = k; L < 64; L++) n += b.len_hist[L] * (L - k + 1);
    return n;
}

static uint64_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Autotuned choice: internal partition count = P_user * 2^j so that the
// expected k-mer instances per partition stay near `target` (the LDS table
// holds CNT_SLOTS distinct k-mers); partition q maps to user partition
// q % P_user, so the reference's output order is kept.
static uint32_t choose_partitions(uint64_t n_windows, uint32_t p_user, double target) {
    uint64_t p = p_user;
    while ((double)n_windows / (double)p > target && p * 2 <= MAX_INTERNAL_PARTITIONS) p *= 2;
    return (uint32_t)p;
}

static double tuned_target(const McPlan &p) {
    const char *e = getenv("MCZI_PART_TARGET");
    if (e && *e) return std::max(64.0, atof(e));
    return 2.0 * (double)(1 << p.cnt_ls);
}

struct Timer {
    hipEvent_t e[12];
    int n = 0;
    bool on;
    explicit Timer(bool enabled) : on(enabled) {
        if (on)
            for (auto &x : e) MCZI_HIP(hipEventCreate(&x));
    }
    ~Timer() {
        if (on)
            for (auto &x : e) (void)hipEventDestroy(x);
    }
    void mark(hipStream_t s) {
        if (on && n < 12) MCZI_HIP(hipEventRecord(e[n++], s));
    }
    double ms(int a, int b) {
        if (!on || b >= n) return 0;
        float v = 0;
        MCZI_HIP(hipEventElapsedTime(&v, e[a], e[b]));
        return v;
    }
};

// Kernel arguments of one pipeline run (pointers depend on the plan's buffers).
static ScanArgs scan_args(Batch &b, McPlan &p, const mczi_counter_config &cfg) {
    const int k = (int)cfg.k, m = (int)cfg.minimizer;
    ScanArgs sa{};
    sa.words = b.words.as<uint32_t>();
    sa.off = b.off.as<uint64_t>();
    sa.dataset = b.has_dataset ? b.dataset.as<uint32_t>() : nullptr;
    sa.tile_start = p.tile_start.as<uint64_t>();
    sa.tile_frag = p.tile_frag.as<uint32_t>();
    sa.tile_len = p.tile_len.as<uint32_t>();
    sa.n_frag = b.n_frag;
    sa.k = k;
    sa.m = m;
    sa.w = k - m + 1;
    sa.x = (uint32_t)cfg.threshold;
    sa.sat = (uint32_t)std::min<uint64_t>(cfg.threshold + 1, 255);
    sa.p_int = p.p_int;
    sa.p_mask = (p.p_int & (p.p_int - 1)) == 0 ? p.p_int - 1 : 0u;
    sa.wrec = (uint32_t)((k > 32 ? REC_MAX_BASES_W : REC_MAX_BASES) - k + 1);
    sa.stats = p.ctr.as<unsigned long long>() + 8;
    sa.stage_rec = p.stage_rec.as<ulonglong2>();
    sa.stage_rec_w = p.stage_rec_w.as<Rec32>();
    sa.stage_cap = p.stage_cap;
    sa.mtab = p.use_mtab ? p.mtab.as<unsigned long long>() : nullptr;
    sa.mtab_mask = p.mtab_mask;
    sa.tile_first = p.want_tiles ? p.tile_first.as<unsigned long long>() : nullptr;
    sa.tile_last = p.want_tiles ? p.tile_last.as<unsigned long long>() : nullptr;
    sa.stage_ctr = p.ctr.as<unsigned long long>() + 11;
    sa.err = p.ctr.as<unsigned long long>() + 4;
    sa.stage_q = p.stage_q.as<uint32_t>();
    sa.grp_n = p.grp_n;
    sa.grp_i = p.grp_i;
    return sa;
}

// Launches the scan (one grid per input chunk with the plain launch, one for
// the batch otherwise); returns the number of blocks, or only counts them
// when s_launch is false.
static unsigned scan_blocks_per_cu(const McPlan &p) {
    static const unsigned scan_res = resident_blocks((const void *)k_scan<STAGE, true>, SCAN_THREADS);
    return p.scan_bpc ? std::min<unsigned>(p.scan_bpc, scan_res) : scan_res;
}

template <int MODE = STAGE>
static uint64_t launch_scan(McPlan &p, const ScanArgs &sa, int launch, hipStream_t s, bool s_launch = true) {
    Device &d = Device::get();
    const unsigned per_cu = scan_blocks_per_cu(p);
    const unsigned scan_grid_max = (unsigned)d.cu_count() * per_cu;
    uint64_t blocks = 0;
    auto go = [&](uint64_t tb, uint64_t te) {
        tb = std::max(tb, p.tile_lo);
        te = std::min(te, p.tile_hi);
        if (te <= tb) return;
        ScanArgs ca = sa;
        ca.tile_begin = tb;
        ca.tile_end = te;
        const unsigned g = (unsigned)std::min<uint64_t>(te - tb, scan_grid_max);
        blocks += g;
        if (!s_launch) return;
        const bool gw = ca.w >= WPT, wide = ca.k > 32;
        if (wide && gw) hipLaunchKernelGGL((k_scan<MODE, true, true>), dim3(g), dim3(SCAN_THREADS), 0, s, ca);
        else if (wide) hipLaunchKernelGGL((k_scan<MODE, false, true>), dim3(g), dim3(SCAN_THREADS), 0, s, ca);
        else if (gw && ca.w == 11) hipLaunchKernelGGL((k_scan<MODE, true, false, 11>), dim3(g), dim3(SCAN_THREADS), 0, s, ca);
        else if (gw) hipLaunchKernelGGL((k_scan<MODE, true>), dim3(g), dim3(SCAN_THREADS), 0, s, ca);
        else hipLaunchKernelGGL((k_scan<MODE, false>), dim3(g), dim3(SCAN_THREADS), 0, s, ca);
    };
    if (launch == MCZI_LAUNCH_PLAIN) {
        for (size_t c = 0; c + 1 < p.chunk_tile.size(); c++) go(p.chunk_tile[c], p.chunk_tile[c + 1]);
    } else {
        go(0, p.n_tiles);
    }
    return blocks;
}

// Stage A: clears, scan + record staging. No host synchronization inside.
static void enqueue_stage_a(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode,
                            int launch, hipStream_t s, Timer *tm) {
    ScanArgs sa = scan_args(b, p, cfg);
    if (mode != MCZI_MODE_DATASETS) sa.dataset = nullptr;
    sa.ds_lo = p.ds_lo;
    unsigned long long *ctr = p.ctr.as<unsigned long long>();
    fill_u64(ctr, 16, 0, s);
    if (p.want_tiles) {
        fill_u64(p.tile_first.as<unsigned long long>(), p.n_tiles, ~0ull, s);
        fill_u64(p.tile_last.as<unsigned long long>(), p.n_tiles, ~0ull, s);
    }
    if (tm) tm->mark(s);
    launch_scan(p, sa, launch, s);
    if (p.want_tiles && p.n_tiles > 1)
        hipLaunchKernelGGL(k_tile_join, dim3(grid_for(p.n_tiles, 256, 4096)), dim3(256), 0, s,
                           p.tile_first.as<unsigned long long>(), p.tile_last.as<unsigned long long>(), p.n_tiles,
                           (uint32_t)cfg.k, ctr + 9);
    if (tm) tm->mark(s);
}

static CountArgs count_args(McPlan &p, const mczi_counter_config &cfg) {
    CountArgs ca{};
    ca.records = p.stage_rec.as<ulonglong2>();  // in partition order after the fine pass
    ca.records_w = p.stage_rec_w.as<Rec32>();
    ca.ovf_hi = p.ovf_hi.as<unsigned long long>();
    ca.out_hi = p.out_hi.as<unsigned long long>();
    ca.base = p.base.as<unsigned long long>();
    ca.cursor = p.cursor.as<unsigned long long>();
    ca.p_int = p.p_int;
    ca.k = (int)cfg.k;
    ca.x = (uint32_t)cfg.threshold;
    ca.ovf_keys = p.ovf_keys.as<unsigned long long>();
    ca.ovf_vals = p.ovf_vals.as<uint32_t>();
    ca.ovf_part = p.ovf_part.as<uint32_t>();
    ca.ovf_mask = p.ovf_slots - 1;
    ca.out_kmer = p.out_kmer.as<unsigned long long>();
    ca.out_count = p.out_count.as<uint32_t>();
    ca.out_part = p.out_part.as<uint32_t>();
    ca.out_cap = p.out_cap;
    ca.carry_lo = p.carry.as<unsigned long long>();
    ca.carry_hi = ca.carry_lo + CARRY_CAP;
    ca.ctr = p.ctr.as<unsigned long long>();
    return ca;
}

static unsigned bits_for(uint64_t n) {
    unsigned b = 1;
    while ((1ull << b) < n) b++;
    return b;
}

// Partition-pass geometry: coarse digit q >> fbits, fine digit q & (2^fbits - 1).
struct PartGeom {
    int fbits = 0;
    uint32_t nc = 1, nf = 1;
};
static PartGeom part_geom(uint32_t p_int) {
    PartGeom g;
    const unsigned bits = bits_for(p_int);
    g.fbits = (int)(bits / 2);
    g.nf = 1u << g.fbits;
    g.nc = (uint32_t)(((uint64_t)p_int + g.nf - 1) >> g.fbits);
    return g;
}

static PartArgs part_args(McPlan &p) {
    const PartGeom g = part_geom(p.p_int);
    PartArgs pa{};
    pa.n_slots = p.ctr.as<unsigned long long>() + 11;
    pa.cap = p.stage_cap;
    pa.p_int = p.p_int;
    pa.fbits = g.fbits;
    pa.nc = g.nc;
    pa.nf = g.nf;
    pa.nblk = p.part_blocks;
    pa.hist = p.hist.as<uint32_t>();
    pa.cbase = p.cbase.as<unsigned long long>();
    pa.pstart = p.base.as<unsigned long long>();
    pa.pend = p.cursor.as<unsigned long long>();
    pa.err = p.ctr.as<unsigned long long>() + 4;
    pa.stage_q = p.stage_q.as<uint32_t>();
    pa.coarse_q = p.coarse_q.as<uint32_t>();
    return pa;
}

// Stage B: records into partition buckets (two counting passes), then per
// partition LDS counting with HBM spill and the emit. Every launch is a
// kernel of this file (no library call, no memset), and no size is read back
// by the host, so the stage can be captured in a graph with stage A.
// The count kernel with a table of 2^ls slots (2^10, 2^11 or 2^12).
using CountFn = void (*)(CountArgs);
template <int MODE>
static CountFn count_kernel(int ls) {
    return ls == 10 ? k_count<MODE, 10> : ls == 12 ? k_count<MODE, 12> : k_count<MODE, 11>;
}

// Blocks of the count grid: every resident block, at most one per partition
// (the output padding is at most one chunk per block).
static uint64_t count_grid_blocks(const McPlan &p, bool wide) {
    Device &d = Device::get();
    static const unsigned res[3] = {resident_blocks((const void *)k_count<0, 10>, CNT_THREADS_K),
                                    resident_blocks((const void *)k_count<0, 11>, CNT_THREADS_K),
                                    resident_blocks((const void *)k_count<0, 12>, CNT_THREADS_K)};
    static const unsigned wide_res = resident_blocks((const void *)k_count_wide, CNT_THREADS);
    const unsigned r = wide ? wide_res : res[std::min(std::max(p.cnt_ls, 10), 12) - 10];
    return std::min<uint64_t>(p.p_int, (uint64_t)d.cu_count() * r);
}

static void enqueue_stage_b(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode,
                            int launch, hipStream_t s, Timer *tm) {
    Device &d = Device::get();
    (void)b;
    (void)launch;
    const bool wide = cfg.k > 32;
    fill_u64(p.ovf_keys.as<unsigned long long>(), p.ovf_slots, EMPTY_KEY, s);
    if (wide) fill_u64(p.ovf_hi.as<unsigned long long>(), p.ovf_slots, EMPTY_KEY, s);
    fill_u32(p.ovf_vals.as<uint32_t>(), p.ovf_slots, 0, s);
    if (tm) tm->mark(s);
    const PartArgs pa = part_args(p);
    if (wide) {
        fill_u32(pa.hist, (uint64_t)pa.nc * pa.nblk, 0u, s);
        hipLaunchKernelGGL(k_part_count<Rec32>, dim3(pa.nblk * PT_COUNT_SPLIT), dim3(PT_THREADS), 0, s, pa);
        hipLaunchKernelGGL(k_part_colscan, dim3(pa.nc), dim3(256), 0, s, pa);
        hipLaunchKernelGGL(k_part_bucketscan, dim3(1), dim3(1024), 0, s, pa);
        hipLaunchKernelGGL(k_part_coarse<Rec32>, dim3(pa.nblk), dim3(PT_THREADS), 0, s, p.stage_rec_w.as<Rec32>(),
                           p.records_w.as<Rec32>(), pa);
        hipLaunchKernelGGL(k_part_fine<Rec32>, dim3(pa.nc), dim3(PT_THREADS), 0, s, p.records_w.as<Rec32>(),
                           p.stage_rec_w.as<Rec32>(), pa);
    } else {
        fill_u32(pa.hist, (uint64_t)pa.nc * pa.nblk, 0u, s);
        hipLaunchKernelGGL(k_part_count<ulonglong2>, dim3(pa.nblk * PT_COUNT_SPLIT), dim3(PT_THREADS), 0, s, pa);
        hipLaunchKernelGGL(k_part_colscan, dim3(pa.nc), dim3(256), 0, s, pa);
        hipLaunchKernelGGL(k_part_bucketscan, dim3(1), dim3(1024), 0, s, pa);
        hipLaunchKernelGGL(k_part_coarse<ulonglong2>, dim3(pa.nblk), dim3(PT_THREADS), 0, s,
                           p.stage_rec.as<ulonglong2>(), p.records.as<ulonglong2>(), pa);
        hipLaunchKernelGGL(k_part_fine<ulonglong2>, dim3(pa.nc), dim3(PT_THREADS), 0, s, p.records.as<ulonglong2>(),
                           p.stage_rec.as<ulonglong2>(), pa);
    }
    if (tm) tm->mark(s);
    CountArgs ca = count_args(p, cfg);
    // persistent grid: exactly the blocks that are resident at once
    const unsigned cgrid = (unsigned)count_grid_blocks(p, wide);
    const unsigned ogrid = grid_for(p.ovf_slots, 256, 8192);
    if (wide) {
        static const unsigned wide_res = resident_blocks((const void *)k_count_wide, CNT_THREADS);
        const unsigned wgrid = (unsigned)std::min<uint64_t>(p.p_int, (uint64_t)d.cu_count() * wide_res);
        hipLaunchKernelGGL(k_count_wide, dim3(wgrid), dim3(CNT_THREADS), 0, s, ca);
        if (tm) tm->mark(s);
        hipLaunchKernelGGL(k_overflow_emit_wide, dim3(ogrid), dim3(256), 0, s, ca);
    } else if (mode == MCZI_MODE_DATASETS) {
        hipLaunchKernelGGL(count_kernel<1>(p.cnt_ls), dim3(cgrid), dim3(CNT_THREADS_K), 0, s, ca);
        if (tm) tm->mark(s);
        hipLaunchKernelGGL(k_overflow_emit<1>, dim3(ogrid), dim3(256), 0, s, ca);
    } else {
        hipLaunchKernelGGL(count_kernel<0>(p.cnt_ls), dim3(cgrid), dim3(CNT_THREADS_K), 0, s, ca);
        if (tm) tm->mark(s);
        hipLaunchKernelGGL(k_overflow_emit<0>, dim3(ogrid), dim3(256), 0, s, ca);
    }
    if (tm) tm->mark(s);
}


// Output entries [0, n_slots) without the padding (the unsorted output
// order: no order is kept). 256-thread blocks.
__global__ __launch_bounds__(SCAN_THREADS) void k_compact_out(const unsigned long long *kin,
                                                              const unsigned long long *hin, const uint32_t *cin,
                                                              const uint32_t *pin, uint64_t n,
                                                              unsigned long long *kout, unsigned long long *hout,
                                                              uint32_t *cout, unsigned long long *cursor) {
    block_compact(
        n, cursor, [&](uint64_t i) { return pin[i] != OUT_PAD; },
        [&](uint64_t i, unsigned long long pos) {
            kout[pos] = kin[i];
            if (hin) hout[pos] = hin[i];
            cout[pos] = cin[i];
        });
}

// The n output slots (n_real k-mers and the padding, which sorts last) into
// b.out_* in the requested order; the first n_real entries are the k-mers.
static void sort_output(Batch &b, McPlan &p, uint64_t n, uint64_t n_real, int order, uint32_t p_user, int k,
                        hipStream_t s) {
    if (order == MCZI_ORDER_NONE || n_real < 2) {
        // drop the padding, keep the emit order
        b.out_kmer.ensure(std::max<uint64_t>(n_real, 1) * 8);
        b.out_count.ensure(std::max<uint64_t>(n_real, 1) * 4);
        if (k > 32) b.out_hi.ensure(std::max<uint64_t>(n_real, 1) * 8);
        if (!n) return;
        unsigned long long *cur = p.ctr.as<unsigned long long>() + 14;
        fill_u64(cur, 1, 0, s);
        hipLaunchKernelGGL(k_compact_out, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                           p.out_kmer.as<unsigned long long>(), k > 32 ? p.out_hi.as<unsigned long long>() : nullptr,
                           p.out_count.as<uint32_t>(), p.out_part.as<uint32_t>(), n,
                           b.out_kmer.as<unsigned long long>(), k > 32 ? b.out_hi.as<unsigned long long>() : nullptr,
                           b.out_count.as<uint32_t>(), cur);
        return;
    }
    require(n < (1ull << 32), MCZI_E_INVALID, "too many output k-mers for one sort");
    unsigned long long *kin = p.out_kmer.as<unsigned long long>();
    uint32_t *cin = p.out_count.as<uint32_t>();
    b.out_kmer.ensure(n * 8);
    b.out_count.ensure(n * 4);
    unsigned long long *kout = b.out_kmer.as<unsigned long long>();
    uint32_t *cout = b.out_count.as<uint32_t>();
    size_t tmp = 0;
    if (k > 32) {
        // 128-bit keys in encoded order: LSD over (lo, then hi) carrying the
        // original positions, then a gather of lo and counts
        require(order == MCZI_ORDER_ENCODED, MCZI_E_INVALID, "KFF u8 record order needs k <= 32");
        const unsigned long long *hin = p.out_hi.as<unsigned long long>();
        b.out_hi.ensure(n * 8);
        unsigned long long *hout = b.out_hi.as<unsigned long long>();
        p.idx.ensure(n * 4);
        p.idx2.ensure(n * 4);
        p.key2.ensure(n * 8);
        p.k2out.ensure(n * 8);
        const unsigned g = grid_for(n, 256, 8192);
        hipLaunchKernelGGL(k_iota_u32, dim3(g), dim3(256), 0, s, p.idx.as<uint32_t>(), n);
        size_t t2 = 0;
        MCZI_HIP(rocprim::radix_sort_pairs(nullptr, tmp, kin, p.key2.as<unsigned long long>(), p.idx.as<uint32_t>(),
                                           p.idx2.as<uint32_t>(), (size_t)n, 0, 64, s));
        MCZI_HIP(rocprim::radix_sort_pairs(nullptr, t2, p.k2out.as<unsigned long long>(), hout, p.idx2.as<uint32_t>(),
                                           p.idx.as<uint32_t>(), (size_t)n, 0, (unsigned)(2 * k - 64), s));
        p.sort_tmp.ensure(std::max(tmp, t2) + 16);
        MCZI_HIP(rocprim::radix_sort_pairs(p.sort_tmp.as<void>(), tmp, kin, p.key2.as<unsigned long long>(),
                                           p.idx.as<uint32_t>(), p.idx2.as<uint32_t>(), (size_t)n, 0, 64, s));
        hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, s, hin, p.idx2.as<uint32_t>(),
                           p.k2out.as<unsigned long long>(), n);
        MCZI_HIP(rocprim::radix_sort_pairs(p.sort_tmp.as<void>(), t2, p.k2out.as<unsigned long long>(), hout,
                                           p.idx2.as<uint32_t>(), p.idx.as<uint32_t>(), (size_t)n, 0,
                                           (unsigned)(2 * k - 64), s));
        hipLaunchKernelGGL(k_gather_kmer, dim3(g), dim3(256), 0, s, kin, cin, p.idx.as<uint32_t>(), kout, cout, n);
        return;
    }
    if (order == MCZI_ORDER_ENCODED) {
        // keys use 2k <= 62 bits (k odd); bit 2k is set only in the padding,
        // which therefore sorts after every k-mer in any top-bit pass
        const unsigned end_bit = (unsigned)std::min(64, 2 * k + 1);
        // Radix passes over the top 24 bits only, then the runs of equal
        // 24-bit prefixes (about n / 2^24 k-mers each) finished in place;
        // keys up to 56 bits, or a run too long to finish, take the full sort.
        constexpr unsigned PREFIX_BITS = 24;
        const bool split = end_bit >= 56 && (n >> PREFIX_BITS) < 64;
        const unsigned begin_bit = split ? end_bit - PREFIX_BITS : 0;
        size_t t_full = 0;
        MCZI_HIP(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, cin, cout, (size_t)n, begin_bit, end_bit, s));
        MCZI_HIP(rocprim::radix_sort_pairs(nullptr, t_full, kin, kout, cin, cout, (size_t)n, 0, end_bit, s));
        p.sort_tmp.ensure(std::max(tmp, t_full) + 16);
        MCZI_HIP(rocprim::radix_sort_pairs(p.sort_tmp.as<void>(), tmp, kin, kout, cin, cout, (size_t)n,
                                           begin_bit, end_bit, s));
        if (!split) return;
        unsigned long long *flag = p.ctr.as<unsigned long long>() + 15;
        fill_u64(flag, 1, 0, s);
        hipLaunchKernelGGL(k_finish_runs<true>, dim3(grid_for(n_real, FIN_CHUNK, 1u << 20)), dim3(256), 0, s, kout,
                           cout, n_real, (int)begin_bit, flag);
        unsigned long long h = 0;
        MCZI_HIP(hipMemcpyAsync(&h, flag, 8, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipStreamSynchronize(s));
        if (h) {
            t_full = std::max(tmp, t_full);
            MCZI_HIP(rocprim::radix_sort_pairs(p.sort_tmp.as<void>(), t_full, kin, kout, cin, cout, (size_t)n, 0,
                                               end_bit, s));
        }
        return;
    }
    // KFF u8 order: (partition % P_user, encoded & 255, encoded); two stable
    // LSD sorts: by (encoded & 255, encoded >> 8), then by user partition.
    p.key2.ensure(n * 8);
    p.k2out.ensure(n * 8);
    p.pu.ensure(n * 4);
    p.pu2.ensure(n * 4);
    p.idx.ensure(n * 4);
    p.idx2.ensure(n * 4);
    unsigned g = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_kff_keys, dim3(g), dim3(256), 0, s, kin, p.out_part.as<uint32_t>(), n, p_user,
                       p.key2.as<unsigned long long>(), p.pu.as<uint32_t>(), p.idx.as<uint32_t>());
    MCZI_HIP(rocprim::radix_sort_pairs(nullptr, tmp, p.key2.as<unsigned long long>(),
                                       p.k2out.as<unsigned long long>(), p.idx.as<uint32_t>(),
                                       p.idx2.as<uint32_t>(), (size_t)n, 0, 64, s));
    size_t tmp2 = 0;
    MCZI_HIP(rocprim::radix_sort_pairs(nullptr, tmp2, p.pu2.as<uint32_t>(), p.pu.as<uint32_t>(),
                                       p.idx2.as<uint32_t>(), p.idx.as<uint32_t>(), (size_t)n, 0, 32, s));
    p.sort_tmp.ensure(std::max(tmp, tmp2) + 16);
    MCZI_HIP(rocprim::radix_sort_pairs(p.sort_tmp.as<void>(), tmp, p.key2.as<unsigned long long>(),
                                       p.k2out.as<unsigned long long>(), p.idx.as<uint32_t>(),
                                       p.idx2.as<uint32_t>(), (size_t)n, 0, 64, s));
    hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(256), 0, s, p.pu.as<uint32_t>(), p.idx2.as<uint32_t>(),
                       p.pu2.as<uint32_t>(), n);
    const unsigned pbits = bits_for((uint64_t)p_user + 1);  // the padding sorts as p_user
    MCZI_HIP(rocprim::radix_sort_pairs(p.sort_tmp.as<void>(), tmp2, p.pu2.as<uint32_t>(),
                                       p.pu.as<uint32_t>(), p.idx2.as<uint32_t>(), p.idx.as<uint32_t>(),
                                       (size_t)n, 0, pbits, s));
    hipLaunchKernelGGL(k_gather_kmer, dim3(g), dim3(256), 0, s, kin, cin, p.idx.as<uint32_t>(), kout,
                       cout, n);
}

static uint64_t mix_sig(uint64_t h, uint64_t v) {
    return spread_hash_u64(h ^ (v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)));
}

constexpr uint32_t MAX_GROUPS = 4096;

static uint64_t plan_held_bytes(const McPlan &p) {
    uint64_t t = 0;
    for (const DevBuf *db : {&p.stage_rec, &p.records, &p.stage_rec_w, &p.records_w, &p.out_kmer, &p.out_hi,
                             &p.out_count, &p.out_part, &p.ovf_keys, &p.ovf_vals, &p.ovf_part, &p.ovf_hi})
        t += db->size();
    return t;
}

// Device memory one counting pass may use: MCZI_DEVICE_BUDGET bytes, or 85%
// of what is free, pooled for reuse, or held by this plan's pass buffers.
static uint64_t device_budget(const McPlan &p) {
    if (const char *e = getenv("MCZI_DEVICE_BUDGET")) return std::max<uint64_t>(strtoull(e, nullptr, 10), 1 << 20);
    size_t fr = 0, tot = 0;
    MCZI_HIP(hipMemGetInfo(&fr, &tot));
    return (uint64_t)(0.85 * (double)(fr + Device::get().pooled_bytes() + plan_held_bytes(p)));
}

// One count of the batch on the device (for fof, of the datasets
// [p.ds_lo, p.ds_lo + 32); with partition groups, of group p.grp_i): scan,
// partitioning, counting and emit, as one graph with the batched launch. The
// output slots (padding included) stay in p.out_*; h receives the counters.
// Staging that outgrew its capacity, or a spill table that filled, is redone
// with more room. Returns false, with p.grp_n raised, when the pass would not
// fit in device memory or would stage 2^32 records or more: the caller then
// counts in that many partition-group passes.
static bool count_pass_one(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode, int launch, Timer &tm,
                           bool want_times, unsigned long long *h) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    const uint32_t k = cfg.k;
    const bool wide = k > 32;
    const bool use_graph = launch == MCZI_LAUNCH_BATCHED && getenv("MCZI_NO_GRAPH") == nullptr;
    double target = p.target_override > 0 ? p.target_override
                    : p.cache_target > 0  ? p.cache_target
                                          : tuned_target(p);
    static const unsigned part_res =
        std::min(resident_blocks((const void *)k_part_coarse<ulonglong2>, PT_THREADS),
                 resident_blocks((const void *)k_part_coarse<Rec32>, PT_THREADS));
    const uint32_t G = std::max<uint32_t>(p.grp_n, 1);
    const uint64_t win = std::max<uint64_t>(p.n_windows / G, 1);  // windows of this pass (expected)
    const uint64_t budget = device_budget(p);
    for (int attempt = 0;; attempt++) {
        p.p_int = choose_partitions(win, cfg.partition_count, target);
        const PartGeom g = part_geom(p.p_int);
        require(g.nc <= PT_MAX_DIGITS && g.nf <= PT_MAX_DIGITS, MCZI_E_INVALID,
                "too many internal partitions for the device partitioning");
        p.base.ensure(((size_t)p.p_int + 1) * 8);
        p.cursor.ensure(((size_t)p.p_int + 1) * 8);
        p.ctr.ensure(16 * 8);
        p.carry.ensure(2 * CARRY_CAP * 8);
        p.part_blocks = (uint32_t)d.cu_count() * part_res;
        p.hist.ensure((size_t)g.nc * p.part_blocks * 4);
        p.cbase.ensure(((size_t)g.nc + 1) * 8);
        // a normal count spills well under one distinct k-mer per 10^4
        // windows; the table takes new keys up to an eighth of its slots, and
        // a count that needs more is redone with smaller partitions and a
        // table twice the size (fill and emit scan the whole table each step)
        const uint64_t ovf = next_pow2(std::max<uint64_t>(win / 256, 1 << 20)) << attempt;
        // every emitted k-mer has more than x windows (out_scale > 1 covers a
        // partition group holding more than its share); plus one padded chunk
        // per count block
        // (threshold 0 keeps every distinct k-mer: a partition's survivors
        // may then abandon up to their own number of slots of a chunk, so
        // the room doubles)
        const uint64_t out_cap = (uint64_t)((double)(win / (cfg.threshold + 1)) * p.out_scale *
                                            (cfg.threshold == 0 ? 2.0 : 1.0)) +
                                 16 + count_grid_blocks(p, wide) * OUT_CHUNK;

        // Staging capacity: the expected record count of a random minimizer
        // scheme (density 2 / (w + 1) per window, plus record splits, one
        // start per fragment and one cut per tile) with a 25% margin, never
        // more than one record per window, plus the padding of each scan
        // block's last chunk. The scan counts every slot it reserves; if the
        // total exceeds the capacity, the step runs again with room for twice
        // the records (the padding of an abandoned chunk is less than its
        // records) plus a full chunk per block.
        const uint64_t cfg_sig = mix_sig(mix_sig(mix_sig(k, cfg.minimizer), b.n_bases), G);
        const uint32_t wrec = (uint32_t)((wide ? REC_MAX_BASES_W : REC_MAX_BASES) - k + 1);
        const uint64_t chunk_pad = launch_scan(p, ScanArgs{}, launch, s, false) * STAGE_CHUNK;
        if (p.rec_sig != cfg_sig) {
            const double w = (double)(k - cfg.minimizer + 1);
            const double expect =
                ((double)p.n_windows * (2.0 / (w + 1.0) + 1.0 / wrec) + (double)b.n_frag + (double)p.n_tiles) / G;
            p.stage_cap = std::min<uint64_t>((uint64_t)(1.25 * expect) + 4096, p.n_windows) + chunk_pad + 16;
            if (const char *e = getenv("MCZI_STAGE_CAPACITY"))  // initial capacity in slots
                p.stage_cap = std::max<uint64_t>(strtoull(e, nullptr, 10), 1);
            p.rec_sig = cfg_sig;
        }
        // device memory of this pass; more than the budget (or 2^32 staging
        // slots) splits the count into partition groups
        const uint64_t rec_bytes = (wide ? sizeof(Rec32) : 16) + 4;  // + the partition word
        const uint64_t need = 2 * p.stage_cap * rec_bytes + out_cap * (wide ? 24 : 16) + ovf * (wide ? 24 : 16);
        if (attempt == 0 && (need > budget || p.stage_cap >= (1ull << 32)) && G < MAX_GROUPS) {
            const uint64_t f = std::max<uint64_t>((need + budget - 1) / std::max<uint64_t>(budget, 1),
                                                  (p.stage_cap >> 31) + 1);
            p.grp_n = (uint32_t)std::min<uint64_t>(MAX_GROUPS, (uint64_t)G * next_pow2(std::max<uint64_t>(f, 2)));
            return false;
        }
        p.ovf_slots = ovf;
        p.ovf_keys.ensure(ovf * 8);
        p.ovf_vals.ensure(ovf * 4);
        p.ovf_part.ensure(ovf * 4);
        if (wide) p.ovf_hi.ensure(ovf * 8);
        p.out_cap = out_cap;
        p.out_kmer.ensure(p.out_cap * 8);
        if (wide) p.out_hi.ensure(p.out_cap * 8);
        p.out_count.ensure(p.out_cap * 4);
        p.out_part.ensure(p.out_cap * 4);
        require(p.stage_cap < (1ull << 32), MCZI_E_INVALID,
                "too many super-k-mer records for one pass (limit 2^32)");
        if (wide) {
            p.stage_rec_w.ensure(p.stage_cap * sizeof(Rec32));
            p.records_w.ensure(p.stage_cap * sizeof(Rec32));
        } else {
            p.stage_rec.ensure(p.stage_cap * 16);
            p.records.ensure(p.stage_cap * 16);
        }
        p.stage_q.ensure(p.stage_cap * 4);
        p.coarse_q.ensure(p.stage_cap * 4);

        tm.n = 0;
        tm.mark(s);
        if (use_graph) {
            uint64_t sig = cfg_sig;
            sig = mix_sig(mix_sig(mix_sig(mix_sig(sig, cfg.threshold), (uint64_t)mode), p.p_int), p.ds_lo);
            sig = mix_sig(mix_sig(sig, p.grp_n), p.grp_i);
            sig = mix_sig(mix_sig(sig, (uint64_t)p.cnt_ls), p.scan_bpc);
            for (const DevBuf *db : {&b.words, &b.off, &b.dataset, &p.base, &p.cursor, &p.records, &p.ctr,
                                     &p.ovf_keys, &p.ovf_vals, &p.ovf_part, &p.out_kmer, &p.out_count,
                                     &p.out_part, &p.tile_start, &p.tile_frag, &p.tile_len, &p.stage_rec,
                                     &p.hist, &p.cbase, &p.records_w, &p.stage_rec_w, &p.ovf_hi, &p.out_hi,
                                     &p.stage_q, &p.coarse_q})
                sig = mix_sig(sig, (uint64_t)(uintptr_t)db->as<void>());
            sig = mix_sig(mix_sig(mix_sig(sig, p.ovf_slots), p.stage_cap), p.out_cap);
            sig = mix_sig(mix_sig(sig, p.part_blocks), (uint64_t)launch);
            sig = mix_sig(mix_sig(sig, p.use_mtab ? (uint64_t)(uintptr_t)p.mtab.as<void>() ^ p.mtab_mask : 0ull),
                          p.want_tiles ? (uint64_t)(uintptr_t)p.tile_first.as<void>() : 0ull);
            if (!p.graph || p.graph_sig != sig) {
                if (p.graph) {
                    MCZI_HIP(hipGraphExecDestroy(p.graph));
                    p.graph = nullptr;
                }
                hipGraph_t gr;
                MCZI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                enqueue_stage_a(b, p, cfg, mode, launch, s, nullptr);
                enqueue_stage_b(b, p, cfg, mode, launch, s, nullptr);
                MCZI_HIP(hipStreamEndCapture(s, &gr));
                MCZI_HIP(hipGraphInstantiate(&p.graph, gr, nullptr, nullptr, 0));
                MCZI_HIP(hipGraphDestroy(gr));
                p.graph_sig = sig;
            }
            MCZI_HIP(hipGraphLaunch(p.graph, s));
        } else {
            enqueue_stage_a(b, p, cfg, mode, launch, s, want_times ? &tm : nullptr);
            enqueue_stage_b(b, p, cfg, mode, launch, s, want_times ? &tm : nullptr);
        }
        tm.mark(s);
        MCZI_HIP(hipMemcpyAsync(h, p.ctr.as<void>(), 16 * 8, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipStreamSynchronize(s));
        MCZI_HIP(hipGetLastError());
        if (h[11] > p.stage_cap || (h[4] & 2ull)) {
            // staging outgrew its capacity: everything after the scan ran on
            // the slots that fit, and is redone (in more partition groups if
            // the room it needs is beyond the budget)
            const uint64_t cap2 = 2 * h[12] + chunk_pad + 16;
            const uint64_t need2 = 2 * cap2 * rec_bytes + p.out_cap * (wide ? 24 : 16) + ovf * (wide ? 24 : 16);
            if ((need2 > budget || cap2 >= (1ull << 32)) && G < MAX_GROUPS) {
                const uint64_t f = std::max<uint64_t>((need2 + budget - 1) / std::max<uint64_t>(budget, 1),
                                                      (cap2 >> 31) + 1);
                p.grp_n = (uint32_t)std::min<uint64_t>(MAX_GROUPS, (uint64_t)G * next_pow2(std::max<uint64_t>(f, 2)));
                return false;
            }
            require(cap2 < (1ull << 32), MCZI_E_INVALID, "too many super-k-mer records for one pass (limit 2^32)");
            require(p.stage_cap < cap2, MCZI_E_DEVICE, "super-k-mer staging exceeded its bound");
            p.stage_cap = cap2;
            attempt--;
            continue;
        }
        require((h[4] & 16ull) == 0, MCZI_E_DEVICE, "partitioning moved a record out of its bucket");
        require((h[4] & 8ull) == 0, MCZI_E_DEVICE, "k-mer table insertion did not settle");
        require((h[4] & 32ull) == 0, MCZI_E_NOMEM,
                "more than 4096 k-mer counts passed 2^32 in one batch (count carry list full)");
        p.n_staged = h[11];
        p.n_records = h[12];
        if (h[1] > p.out_cap && (G > 1 || !(h[4] & 1ull))) {
            // this group holds more than its share of the output, or the
            // chunk padding took more room than planned
            require(attempt < 8, MCZI_E_NOMEM, "output slots overflowed repeatedly");
            p.out_scale *= 2;
            continue;
        }
        if ((h[4] & 1ull) || h[1] > p.out_cap) {
            require(attempt < 4, MCZI_E_NOMEM, "k-mer spill table overflowed repeatedly");
            target /= 4;
            continue;
        }
        // a fill that had to shrink is where the next count of this
        // configuration starts (an explicit override stays as it is)
        if (attempt > 0 && p.target_override <= 0) p.cache_target = target;
        p.effective_target = target;
        return true;
    }
}

// Carries of the pass just counted (ctr[7] entries) into b.carries.
static void collect_carries(Batch &b, McPlan &p, uint64_t n) {
    if (!n) return;
    Device &d = Device::get();
    std::vector<unsigned long long> c(2 * CARRY_CAP);
    d.download(c.data(), p.carry.as<void>(), c.size() * 8, d.stream());
    for (uint64_t i = 0; i < n && i < CARRY_CAP; i++) b.carries[{c[i], c[CARRY_CAP + i]}] += 1ull << 32;
}

// Output entries [0, n) without the padding appended at cursor (k-mer, high
// word, count, partition).
__global__ __launch_bounds__(SCAN_THREADS) void k_append_out(const unsigned long long *kin, const unsigned long long *hin, const uint32_t *cin,
                             const uint32_t *pin, uint64_t n, unsigned long long *kout, unsigned long long *hout,
                             uint32_t *cout, uint32_t *pout, unsigned long long *cursor) {
    block_compact(
        n, cursor, [&](uint64_t i) { return pin[i] != OUT_PAD; },
        [&](uint64_t i, unsigned long long pos) {
            kout[pos] = kin[i];
            if (hout) hout[pos] = hin[i];
            cout[pos] = cin[i];
            pout[pos] = pin[i];
        });
}

// A count that does not fit one device pass (src/lib.rs spills the
// super-k-mers of every partition to disk and counts the partitions one by
// one, 2803-2845): the partitions are split into p.grp_n groups by their key
// hash, and each group is one pass over the resident batch that stages only
// its own super-k-mers, so each pass needs about 1 / grp_n of the staging,
// table and output memory. Every k-mer lives in one group (its minimizer
// fixes its group), so the groups' outputs are disjoint; they gather in HBM
// and leave in p.out_* (h[1] entries, no padding) like one pass's output.
// MCZI_GROUP_PASSES sets the group count; otherwise it is 1 unless a pass
// would exceed the device budget (free device memory, or
// MCZI_DEVICE_BUDGET bytes), and it is kept for later counts of the same
// configuration.
static void count_pass(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode, int launch, Timer &tm,
                       bool want_times, unsigned long long *h) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    const bool wide = cfg.k > 32;
    const uint64_t gsig = mix_sig(mix_sig(mix_sig(mix_sig(cfg.k, cfg.minimizer), cfg.threshold), b.n_bases), mode);
    if (p.grp_sig != gsig) {
        p.grp_n = 1;
        if (const char *e = getenv("MCZI_GROUP_PASSES"))
            p.grp_n = (uint32_t)std::min<uint64_t>(MAX_GROUPS, std::max<uint64_t>(strtoull(e, nullptr, 10), 1));
        p.out_scale = p.grp_n > 1 ? 1.5 : 1.0;
        p.grp_sig = gsig;
    }
    b.carries.clear();
    for (;;) {
        const uint32_t G = p.grp_n;
        if (G <= 1) {
            p.grp_i = 0;
            if (!count_pass_one(b, p, cfg, mode, launch, tm, want_times, h)) {
                p.out_scale = 1.5;
                continue;
            }
            collect_carries(b, p, h[7]);
            return;
        }
        unsigned long long hs[16] = {};
        uint64_t used = 0, cap = 0;
        bool again = false;
        for (uint32_t gi = 0; gi < G; gi++) {
            p.grp_i = gi;
            if (!count_pass_one(b, p, cfg, mode, launch, tm, want_times, h)) {
                again = true;
                break;
            }
            collect_carries(b, p, h[7]);
            const uint64_t n_slots = h[1], n_real = h[1] - h[6];
            if (used + n_real > cap) {  // grow, keeping what is there
                const uint64_t nc = std::max<uint64_t>(2 * cap, used + n_real) + 1024;
                DevBuf nk(nc * 8), nc4(nc * 4), np(nc * 4), nh(wide ? nc * 8 : 0);
                if (used) {
                    MCZI_HIP(hipMemcpyAsync(nk.as<void>(), p.acc_kmer.as<void>(), used * 8, hipMemcpyDeviceToDevice, s));
                    MCZI_HIP(hipMemcpyAsync(nc4.as<void>(), p.acc_count.as<void>(), used * 4, hipMemcpyDeviceToDevice, s));
                    MCZI_HIP(hipMemcpyAsync(np.as<void>(), p.acc_part.as<void>(), used * 4, hipMemcpyDeviceToDevice, s));
                    if (wide)
                        MCZI_HIP(hipMemcpyAsync(nh.as<void>(), p.acc_hi.as<void>(), used * 8, hipMemcpyDeviceToDevice, s));
                }
                MCZI_HIP(hipStreamSynchronize(s));
                p.acc_kmer = std::move(nk);
                p.acc_count = std::move(nc4);
                p.acc_part = std::move(np);
                p.acc_hi = std::move(nh);
                cap = nc;
            }
            if (n_slots) {
                unsigned long long *cur = p.ctr.as<unsigned long long>() + 14;
                fill_u64(cur, 1, 0, s);
                hipLaunchKernelGGL(k_append_out, dim3(grid_for(n_slots, 256, 8192)), dim3(256), 0, s,
                                   p.out_kmer.as<unsigned long long>(),
                                   wide ? p.out_hi.as<unsigned long long>() : nullptr, p.out_count.as<uint32_t>(),
                                   p.out_part.as<uint32_t>(), n_slots, p.acc_kmer.as<unsigned long long>() + used,
                                   wide ? p.acc_hi.as<unsigned long long>() + used : nullptr,
                                   p.acc_count.as<uint32_t>() + used, p.acc_part.as<uint32_t>() + used, cur);
                MCZI_HIP(hipGetLastError());
            }
            used += n_real;
            for (int i : {2, 3, 9, 11, 12, 13}) hs[i] += h[i];
            if (gi == 0) hs[8] = h[8];  // every pass scans every piece
        }
        if (again) {
            b.carries.clear();
            continue;
        }
        MCZI_HIP(hipStreamSynchronize(s));
        p.out_kmer.swap(p.acc_kmer);
        p.out_count.swap(p.acc_count);
        p.out_part.swap(p.acc_part);
        if (wide) p.out_hi.swap(p.acc_hi);
        p.out_cap = cap;
        for (int i = 0; i < 16; i++) h[i] = hs[i];
        h[1] = used;
        p.n_records = hs[12];
        p.n_staged = hs[11];
        return;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_gt_compact(const unsigned long long *keys, const uint32_t *vals,
                                                             const unsigned long long *n_ptr, uint32_t x,
                                                             unsigned long long *kout, uint32_t *cout,
                                                             unsigned long long *cursor) {
    const uint64_t n = *n_ptr;
    block_compact(
        n, cursor, [&](uint64_t i) { return vals[i] > x; },
        [&](uint64_t i, unsigned long long pos) {
            kout[pos] = keys[i];
            cout[pos] = vals[i];
        });
}

// Sorts the accumulated (k-mer, value) pairs by k-mer, sums the values of
// each k-mer, and writes the sums > x to the batch's output in encoded
// order; n_distinct = the distinct k-mers. cur: a device counter.
static void merge_kmer_values(Batch &b, DevBuf &acc_k, DevBuf &acc_c, uint64_t used, const mczi_counter_config &cfg,
                              unsigned long long *cur, hipStream_t s, uint64_t &n_out, uint64_t &n_distinct) {
    // sort the pairs by k-mer, sum per k-mer, keep sums > x
    const unsigned end_bit = (unsigned)std::min<uint32_t>(64, 2 * cfg.k);
    DevBuf sk(std::max<uint64_t>(used, 1) * 8), sv(std::max<uint64_t>(used, 1) * 4);
    DevBuf uk(std::max<uint64_t>(used, 1) * 8), uv(std::max<uint64_t>(used, 1) * 4), nu(16);
    size_t t1 = 0, t2 = 0;
    MCZI_HIP(rocprim::radix_sort_pairs(nullptr, t1, acc_k.as<unsigned long long>(), sk.as<unsigned long long>(),
                                       acc_c.as<uint32_t>(), sv.as<uint32_t>(), (size_t)used, 0, end_bit, s));
    MCZI_HIP(rocprim::reduce_by_key(nullptr, t2, sk.as<unsigned long long>(), sv.as<uint32_t>(), (size_t)used,
                                    uk.as<unsigned long long>(), uv.as<uint32_t>(), nu.as<unsigned long long>(),
                                    rocprim::plus<uint32_t>(), rocprim::equal_to<unsigned long long>(), s));
    DevBuf tmp(std::max(t1, t2) + 16);
    MCZI_HIP(rocprim::radix_sort_pairs(tmp.as<void>(), t1, acc_k.as<unsigned long long>(), sk.as<unsigned long long>(),
                                       acc_c.as<uint32_t>(), sv.as<uint32_t>(), (size_t)used, 0, end_bit, s));
    MCZI_HIP(rocprim::reduce_by_key(tmp.as<void>(), t2, sk.as<unsigned long long>(), sv.as<uint32_t>(), (size_t)used,
                                    uk.as<unsigned long long>(), uv.as<uint32_t>(), nu.as<unsigned long long>(),
                                    rocprim::plus<uint32_t>(), rocprim::equal_to<unsigned long long>(), s));
    // the sums > x (appended in any order), then sorted by k-mer
    fill_u64(cur, 1, 0, s);
    hipLaunchKernelGGL(k_gt_compact, dim3(grid_for(std::max<uint64_t>(used, 1), 256, 8192)), dim3(256), 0, s,
                       uk.as<unsigned long long>(), uv.as<uint32_t>(), nu.as<unsigned long long>(),
                       (uint32_t)cfg.threshold, sk.as<unsigned long long>(), sv.as<uint32_t>(), cur);
    unsigned long long r[2] = {0, 0};
    MCZI_HIP(hipMemcpyAsync(&r[0], cur, 8, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipMemcpyAsync(&r[1], nu.as<void>(), 8, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipStreamSynchronize(s));
    n_out = r[0];
    n_distinct = r[1];
    b.out_kmer.ensure(std::max<uint64_t>(n_out, 1) * 8);
    b.out_count.ensure(std::max<uint64_t>(n_out, 1) * 4);
    size_t t3 = 0;
    MCZI_HIP(rocprim::radix_sort_pairs(nullptr, t3, sk.as<unsigned long long>(), b.out_kmer.as<unsigned long long>(),
                                       sv.as<uint32_t>(), b.out_count.as<uint32_t>(), (size_t)n_out, 0, end_bit, s));
    tmp.ensure(t3 + 16);
    MCZI_HIP(rocprim::radix_sort_pairs(tmp.as<void>(), t3, sk.as<unsigned long long>(),
                                       b.out_kmer.as<unsigned long long>(), sv.as<uint32_t>(),
                                       b.out_count.as<uint32_t>(), (size_t)n_out, 0, end_bit, s));
    MCZI_HIP(hipStreamSynchronize(s));
}

// fof with more than 32 datasets (count_dataset_presence_counts_u64,
// src/lib.rs:637-703, keeps u32 presence counts for any number of
// datasets): each group of 32 datasets is counted as its own pass
// (presence = popcount of the group's 32-bit masks, every k-mer present in
// the group kept); the per-group (k-mer, presence) pairs accumulate in HBM,
// are sorted by k-mer and summed per k-mer, and the sums > x are emitted in
// encoded order.
static void count_dataset_groups(Batch &b, McPlan &p, const mczi_counter_config &cfg, int launch, Timer &tm,
                                 unsigned long long *h, uint64_t &n_out, uint64_t &n_distinct) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    mczi_counter_config gcfg = cfg;
    gcfg.threshold = 0;  // keep every k-mer present in the group
    const uint32_t groups = (b.n_datasets + 31) / 32;
    DevBuf acc_k, acc_c;
    uint64_t used = 0, cap = 0;
    unsigned long long *cur = nullptr;
    for (uint32_t gi = 0; gi < groups; gi++) {
        p.ds_lo = gi * 32;
        count_pass(b, p, gcfg, MCZI_MODE_DATASETS, launch, tm, false, h);
        cur = p.ctr.as<unsigned long long>() + 14;  // allocated by the pass
        const uint64_t n_slots = h[1], n_real = h[1] - h[6];
        if (used + n_real > cap) {  // grow, keeping what is there
            const uint64_t nc = std::max<uint64_t>(2 * cap, used + n_real) + 1024;
            DevBuf nk(nc * 8), nv(nc * 4);
            if (used) {
                MCZI_HIP(hipMemcpyAsync(nk.as<void>(), acc_k.as<void>(), used * 8, hipMemcpyDeviceToDevice, s));
                MCZI_HIP(hipMemcpyAsync(nv.as<void>(), acc_c.as<void>(), used * 4, hipMemcpyDeviceToDevice, s));
            }
            acc_k = std::move(nk);
            acc_c = std::move(nv);
            cap = nc;
        }
        fill_u64(cur, 1, used, s);
        if (n_slots)
            hipLaunchKernelGGL(k_compact_out, dim3(grid_for(n_slots, 256, 8192)), dim3(256), 0, s,
                               p.out_kmer.as<unsigned long long>(), nullptr, p.out_count.as<uint32_t>(),
                               p.out_part.as<uint32_t>(), n_slots, acc_k.as<unsigned long long>(), nullptr,
                               acc_c.as<uint32_t>(), cur);
        used += n_real;
        MCZI_HIP(hipStreamSynchronize(s));
    }
    p.ds_lo = 0;
    p.graph_sig = 0;  // the captured step belongs to the last group
    merge_kmer_values(b, acc_k, acc_c, used, cfg, cur, s, n_out, n_distinct);
}

// Windows of the fragments summarised by one input's length histogram.
static uint64_t windows_of(const std::array<uint64_t, 66> &h, uint32_t k) {
    uint64_t n = h[65] - h[64] * (k - 1);
    for (uint32_t L = k; L < 64; L++) n += h[L] * (L - k + 1);
    return n;
}

// The plain launch over several inputs counts each input (file) as a job of
// its own, as one call per file would: only its tiles are scanned, then
// partitioned and counted with every k-mer kept (threshold 0), and its
// (k-mer, count) pairs -- fof, an input of one dataset: (k-mer, 1) --
// accumulate in HBM. The pairs of all inputs are then merged by k-mer
// (sorted, summed: the count over the inputs, or the number of datasets
// holding the k-mer, as the reference's per-dataset passes give it,
// src/lib.rs:637-703, 2227-2247) and the sums > x emitted in encoded order.
// The result equals the batched count's; the batched launch does not merge.
static void count_per_input(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode, Timer &tm,
                            unsigned long long *h, uint64_t &n_out, uint64_t &n_distinct) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    mczi_counter_config icfg = cfg;
    icfg.threshold = 0;
    const uint64_t all_windows = p.n_windows;
    DevBuf acc_k, acc_c;
    uint64_t used = 0, cap = 0;
    unsigned long long *cur = nullptr;
    size_t c = 0;
    uint64_t win_done = 0;
    unsigned long long hs[16] = {};
    const bool trace = getenv("MCZI_PLAIN_TRACE") != nullptr;  // per-input pass and merge times on stderr
    // An input alone repeats its k-mers less than the whole batch, which the
    // tuned table fill was chosen for: its partitions hold at most one k-mer
    // instance per two LDS slots (so the table stays at most half full even
    // when every k-mer is distinct), and the batch's tuned fill is left as
    // it is.
    struct Restore {  // the plan's batch-wide state, also when a pass throws
        McPlan &p;
        const uint64_t windows;
        const double override_, effective;
        ~Restore() {
            p.target_override = override_;
            p.effective_target = effective;
            p.tile_lo = 0;
            p.tile_hi = ~0ull;
            p.n_windows = windows;
            p.rec_sig = 0;
            p.ds_lo = 0;
            p.graph_sig = 0;
        }
    } restore{p, all_windows, p.target_override, p.effective_target};
    {
        const double t = p.target_override > 0 ? p.target_override
                         : p.cache_target > 0  ? p.cache_target
                                               : tuned_target(p);
        p.target_override = std::min(t, 0.5 * (double)(1u << p.cnt_ls));
    }
    for (size_t i = 0; i + 1 < b.input_frag.size(); i++) {
        const uint64_t f0 = b.input_frag[i], f1 = b.input_frag[i + 1];
        // every input starts a chunk: its tiles are those of chunks [c, ce)
        while (c + 1 < b.chunk_frag.size() && b.chunk_frag[c] < f0) c++;
        size_t ce = c;
        while (ce + 1 < b.chunk_frag.size() && b.chunk_frag[ce] < f1) ce++;
        require(b.chunk_frag[c] == f0 && b.chunk_frag[ce] == f1, MCZI_E_DEVICE, "an input does not start a chunk");
        const uint64_t win = windows_of(b.input_hist[i], cfg.k);
        if (win == 0) continue;
        p.tile_lo = p.chunk_tile[c];
        p.tile_hi = p.chunk_tile[ce];
        p.n_windows = win;
        p.rec_sig = 0;  // staging capacity for this input's size
        p.ds_lo = mode == MCZI_MODE_DATASETS ? (b.input_ds[i] & ~31u) : 0u;
        const auto ti = std::chrono::steady_clock::now();
        count_pass(b, p, icfg, mode, MCZI_LAUNCH_PLAIN, tm, false, h);
        if (trace)
            fprintf(stderr, "MC_PLAIN_TRACE input %zu windows %llu pass_ms %.2f out %llu\n", i,
                    (unsigned long long)win,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ti).count(),
                    (unsigned long long)(h[1] - h[6]));
        cur = p.ctr.as<unsigned long long>() + 14;
        const uint64_t n_slots = h[1], n_real = h[1] - h[6];
        win_done += win;
        if (used + n_real > cap) {  // grow, keeping what is there: room for the
            // inputs still to come at the rate of those counted so far
            const double proj = 1.1 * (double)(used + n_real) * (double)all_windows / (double)win_done;
            const uint64_t nc = std::max<uint64_t>(std::max<uint64_t>(2 * cap, used + n_real), (uint64_t)proj) + 1024;
            DevBuf nk(nc * 8), nv(nc * 4);
            if (used) {
                MCZI_HIP(hipMemcpyAsync(nk.as<void>(), acc_k.as<void>(), used * 8, hipMemcpyDeviceToDevice, s));
                MCZI_HIP(hipMemcpyAsync(nv.as<void>(), acc_c.as<void>(), used * 4, hipMemcpyDeviceToDevice, s));
                MCZI_HIP(hipStreamSynchronize(s));
            }
            acc_k = std::move(nk);
            acc_c = std::move(nv);
            cap = nc;
        }
        fill_u64(cur, 1, used, s);
        if (n_slots)
            hipLaunchKernelGGL(k_compact_out, dim3(grid_for(n_slots, 256, 8192)), dim3(256), 0, s,
                               p.out_kmer.as<unsigned long long>(), nullptr, p.out_count.as<uint32_t>(),
                               p.out_part.as<uint32_t>(), n_slots, acc_k.as<unsigned long long>(), nullptr,
                               acc_c.as<uint32_t>(), cur);
        used += n_real;
        for (int j : {8, 11, 12, 13}) hs[j] += h[j];
        c = ce;
    }
    MCZI_HIP(hipStreamSynchronize(s));
    if (!cur) {  // no input has a window
        p.ctr.ensure(16 * 8);
        cur = p.ctr.as<unsigned long long>() + 14;
    }
    const auto tm0 = std::chrono::steady_clock::now();
    merge_kmer_values(b, acc_k, acc_c, used, cfg, cur, s, n_out, n_distinct);
    if (trace)
        fprintf(stderr, "MC_PLAIN_TRACE merge pairs %llu ms %.2f\n", (unsigned long long)used,
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tm0).count());
    for (int j = 0; j < 16; j++) h[j] = hs[j];
}

// Pass 1 on the device (count_input_minimizers, src/lib.rs:1726-1797; fof:
// count_dataset_minimizer_presence, src/lib.rs:1799-1822 with the driver loop
// of 659-665): the run keys of the whole batch with their windows (fof:
// datasets), sorted, then the distinct and passing keys; the passing keys
// become the set that filters staging. unique = distinct keys, retained =
// passing keys.
static void minimizer_pass(Batch &b, McPlan &p, const mczi_counter_config &cfg, bool fof, int launch,
                           uint64_t &unique, uint64_t &retained, double &ms) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    const uint32_t k = cfg.k;
    hipEvent_t e0, e1;
    MCZI_HIP(hipEventCreate(&e0));
    MCZI_HIP(hipEventCreate(&e1));
    MCZI_HIP(hipEventRecord(e0, s));
    p.ctr.ensure(16 * 8);
    unsigned long long *ctr = p.ctr.as<unsigned long long>();
    if (!p.run_cap) {
        const double w = (double)(k - cfg.minimizer + 1);
        p.run_cap = (uint64_t)(1.25 * (double)p.n_windows * 2.0 / (w + 1.0)) + b.n_frag + p.n_tiles + 4096;
    }
    uint64_t n = 0;
    for (;;) {
        p.runs.ensure(p.run_cap * 8);
        ScanArgs sa = scan_args(b, p, cfg);
        sa.dataset = fof ? sa.dataset : nullptr;
        sa.mtab = nullptr;
        sa.tile_first = sa.tile_last = nullptr;
        sa.run_out = p.runs.as<unsigned long long>();
        sa.run_cap = p.run_cap;
        sa.run_ctr = ctr + 10;
        fill_u64(ctr, 16, 0, s);
        launch_scan<RUNS>(p, sa, launch, s);
        unsigned long long h[16];
        MCZI_HIP(hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, s));
        MCZI_HIP(hipStreamSynchronize(s));
        MCZI_HIP(hipGetLastError());
        n = h[10];
        if (n <= p.run_cap) break;
        p.run_cap = n + n / 8 + 4096;
    }
    p.runs_sorted.ensure(std::max<uint64_t>(n, 1) * 8);
    size_t tb = 0;
    // whole entries: by key, then windows (dataset)
    MCZI_HIP(rocprim::radix_sort_keys(nullptr, tb, p.runs.as<unsigned long long>(),
                                      p.runs_sorted.as<unsigned long long>(), (size_t)n, 0u, 64u, s));
    p.run_tmp.ensure(tb + 16);
    MCZI_HIP(rocprim::radix_sort_keys(p.run_tmp.as<void>(), tb, p.runs.as<unsigned long long>(),
                                      p.runs_sorted.as<unsigned long long>(), (size_t)n, 0u, 64u, s));
    const unsigned g = grid_for(n, 256, 8192);
    fill_u64(ctr, 4, 0, s);
    hipLaunchKernelGGL(k_mtab_scan, dim3(g), dim3(256), 0, s, p.runs_sorted.as<unsigned long long>(), n, fof ? 1 : 0,
                       (uint64_t)cfg.threshold, nullptr, 0ull, ctr);
    unsigned long long c[2];
    MCZI_HIP(hipMemcpyAsync(c, ctr, sizeof c, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipStreamSynchronize(s));
    unique = c[0];
    retained = c[1];
    const uint64_t slots = next_pow2(std::max<uint64_t>(2 * retained, 1024));
    p.mtab.ensure(slots * 8);
    p.mtab_mask = slots - 1;
    fill_u64(p.mtab.as<unsigned long long>(), slots, 0ull, s);
    hipLaunchKernelGGL(k_mtab_scan, dim3(g), dim3(256), 0, s, p.runs_sorted.as<unsigned long long>(), n, fof ? 1 : 0,
                       (uint64_t)cfg.threshold, p.mtab.as<unsigned long long>(), p.mtab_mask, ctr + 2);
    MCZI_HIP(hipEventRecord(e1, s));
    MCZI_HIP(hipEventSynchronize(e1));
    float f = 0;
    MCZI_HIP(hipEventElapsedTime(&f, e0, e1));
    ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

static void apply_tuning_cache(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode, int order);

void batch_count(Batch &b, const mczi_counter_config &cfg, int mode, int launch, int order,
                 mczi_run_stats *stats) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    (void)hipGetLastError();  // an earlier failed call's error is not this call's
    const uint32_t k = cfg.k;
    require(k >= 1 && k <= 64, MCZI_E_INVALID, "k must be <= 64 because MC stores k-mers in u128");
    const bool wide = k > 32;  // 128-bit k-mers, 256-bit records
    require(cfg.minimizer >= 1 && cfg.minimizer <= k, MCZI_E_INVALID, "minimizer size must be <= k");
    require(k - cfg.minimizer + 1 <= MAXW, MCZI_E_INVALID, "window k - m + 1 must be <= 64");
    require(k % 2 == 1, MCZI_E_INVALID,
            "k must be odd: the canonical window length l = k must be odd (simd-minimizers)");
    require(cfg.partition_count > 0, MCZI_E_INVALID, "partition count must be greater than 0");
    require(cfg.partition_count <= MAX_INTERNAL_PARTITIONS, MCZI_E_INVALID,
            "partition count above 4194304 is not supported on the device");
    require(mode >= MCZI_MODE_TWO_PASS && mode <= MCZI_MODE_DATASETS_TWO_PASS, MCZI_E_INVALID, "unknown counting mode");
    const bool two_pass = mode == MCZI_MODE_TWO_PASS || mode == MCZI_MODE_DATASETS_TWO_PASS;
    if (mode == MCZI_MODE_DATASETS_TWO_PASS) mode = MCZI_MODE_DATASETS;
    if (mode == MCZI_MODE_DATASETS) require(!wide, MCZI_E_INVALID, "FOFN dataset-presence mode currently requires k <= 32");
    require(!wide || order != MCZI_ORDER_KFF_U8, MCZI_E_INVALID, "KFF u8 record order needs k <= 32");
    McPlan &p = *b.plan;
    p.n_windows = windows_for(b, k);
    p.k = k;
    if (p.target_override <= 0) apply_tuning_cache(b, p, cfg, mode, order);
    if (wide) p.cnt_ls = CNT_LOG2_SLOTS;  // the 128-bit table has one size
    uint64_t unique_min = 0, retained_min = 0;
    double ms_pass1 = 0;
    p.use_mtab = false;
    if (two_pass) {
        minimizer_pass(b, p, cfg, mode == MCZI_MODE_DATASETS, launch, unique_min, retained_min, ms_pass1);
        p.use_mtab = true;
    }
    // exact partition bytes: per-tile edge entries joined on the device (one dataset group)
    p.want_tiles = !(mode == MCZI_MODE_DATASETS && b.n_datasets > 32);
    if (p.want_tiles) {
        p.tile_first.ensure(std::max<uint64_t>(p.n_tiles, 1) * 8);
        p.tile_last.ensure(std::max<uint64_t>(p.n_tiles, 1) * 8);
    }
    const bool use_graph = launch == MCZI_LAUNCH_BATCHED && getenv("MCZI_NO_GRAPH") == nullptr;
    const bool want_times = stats != nullptr && !use_graph;
    Timer tm(stats != nullptr);
    unsigned long long h[16];
    mczi_run_stats st{};
    b.k = k;
    // the plain launch over several inputs: each input counted on its own,
    // then merged (encoded order, k <= 32; a k-mer's count over the inputs
    // fits the 32-bit sums when the batch has fewer than 2^32 windows; fof:
    // every input holds one dataset)
    bool per_input = launch == MCZI_LAUNCH_PLAIN && order == MCZI_ORDER_ENCODED && !wide &&
                     b.input_frag.size() > 2 && (mode == MCZI_MODE_DATASETS || p.n_windows < (1ull << 32));
    if (per_input && mode == MCZI_MODE_DATASETS)
        for (uint32_t dsi : b.input_ds) per_input = per_input && dsi != Batch::NO_DATASET;
    if (per_input || (mode == MCZI_MODE_DATASETS && b.n_datasets > 32)) {
        auto t0 = std::chrono::steady_clock::now();
        uint64_t n_out = 0, n_distinct = 0;
        if (per_input) count_per_input(b, p, cfg, mode, tm, h, n_out, n_distinct);
        else count_dataset_groups(b, p, cfg, launch, tm, h, n_out, n_distinct);
        b.carries.clear();  // the merged sums are whole counts
        b.n_out = n_out;
        st.n_bases = b.n_bases;
        st.n_windows = p.n_windows;
        st.n_distinct = n_distinct;
        st.n_out = n_out;
        st.internal_partitions = p.p_int;
        st.lds_slots = 1u << p.cnt_ls;
        st.tile_windows = TILE;
        st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st.unique_minimizers = unique_min;
        st.minimizers_retained = retained_min;
        st.ms_pass1 = ms_pass1;
        st.partition_groups = p.grp_n;
        st.scan_blocks_per_cu = scan_blocks_per_cu(p);
        b.last = st;
        if (stats) *stats = st;
        return;
    }
    p.ds_lo = 0;
    count_pass(b, p, cfg, mode, launch, tm, want_times, h);
    const int t_dev_end = tm.n - 1;
    const uint64_t n_slots = h[1], n_real = h[1] - h[6];
    const int tsort0 = tm.n;
    tm.mark(s);
    sort_output(b, p, n_slots, n_real, order, cfg.partition_count, (int)k, s);
    tm.mark(s);
    MCZI_HIP(hipStreamSynchronize(s));
    MCZI_HIP(hipGetLastError());
    b.n_out = n_real;
    st.n_bases = b.n_bases;
    st.n_windows = p.n_windows;
    st.n_pieces = h[8];
    st.n_records = p.n_records;
    st.tile_windows = TILE;
    st.n_distinct = h[2] + h[3];
    st.n_overflow = h[3];
    st.n_out = n_real;
    st.partition_bytes = h[13] + h[9];  // piece bytes + the join correction (two's complement)
    st.unique_minimizers = unique_min;
    st.minimizers_retained = retained_min;
    st.ms_pass1 = ms_pass1;
    st.internal_partitions = p.p_int;
    st.lds_slots = wide ? CNT_SLOTS : 1u << p.cnt_ls;
    st.partition_groups = p.grp_n;
    st.scan_blocks_per_cu = scan_blocks_per_cu(p);
    if (want_times && t_dev_end >= 7) {
        // marks: 0 start, 1 clears, 2 scan + staging, 3 clears,
        // 4 partitioning, 5 count, 6 spill emit
        st.ms_scan1 = tm.ms(1, 2);
        st.ms_scan2 = tm.ms(3, 4);
        st.ms_count = tm.ms(4, 6);
    }
    st.ms_total = tm.ms(0, t_dev_end);
    st.ms_sort = tm.ms(tsort0, tsort0 + 1);
    b.last = st;
    if (stats) *stats = st;
}

uint64_t batch_minimizer_count(Batch &b, const mczi_counter_config &cfg) {
    require(cfg.k % 2 == 1 && cfg.k <= 63, MCZI_E_INVALID,
            "k must be odd: the canonical window length l = k must be odd (simd-minimizers)");
    require(cfg.k - cfg.minimizer + 1 <= MAXW, MCZI_E_INVALID, "window k - m + 1 must be <= 64");
    McPlan &p = *b.plan;
    p.n_windows = windows_for(b, cfg.k);
    uint64_t unique = 0, retained = 0;
    double ms = 0;
    minimizer_pass(b, p, cfg, false, MCZI_LAUNCH_BATCHED, unique, retained, ms);
    return unique;
}

// ---- autotuner and its cache ------------------------------------------------
// Tuned dimensions (none changes a result): the count's LDS table size
// (2^10, 2^11 or 2^12 slots; k <= 32), the k-mer instances per internal
// partition (the table fill, which sets the partition count), and the scan's
// blocks per CU (its persistent grid; each block is four wavefronts over a
// 1920-window tile, so a read gets one wavefront per 480 of its windows).

// Tuning-cache key: device, k, m, mode, order and the batch's shape (log2 of
// its bases and of its mean fragment length, its dataset count class).
static std::string tune_key(const Batch &b, const mczi_counter_config &cfg, int mode, int order) {
    static const std::string dev = [] {
        hipDeviceProp_t pr;
        MCZI_HIP(hipGetDeviceProperties(&pr, Device::get().device()));
        return std::string(pr.gcnArchName) + ":" + std::to_string(pr.multiProcessorCount);
    }();
    auto lg = [](uint64_t v) {
        int r = 0;
        while (v > 1) {
            v >>= 1;
            r++;
        }
        return r;
    };
    const uint64_t mean = b.n_frag ? b.n_bases / b.n_frag : 0;
    const int dsc = b.n_datasets <= 1 ? 1 : b.n_datasets <= 32 ? 32 : 33;
    char buf[160];
    snprintf(buf, sizeof buf, "%s:k%u:m%u:mode%d:order%d:b%d:L%d:d%d", dev.c_str(), cfg.k, cfg.minimizer, mode, order,
             lg(b.n_bases), lg(mean), dsc);
    return buf;
}

// MCZI_TUNE_CACHE, else $HOME/.cache/mczi/autotune.tsv; lines of
// key, table bits, instances per partition, scan blocks per CU, ms.
static std::string tune_path() {
    if (const char *e = getenv("MCZI_TUNE_CACHE")) return e;
    if (const char *h = getenv("HOME")) return std::string(h) + "/.cache/mczi/autotune.tsv";
    return "";
}

static std::mutex g_tune_mu;

static bool tune_lookup(const std::string &key, int &ls, double &target, unsigned &bpc) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string path = tune_path();
    if (path.empty()) return false;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
        char k[256];
        int l = 0;
        double t = 0, ms = 0;
        unsigned c = 0;
        if (sscanf(line, "%255s %d %lf %u %lf", k, &l, &t, &c, &ms) == 5 && key == k && l >= 10 && l <= 12 &&
            t > 0) {
            ls = l;
            target = t;
            bpc = c;
            found = true;  // the last entry for a key wins
        }
    }
    fclose(f);
    return found;
}

static void tune_store(const std::string &key, int ls, double target, unsigned bpc, double ms) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string path = tune_path();
    if (path.empty()) return;
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) {
        // create the directory chain (best effort: the cache is optional)
        for (size_t i = 1; i <= slash; i++)
            if (path[i] == '/' || i == slash) (void)mkdir(path.substr(0, i).c_str(), 0755);
    }
    if (FILE *f = fopen(path.c_str(), "a")) {
        fprintf(f, "%s\t%d\t%.1f\t%u\t%.3f\n", key.c_str(), ls, target, bpc, ms);
        fclose(f);
    }
}

// Applies the cached tuning of this configuration, if any, once per
// configuration (an explicit override or an autotune run wins).
static void apply_tuning_cache(Batch &b, McPlan &p, const mczi_counter_config &cfg, int mode, int order) {
    const std::string key = tune_key(b, cfg, mode, order);
    const uint64_t sig = spread_hash_u64(std::hash<std::string>()(key));
    if (p.tune_sig == sig) return;
    p.tune_sig = sig;
    p.cnt_ls = CNT_LOG2_SLOTS;
    p.scan_bpc = 0;
    p.cache_target = 0;
    int ls;
    double target;
    unsigned bpc;
    if (cfg.k <= 32 && tune_lookup(key, ls, target, bpc)) {
        p.cnt_ls = ls;
        p.cache_target = target;
        p.scan_bpc = bpc;
    }
}

// Autotuner: times the count over the table sizes and table fills, then the
// scan grid, keeps the fastest, and records it in the tuning cache.
double batch_autotune(Batch &b, const mczi_counter_config &cfg, int mode, int launch, int order) {
    if (!b.plan) b.plan.reset(new McPlan());
    McPlan &p = *b.plan;
    const bool wide = cfg.k > 32;
    auto timed = [&](int ls, double target, unsigned bpc) {
        p.cnt_ls = ls;
        p.scan_bpc = bpc;
        // a candidate whose tables spill past an eighth of the spill table is retried
        // by the count at a quarter of its size; the size that ran is what
        // counts
        p.target_override = target;
        mczi_run_stats st{};
        batch_count(b, cfg, mode, launch, order, &st);  // calibrates / captures
        p.target_override = p.effective_target;
        batch_count(b, cfg, mode, launch, order, &st);
        return st.ms_total + st.ms_sort;
    };
    int best_ls = CNT_LOG2_SLOTS;
    unsigned best_bpc = 0;
    double best = 0, best_ms = 1e300;
    const std::vector<int> sizes = wide ? std::vector<int>{CNT_LOG2_SLOTS} : std::vector<int>{10, 11, 12};
    for (int ls : sizes)
        for (double c : {0.5, 1.0, 2.0, 4.0}) {
            const double ms = timed(ls, c * (double)(1 << ls), 0);
            if (ms < best_ms) {
                best_ms = ms;
                best = p.effective_target;
                best_ls = ls;
            }
        }
    for (unsigned bpc : {2u, 3u}) {
        const double ms = timed(best_ls, best, bpc);
        if (ms < best_ms) {
            best_ms = ms;
            best_bpc = bpc;
        }
    }
    p.cnt_ls = best_ls;
    p.scan_bpc = best_bpc;
    p.target_override = best;
    const int m = mode == MCZI_MODE_DATASETS_TWO_PASS ? MCZI_MODE_DATASETS
                  : mode == MCZI_MODE_TWO_PASS        ? MCZI_MODE_SINGLE_PASS
                                                      : mode;
    if (!wide) tune_store(tune_key(b, cfg, m, order), best_ls, best, best_bpc, best_ms);
    return best;
}

void batch_download(Batch &b, std::vector<mczi_counted_kmer> &out) {
    out.resize(b.n_out);
    batch_download(b, out.data());
}

void batch_download(Batch &b, mczi_counted_kmer *out) {
    Device &d = Device::get();
    if (!b.n_out) return;
    static_assert(sizeof(mczi_counted_kmer) == 24, "record layout");
    DevBuf rec(b.n_out * sizeof(mczi_counted_kmer));
    const unsigned g = (unsigned)std::min<uint64_t>((b.n_out + 255) / 256, (uint64_t)d.cu_count() * 16);
    hipLaunchKernelGGL(k_pack_counted, dim3(g), dim3(256), 0, d.stream(), b.out_kmer.as<unsigned long long>(),
                       b.k > 32 ? b.out_hi.as<unsigned long long>() : nullptr, b.out_count.as<uint32_t>(),
                       rec.as<unsigned long long>(), b.n_out);
    MCZI_HIP(hipGetLastError());
    d.download(out, rec.as<void>(), b.n_out * sizeof(mczi_counted_kmer), d.stream());
    if (!b.carries.empty())
        for (uint64_t i = 0; i < b.n_out; i++) {
            const auto it = b.carries.find({out[i].lo, out[i].hi});
            if (it != b.carries.end()) out[i].count += it->second;
        }
}

void batch_download_records(Batch &b, std::vector<uint64_t> &rec, std::vector<uint32_t> &part, uint32_t &rec_words,
                            uint32_t &p_int) {
    require(b.plan != nullptr && b.plan->k != 0, MCZI_E_INVALID, "the batch has not been counted");
    McPlan &p = *b.plan;
    require(p.grp_n <= 1, MCZI_E_INVALID, "a count in partition groups keeps no super-k-mer records");
    Device &d = Device::get();
    const bool wide = p.k > 32;
    rec_words = wide ? 4 : 2;
    p_int = p.p_int;
    const uint64_t n = p.n_records;  // partition order, no padding
    rec.resize(n * rec_words);
    part.resize(n);
    if (!n) return;
    d.download(rec.data(), wide ? p.stage_rec_w.as<void>() : p.stage_rec.as<void>(), n * rec_words * 8, d.stream());
    // the device layout to the one the interop writer reads: bases from bit 0,
    // base count at bits 48..55 of the last word; partition beside
    parallel_ranges(n, 1u << 16, host_workers(n), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            uint64_t *r = rec.data() + i * rec_words;
            if (wide) {
                const Rec32 x{{r[0], r[1], r[2], r[3]}};
                part[i] = rec_q(x);
                r[3] = (r[3] & 0xffffull) | ((uint64_t)rec_len(x) << 48);
            } else {
                const ulonglong2 x = make_ulonglong2(r[0], r[1]);
                part[i] = rec_q(x);
                r[1] = (r[1] & 0xffffffull) | ((uint64_t)rec_len(x) << 48);
            }
        }
    });
}

void batch_download_keys(Batch &b, uint64_t *lo, uint64_t *hi) {
    if (!b.n_out) return;
    Device &d = Device::get();
    d.download(lo, b.out_kmer.as<void>(), b.n_out * 8, d.stream());
    if (hi) {
        if (b.k > 32)
            d.download(hi, b.out_hi.as<void>(), b.n_out * 8, d.stream());
        else
            memset(hi, 0, b.n_out * 8);
    }
}

Batch::~Batch() = default;

// ---------------------------------------------------------------------------
// Plain device copy (16-byte vectors) used to calibrate HBM byte counters and
// to measure achievable bandwidth on the box.
__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

double device_copy_probe(uint64_t bytes, int reps) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    const uint64_t n16 = std::max<uint64_t>(bytes / 16, 1);
    DevBuf a(n16 * 16), b(n16 * 16);
    fill_zero_bytes(a.as<void>(), n16 * 16, s);
    const unsigned g = (unsigned)std::min<uint64_t>((n16 + 255) / 256, (uint64_t)d.cu_count() * 16);
    hipLaunchKernelGGL(k_copy16, dim3(g), dim3(256), 0, s, a.as<uint4>(), b.as<uint4>(), n16);
    hipEvent_t e0, e1;
    MCZI_HIP(hipEventCreate(&e0));
    MCZI_HIP(hipEventCreate(&e1));
    MCZI_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy16, dim3(g), dim3(256), 0, s, a.as<uint4>(), b.as<uint4>(), n16);
    MCZI_HIP(hipEventRecord(e1, s));
    MCZI_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MCZI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / std::max(reps, 1);
}

// ---------------------------------------------------------------------------
// run_hashes boundary: per-window selected hashes of one sequence on the
// device, deduplicated into runs on the host.
void device_run_hashes(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m,
                       std::vector<uint32_t> &hashes, std::vector<uint32_t> &starts) {
    hashes.clear();
    starts.clear();
    require(k >= 1 && k <= 64 && m >= 1 && m <= k && k - m + 1 <= MAXW, MCZI_E_INVALID,
            "invalid k / m for the minimizer scan");
    require(k % 2 == 1, MCZI_E_INVALID, "window length l = k must be odd");
    if (n_bases < k) return;
    HostReads r;
    for (uint64_t i = 0; i < n_bases; i++) r.append_base((packed[i / 4] >> (2 * (i % 4))) & 3u);
    r.close_fragment(k, 0);
    auto b = make_batch(std::move(r));
    McPlan &p = *b->plan;
    Device &d = Device::get();
    hipStream_t s = d.stream();
    const uint64_t nw = n_bases - k + 1;
    DevBuf win(n_bases * 4 + 64);
    DevBuf ctr(16 * 8);
    fill_u64(ctr.as<unsigned long long>(), 16, 0, s);
    ScanArgs sa{};
    sa.words = b->words.as<uint32_t>();
    sa.off = b->off.as<uint64_t>();
    sa.tile_start = p.tile_start.as<uint64_t>();
    sa.tile_frag = p.tile_frag.as<uint32_t>();
    sa.tile_len = p.tile_len.as<uint32_t>();
    sa.tile_begin = 0;
    sa.tile_end = p.n_tiles;
    sa.n_frag = b->n_frag;
    sa.k = (int)k;
    sa.m = (int)m;
    sa.w = (int)(k - m + 1);
    sa.p_int = 1;
    sa.wrec = 1;
    sa.win_out = win.as<uint32_t>();
    sa.stats = ctr.as<unsigned long long>();
    const dim3 wg((unsigned)std::min<uint64_t>(p.n_tiles, 2048));
    if (sa.w >= WPT) hipLaunchKernelGGL((k_scan<WINDOWS, true>), wg, dim3(SCAN_THREADS), 0, s, sa);
    else hipLaunchKernelGGL((k_scan<WINDOWS, false>), wg, dim3(SCAN_THREADS), 0, s, sa);
    MCZI_HIP(hipGetLastError());
    std::vector<uint32_t> v(nw);
    MCZI_HIP(hipMemcpyAsync(v.data(), win.as<void>(), nw * 4, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipStreamSynchronize(s));
    for (uint64_t j = 0; j < nw; j++) {
        if (j == 0 || v[j] != v[j - 1]) {
            hashes.push_back(v[j]);
            starts.push_back((uint32_t)j);
        }
    }
}

// ---------------------------------------------------------------------------
// Allocation-free run hashes (mczi_minimizer_run_hashes_noalloc): the device
// memory of a call lives in a file-scope array of fixed size, so the call
// makes no device allocation. In host code the array's name is only a
// stand-in: each GPU's copy is reached through the address
// hipGetSymbolAddress returns with that GPU current, fetched once per device
// and kept. Calls on one GPU share its copy (the library's call lock
// serializes them; concurrent use of the same GPU's copy outside it would
// race).
constexpr uint64_t NOALLOC_MAX_BASES = 1ull << 24;
constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NA_TILES = (NOALLOC_MAX_BASES + TILE - 1) / TILE;
constexpr size_t NA_CTR = 0;                                        // 16 counters
constexpr size_t NA_OFF = 256;                                      // fragment bounds {0, n}
constexpr size_t NA_TS = 512;                                       // tile starts (u64)
constexpr size_t NA_TF = NA_TS + na_align(NA_TILES * 8);            // tile fragments (u32)
constexpr size_t NA_TL = NA_TF + na_align(NA_TILES * 4);            // tile lengths (u32)
constexpr size_t NA_WORDS = NA_TL + na_align(NA_TILES * 4);         // packed words + 64 slack words
constexpr size_t NA_WIN = NA_WORDS + na_align((NOALLOC_MAX_BASES / 16 + 64) * 4);  // selected hash per window
constexpr size_t NA_BYTES = NA_WIN + na_align(NOALLOC_MAX_BASES * 4);
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// The device address of this GPU's scratch, from hipGetSymbolAddress on the
// first call with the device current.
static unsigned char *noalloc_scratch(int dev) {
    static std::mutex mu;
    static unsigned char *ptr[64] = {};
    require(dev >= 0 && dev < 64, MCZI_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void *a = nullptr;
        MCZI_HIP(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        require(a != nullptr, MCZI_E_DEVICE, "no device address for the scratch array");
        ptr[dev] = static_cast<unsigned char *>(a);
    }
    return ptr[dev];
}

// Counters zeroed, fragment bounds and the tile schedule of one sequence of n
// bases, and the packed words from word w0 on zeroed (the copy of the input
// fills the words below n; the rest is slack the scan reads).
__global__ void k_noalloc_setup(unsigned char *base, uint64_t n, uint64_t n_tiles, uint64_t w0, uint64_t w1) {
    unsigned long long *ctr = reinterpret_cast<unsigned long long *>(base + NA_CTR);
    uint64_t *off = reinterpret_cast<uint64_t *>(base + NA_OFF);
    uint64_t *ts = reinterpret_cast<uint64_t *>(base + NA_TS);
    uint32_t *tf = reinterpret_cast<uint32_t *>(base + NA_TF);
    uint32_t *tl = reinterpret_cast<uint32_t *>(base + NA_TL);
    uint32_t *words = reinterpret_cast<uint32_t *>(base + NA_WORDS);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < 16; i += stride) ctr[i] = 0;
    if (i0 == 0) {
        off[0] = 0;
        off[1] = n;
    }
    for (uint64_t t = i0; t < n_tiles; t += stride) {
        ts[t] = t * TILE;
        tf[t] = 0;
        tl[t] = (uint32_t)min((uint64_t)TILE, n - t * TILE);
    }
    for (uint64_t w = w0 + i0; w < w1; w += stride) words[w] = 0;
}

void device_run_hashes_noalloc(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m, uint32_t *hashes,
                               uint32_t *starts, size_t *n_runs) {
    Device &d = Device::get();
    hipStream_t s = d.stream();
    MCZI_HIP(hipSetDevice(d.device()));
    unsigned char *base = noalloc_scratch(d.device());
    const uint64_t nw = n_bases - k + 1;
    const uint64_t n_tiles = (n_bases + TILE - 1) / TILE;
    const uint64_t n_bytes = (n_bases + 3) / 4;
    const uint64_t w1 = (n_bases + 15) / 16 + 64;
    hipLaunchKernelGGL(k_noalloc_setup, dim3((unsigned)std::min<uint64_t>((w1 + 255) / 256 + 1, 1024)), dim3(256), 0,
                       s, base, n_bases, n_tiles, n_bytes / 4, w1);
    MCZI_HIP(hipMemcpyAsync(base + NA_WORDS, packed, n_bytes, hipMemcpyHostToDevice, s));
    ScanArgs sa{};
    sa.words = reinterpret_cast<const uint32_t *>(base + NA_WORDS);
    sa.off = reinterpret_cast<const uint64_t *>(base + NA_OFF);
    sa.tile_start = reinterpret_cast<const uint64_t *>(base + NA_TS);
    sa.tile_frag = reinterpret_cast<const uint32_t *>(base + NA_TF);
    sa.tile_len = reinterpret_cast<const uint32_t *>(base + NA_TL);
    sa.tile_begin = 0;
    sa.tile_end = n_tiles;
    sa.n_frag = 1;
    sa.k = (int)k;
    sa.m = (int)m;
    sa.w = (int)(k - m + 1);
    sa.p_int = 1;
    sa.wrec = 1;
    sa.win_out = reinterpret_cast<uint32_t *>(base + NA_WIN);
    sa.stats = reinterpret_cast<unsigned long long *>(base + NA_CTR);
    const dim3 g((unsigned)std::min<uint64_t>(n_tiles, 2048));
    if (sa.w >= WPT) hipLaunchKernelGGL((k_scan<WINDOWS, true>), g, dim3(SCAN_THREADS), 0, s, sa);
    else hipLaunchKernelGGL((k_scan<WINDOWS, false>), g, dim3(SCAN_THREADS), 0, s, sa);
    MCZI_HIP(hipGetLastError());
    MCZI_HIP(hipMemcpyAsync(hashes, base + NA_WIN, nw * 4, hipMemcpyDeviceToHost, s));
    MCZI_HIP(hipStreamSynchronize(s));
    // runs of equal selected hash, in place (the write index never passes the read index)
    size_t r = 0;
    uint32_t prev = 0;
    for (uint64_t j = 0; j < nw; j++) {
        const uint32_t v = hashes[j];
        if (j == 0 || v != prev) {
            hashes[r] = v;
            starts[r] = (uint32_t)j;
            r++;
        }
        prev = v;
    }
    *n_runs = r;
}

}  // namespace mczi
