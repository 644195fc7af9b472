/*
 * CPU oracle: plain-C restatement of the MCZI counting / subtraction path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Nothing in the product links it.
 *
 * Pinning (how each function is checked, beyond being a restatement):
 *  - or_nthash32: seq-hash 0.1.2 NtHasher<true> is not vendored. This
 *    restatement (ntHash seeds truncated to 32 bits, indexed by packed code,
 *    rotation 7, canonical = forward + reverse-complement, wrapping) is pinned
 *    by the three position vectors the reference holds
 *    (vendor/simd-minimizers-2.3.1/src/lib.rs:86-137, src/test.rs:401-425):
 *    a sweep over seed orderings, halves, rotations and combiners leaves this
 *    one variant as the only one matching all three.
 *  - or_window_hashes / or_minimizer_runs: restate the scalar specs
 *    (sm/minimizers.rs:167-219, sm/sliding_min.rs:279-340,
 *    sm/canonical.rs:11-29, sm/collect.rs:38-75); pinned by the same position
 *    vectors, the runs-tile-windows invariant (src/lib.rs:5005-5038) and
 *    reverse-complement symmetry (sm/test.rs:112-152).
 *  - or_count_two_pass / or_kff_u8_kmc / or_count_datasets: restate the
 *    reference drivers; their k-mer sets and counts are pinned against
 *    or_count_bruteforce (closed form: {K : c(K) > x}) and the reference test
 *    literals. KFF record ORDER and partition_bytes depend on the hash above.
 *  - or_zi_absent: exact set difference (closed form), pinned by the zi.rs
 *    test literals.
 *  - writers: restate src/lib.rs:1478-1670, 3071-3121; pinned by the
 *    reference's writer literals (src/lib.rs:5876-5991).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <string.h>

typedef unsigned __int128 u128;

void or_free(void *p) { free(p); }

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s ? s : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}
static void *xrealloc(void *p, size_t n) {
    void *q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return q;
}

/* ------------------------------------------------------------------------ */
/* Hashing                                                                  */
/* ------------------------------------------------------------------------ */

static inline uint32_t rol32(uint32_t x, unsigned r) {
    r &= 31u;
    return r ? (x << r) | (x >> (32u - r)) : x;
}

/* ntHash seed values (low 32 bits), indexed by packed code A0 C1 T2 G3.
 * Complement of a packed code is code ^ 2. */
static const uint32_t NT_SEED[4] = {0x95c60474u, 0x62a02b4cu, 0x82572324u, 0x4be24456u};
#define NT_ROT 7u

uint32_t or_nthash32(const uint8_t *c, int m, int canonical) {
    uint32_t fw = 0, rc = 0;
    for (int j = 0; j < m; j++) {
        fw ^= rol32(NT_SEED[c[j]], NT_ROT * (unsigned)(m - 1 - j));
        rc ^= rol32(NT_SEED[c[j] ^ 2u], NT_ROT * (unsigned)j);
    }
    return canonical ? fw + rc : fw;
}

/* src/lib.rs:2984-2990 */
uint32_t or_spread_hash_u32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    return v ^ (v >> 16);
}

/* src/lib.rs:2996-3002 */
uint64_t or_spread_hash_u64(uint64_t v) {
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdull;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ull;
    return v ^ (v >> 33);
}

/* ------------------------------------------------------------------------ */
/* Scan                                                                     */
/* ------------------------------------------------------------------------ */

/* Per window j in [0, n-k]: canonical minimizer selection.
 * sm/minimizers.rs:167-219 (scalar spec): hash of every m-mer, canonical
 * strand vote over l = m + w - 1 = k bases (sm/canonical.rs:11-29: #TG > l/2),
 * left/right sliding minimum on the upper 16 bits (sm/sliding_min.rs:279-340:
 * ties to the leftmost resp. rightmost position), returning the full hash. */
size_t or_window_hashes(const uint8_t *c, size_t n, int k, int m, uint32_t *out) {
    if (n < (size_t)k) return 0;
    int w = k - m + 1;
    size_t nm = n - (size_t)m + 1;
    uint32_t *h = (uint32_t *)xmalloc(nm * sizeof(uint32_t));
    for (size_t i = 0; i < nm; i++) h[i] = or_nthash32(c + i, m, 1);
    size_t nw = n - (size_t)k + 1;
    /* T/G counts via prefix sums */
    uint32_t tg = 0;
    for (int t = 0; t < k; t++) tg += (c[t] & 2u) ? 1u : 0u;
    for (size_t j = 0; j < nw; j++) {
        if (j > 0) {
            tg -= (c[j - 1] & 2u) ? 1u : 0u;
            tg += (c[j + (size_t)k - 1] & 2u) ? 1u : 0u;
        }
        size_t lpos = j, rpos = j;
        uint32_t best = h[j] >> 16;
        for (int i = 1; i < w; i++) {
            uint32_t v = h[j + (size_t)i] >> 16;
            if (v < best) {
                best = v;
                lpos = rpos = j + (size_t)i;
            } else if (v == best) {
                rpos = j + (size_t)i;
            }
        }
        out[j] = (2u * tg > (uint32_t)k) ? h[lpos] : h[rpos];
    }
    free(h);
    return nw;
}

/* collect_and_dedup_with_index (sm/collect.rs:38-75). */
size_t or_minimizer_runs(const uint8_t *c, size_t n, int k, int m, uint32_t *run_hash,
                         uint32_t *run_start) {
    if (n < (size_t)k) return 0;
    size_t nw = n - (size_t)k + 1;
    uint32_t *v = (uint32_t *)xmalloc(nw * sizeof(uint32_t));
    or_window_hashes(c, n, k, m, v);
    size_t r = 0;
    for (size_t j = 0; j < nw; j++) {
        if (j == 0 || v[j] != v[j - 1]) {
            run_hash[r] = v[j];
            run_start[r] = (uint32_t)j;
            r++;
        }
    }
    free(v);
    return r;
}

/* Positions API used by the reference's known-answer vectors
 * (sm/lib.rs:86-137): per window the selected m-mer position, deduplicated. */
size_t or_minimizer_positions(const uint8_t *c, size_t n, int kh, int w, int canonical,
                              uint32_t *out) {
    int l = kh + w - 1;
    if (n < (size_t)l) return 0;
    size_t nm = n - (size_t)kh + 1;
    uint32_t *h = (uint32_t *)xmalloc(nm * sizeof(uint32_t));
    for (size_t i = 0; i < nm; i++) h[i] = or_nthash32(c + i, kh, canonical);
    size_t cnt = 0;
    for (size_t j = 0; j + (size_t)l <= n; j++) {
        size_t lpos = j, rpos = j;
        uint32_t best = h[j] >> 16;
        for (int i = 1; i < w; i++) {
            uint32_t v = h[j + (size_t)i] >> 16;
            if (v < best) {
                best = v;
                lpos = rpos = j + (size_t)i;
            } else if (v == best) {
                rpos = j + (size_t)i;
            }
        }
        size_t p = lpos;
        if (canonical) {
            uint32_t tg = 0;
            for (int t = 0; t < l; t++) tg += (c[j + (size_t)t] & 2u) ? 1u : 0u;
            p = (2u * tg > (uint32_t)l) ? lpos : rpos;
        }
        if (cnt == 0 || out[cnt - 1] != (uint32_t)p) out[cnt++] = (uint32_t)p;
    }
    free(h);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* k-mer helpers                                                            */
/* ------------------------------------------------------------------------ */

static inline uint32_t mc_bits(uint8_t code) { return (uint32_t)(code ^ (code >> 1)); }

static inline u128 kmask128(int k) {
    return k == 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
}

uint64_t or_canonical_u64(const uint8_t *c, int k) {
    u128 fwd = 0, rev = 0, mask = kmask128(k);
    for (int i = 0; i < k; i++) {
        u128 b = mc_bits(c[i]);
        fwd = ((fwd << 2) | b) & mask;
        rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
    }
    return (uint64_t)(fwd < rev ? fwd : rev);
}

static int cmp_u128(const void *a, const void *b) {
    u128 x = *(const u128 *)a, y = *(const u128 *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
static int cmp_counted(const void *a, const void *b) {
    const or_counted *x = (const or_counted *)a, *y = (const or_counted *)b;
    if (x->hi != y->hi) return x->hi < y->hi ? -1 : 1;
    if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
    return 0;
}

/* Canonical k-mers of every window of a fragment (src/lib.rs:3589-3614). */
static size_t canon_windows(const uint8_t *c, size_t n, int k, u128 *out) {
    if (n < (size_t)k) return 0;
    u128 fwd = 0, rev = 0, mask = kmask128(k);
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
        u128 b = mc_bits(c[i]);
        fwd = ((fwd << 2) | b) & mask;
        rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
        if (i + 1 >= (size_t)k) out[o++] = fwd < rev ? fwd : rev;
    }
    return o;
}

/* ------------------------------------------------------------------------ */
/* Open-addressing tables                                                   */
/* ------------------------------------------------------------------------ */

/* u128 -> u64 map with tombstones (used for counts and simplitig removal). */
typedef struct {
    u128 *keys;
    uint64_t *vals;
    uint8_t *state; /* 0 empty, 1 used, 2 deleted */
    size_t cap, used, filled;
} map128;

static inline size_t h128(u128 k) {
    uint64_t a = (uint64_t)k, b = (uint64_t)(k >> 64);
    return (size_t)or_spread_hash_u64(a ^ or_spread_hash_u64(b + 0x9e3779b97f4a7c15ull));
}
static void map_init(map128 *m, size_t want) {
    size_t cap = 64;
    while (cap < want * 2) cap <<= 1;
    m->cap = cap;
    m->used = m->filled = 0;
    m->keys = (u128 *)xmalloc(cap * sizeof(u128));
    m->vals = (uint64_t *)xcalloc(cap, sizeof(uint64_t));
    m->state = (uint8_t *)xcalloc(cap, 1);
}
static void map_free(map128 *m) {
    free(m->keys);
    free(m->vals);
    free(m->state);
}
static void map_grow(map128 *m);
static uint64_t *map_slot(map128 *m, u128 key, int create) {
    if (create && (m->filled + 1) * 10 >= m->cap * 7) map_grow(m);
    size_t mask = m->cap - 1, i = h128(key) & mask, tomb = (size_t)-1;
    for (;;) {
        if (m->state[i] == 0) {
            if (!create) return NULL;
            if (tomb != (size_t)-1) i = tomb;
            else m->filled++;
            m->state[i] = 1;
            m->keys[i] = key;
            m->vals[i] = 0;
            m->used++;
            return &m->vals[i];
        }
        if (m->state[i] == 1 && m->keys[i] == key) return &m->vals[i];
        if (m->state[i] == 2 && tomb == (size_t)-1) tomb = i;
        i = (i + 1) & mask;
    }
}
static void map_grow(map128 *m) {
    map128 old = *m;
    map_init(m, old.used + 1);
    for (size_t i = 0; i < old.cap; i++)
        if (old.state[i] == 1) *map_slot(m, old.keys[i], 1) = old.vals[i];
    map_free(&old);
}
static int map_remove(map128 *m, u128 key) {
    size_t mask = m->cap - 1, i = h128(key) & mask;
    for (;;) {
        if (m->state[i] == 0) return 0;
        if (m->state[i] == 1 && m->keys[i] == key) {
            m->state[i] = 2;
            m->used--;
            return 1;
        }
        i = (i + 1) & mask;
    }
}

/* CompactMinimizerTable restatement (src/lib.rs:4171-4251): u32 key -> u8
 * count saturating at `sat`, linear probing from key & mask, growth at 0.7. */
typedef struct {
    uint32_t *keys;
    uint8_t *counts;
    size_t cap, items;
    uint8_t sat;
} mintab;

static void mintab_init(mintab *t, size_t cap, uint8_t sat) {
    size_t c = 64;
    while (c < cap) c <<= 1;
    t->cap = c;
    t->items = 0;
    t->sat = sat;
    t->keys = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    t->counts = (uint8_t *)xcalloc(c, 1);
}
static void mintab_add_nogrow(mintab *t, uint32_t key, uint8_t delta) {
    size_t mask = t->cap - 1, i = key & mask;
    for (;;) {
        if (t->counts[i] == 0) {
            t->keys[i] = key;
            t->counts[i] = delta < t->sat ? delta : t->sat;
            t->items++;
            return;
        }
        if (t->keys[i] == key) {
            if (t->counts[i] < t->sat) {
                unsigned s = (unsigned)t->counts[i] + delta;
                if (s > 255u) s = 255u;
                t->counts[i] = (uint8_t)(s < t->sat ? s : t->sat);
            }
            return;
        }
        i = (i + 1) & mask;
    }
}
static void mintab_add(mintab *t, uint32_t key, uint8_t delta) {
    if ((t->items + 1) * 10 >= t->cap * 7) {
        mintab old = *t;
        mintab_init(t, old.cap * 2, old.sat);
        for (size_t i = 0; i < old.cap; i++)
            if (old.counts[i]) mintab_add_nogrow(t, old.keys[i], old.counts[i]);
        free(old.keys);
        free(old.counts);
    }
    mintab_add_nogrow(t, key, delta);
}
static uint8_t mintab_get(const mintab *t, uint32_t key) {
    size_t mask = t->cap - 1, i = key & mask;
    for (;;) {
        if (t->counts[i] == 0) return 0;
        if (t->keys[i] == key) return t->counts[i];
        i = (i + 1) & mask;
    }
}

/* ------------------------------------------------------------------------ */
/* Runs of a fragment as (key, start, end) (src/lib.rs:3694-3789)           */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t key;
    uint32_t start, end;
} run_t;

static size_t fragment_runs(const uint8_t *c, size_t n, int k, int m, run_t *runs,
                            uint32_t *tmp_h, uint32_t *tmp_s) {
    size_t r = or_minimizer_runs(c, n, k, m, tmp_h, tmp_s);
    size_t windows = n - (size_t)k + 1;
    for (size_t i = 0; i < r; i++) {
        runs[i].key = or_spread_hash_u32(tmp_h[i]);
        runs[i].start = tmp_s[i];
        runs[i].end = (i + 1 < r) ? tmp_s[i + 1] : (uint32_t)windows;
    }
    return r;
}

/* Packed-seq bytes of a fragment: 4 codes per byte, LSB first. */
static uint8_t *pack_codes(const uint8_t *c, size_t n) {
    uint8_t *p = (uint8_t *)xcalloc(n / 4 + 2, 1);
    for (size_t i = 0; i < n; i++) p[i / 4] |= (uint8_t)(c[i] << (2 * (i % 4)));
    return p;
}

/* byte buffer */
typedef struct {
    uint8_t *d;
    size_t n, cap;
} bbuf;
static void bb_push(bbuf *b, const void *src, size_t n) {
    if (b->n + n > b->cap) {
        size_t c = b->cap ? b->cap * 2 : 256;
        while (c < b->n + n) c *= 2;
        b->d = (uint8_t *)xrealloc(b->d, c);
        b->cap = c;
    }
    memcpy(b->d + b->n, src, n);
    b->n += n;
}
static void bb_u8(bbuf *b, uint8_t v) { bb_push(b, &v, 1); }
static void bb_be32(bbuf *b, uint32_t v) {
    uint8_t t[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
    bb_push(b, t, 4);
}
static void bb_be64(bbuf *b, uint64_t v) {
    uint8_t t[8];
    for (int i = 0; i < 8; i++) t[i] = (uint8_t)(v >> (56 - 8 * i));
    bb_push(b, t, 8);
}

/* append_record_len + append_packed_dna_range (src/lib.rs:3209-3216,
 * 3170-3190): super-k-mer record of bases [start, start+len). */
static void append_superkmer(bbuf *b, const uint8_t *packed, size_t packed_len, size_t start,
                             size_t len) {
    if (len < 255) bb_u8(b, (uint8_t)len);
    else {
        bb_u8(b, 255);
        bb_be32(b, (uint32_t)len);
    }
    size_t sb = start / 4, bytes = (len + 3) / 4, shift = 2 * (start % 4);
    for (size_t i = 0; i < bytes; i++) {
        size_t s = sb + i;
        unsigned word = packed[s] | ((s + 1 < packed_len ? packed[s + 1] : 0u) << 8);
        bb_u8(b, (uint8_t)((word >> shift) & 0xffu));
    }
}

/* Iterate the records of a partition buffer (read_superkmer_partition,
 * src/lib.rs:2692-2711, read_record_len 2756-2769); calls fn per base code. */
typedef void (*rec_fn)(void *ctx, const uint8_t *packed, size_t len);
static void for_each_record(const bbuf *b, rec_fn fn, void *ctx) {
    size_t p = 0;
    while (p < b->n) {
        size_t len = b->d[p++];
        if (len == 255) {
            len = ((size_t)b->d[p] << 24) | ((size_t)b->d[p + 1] << 16) |
                  ((size_t)b->d[p + 2] << 8) | b->d[p + 3];
            p += 4;
        }
        fn(ctx, b->d + p, len);
        p += (len + 3) / 4;
    }
}
static inline uint32_t packed_base_bits(const uint8_t *packed, size_t idx) {
    uint32_t bits = (packed[idx / 4] >> (2 * (idx % 4))) & 3u;
    return bits ^ (bits >> 1);
}

/* ------------------------------------------------------------------------ */
/* Brute-force counting (closed form {K : c(K) > x})                        */
/* ------------------------------------------------------------------------ */

static u128 *all_canonical(const or_frags *f, int k, size_t *n_out) {
    size_t total = 0;
    for (size_t i = 0; i < f->nfrag; i++) {
        size_t n = f->off[i + 1] - f->off[i];
        if (n >= (size_t)k) total += n - (size_t)k + 1;
    }
    u128 *all = (u128 *)xmalloc(total * sizeof(u128));
    size_t o = 0;
    for (size_t i = 0; i < f->nfrag; i++)
        o += canon_windows(f->codes + f->off[i], f->off[i + 1] - f->off[i], k, all + o);
    *n_out = o;
    return all;
}

int or_count_bruteforce(const or_frags *f, int k, uint64_t x, or_counted **out, size_t *n_out) {
    size_t n;
    u128 *all = all_canonical(f, k, &n);
    qsort(all, n, sizeof(u128), cmp_u128);
    or_counted *res = (or_counted *)xmalloc(n * sizeof(or_counted));
    size_t r = 0;
    for (size_t i = 0; i < n;) {
        size_t j = i;
        while (j < n && all[j] == all[i]) j++;
        if ((uint64_t)(j - i) > x) {
            res[r].lo = (uint64_t)all[i];
            res[r].hi = (uint64_t)(all[i] >> 64);
            res[r].count = j - i;
            r++;
        }
        i = j;
    }
    free(all);
    *out = res;
    *n_out = r;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Two-pass partitioned counting (src/lib.rs:920-992)                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    map128 *counts;
    int k;
} count_ctx;

/* add_packed_superkmer_kmer_counts (src/lib.rs:3589-3614) */
static void count_record(void *vctx, const uint8_t *packed, size_t len) {
    count_ctx *ctx = (count_ctx *)vctx;
    int k = ctx->k;
    if (len < (size_t)k) return;
    u128 fwd = 0, rev = 0, mask = kmask128(k);
    for (size_t i = 0; i < len; i++) {
        u128 b = packed_base_bits(packed, i);
        fwd = ((fwd << 2) | b) & mask;
        rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
        if (i + 1 >= (size_t)k) (*map_slot(ctx->counts, fwd < rev ? fwd : rev, 1))++;
    }
}

static size_t max_frag_len(const or_frags *f) {
    size_t mx = 0;
    for (size_t i = 0; i < f->nfrag; i++)
        if (f->off[i + 1] - f->off[i] > mx) mx = f->off[i + 1] - f->off[i];
    return mx;
}

/* Builds the super-k-mer partitions: phase 1 (pass-1 minimizer counts) when
 * `filter`, then phase 2 (src/lib.rs:3483-3560). */
static bbuf *build_partitions(const or_frags *f, int k, int m, uint64_t x, size_t P, int filter,
                              uint64_t *kept) {
    size_t mx = max_frag_len(f) + 1;
    run_t *runs = (run_t *)xmalloc(mx * sizeof(run_t));
    uint32_t *th = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    uint32_t *ts = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    mintab *shards = NULL;
    if (filter) {
        /* ShardedMinimizerCounts (src/lib.rs:4524-4563): saturation x+1 */
        shards = (mintab *)xmalloc(P * sizeof(mintab));
        for (size_t s = 0; s < P; s++) mintab_init(&shards[s], 64, (uint8_t)(x + 1));
        for (size_t i = 0; i < f->nfrag; i++) {
            const uint8_t *c = f->codes + f->off[i];
            size_t n = f->off[i + 1] - f->off[i];
            if (n < (size_t)k) continue;
            size_t r = fragment_runs(c, n, k, m, runs, th, ts);
            for (size_t j = 0; j < r; j++) {
                size_t len = runs[j].end - runs[j].start;
                uint8_t delta = (uint8_t)(len < 255 ? len : 255); /* src/lib.rs:2847-2849 */
                mintab_add(&shards[or_spread_hash_u32(runs[j].key) % P], runs[j].key, delta);
            }
        }
        if (kept) {
            uint64_t u = 0;
            for (size_t s = 0; s < P; s++) u += shards[s].items;
            *kept = u;
        }
    }
    bbuf *parts = (bbuf *)xcalloc(P, sizeof(bbuf));
    for (size_t i = 0; i < f->nfrag; i++) {
        const uint8_t *c = f->codes + f->off[i];
        size_t n = f->off[i + 1] - f->off[i];
        if (n < (size_t)k) continue;
        uint8_t *packed = pack_codes(c, n);
        size_t plen = (n + 3) / 4;
        size_t r = fragment_runs(c, n, k, m, runs, th, ts);
        for (size_t j = 0; j < r; j++) {
            if (filter) {
                size_t shard = or_spread_hash_u32(runs[j].key) % P;
                if (!((uint64_t)mintab_get(&shards[shard], runs[j].key) > x)) continue;
            }
            size_t send = runs[j].end + (size_t)k - 1;
            if (send > n) continue;
            size_t part = or_spread_hash_u32(runs[j].key) % P; /* minimizer_shard */
            append_superkmer(&parts[part], packed, plen, runs[j].start, send - runs[j].start);
        }
        free(packed);
    }
    if (shards) {
        for (size_t s = 0; s < P; s++) {
            free(shards[s].keys);
            free(shards[s].counts);
        }
        free(shards);
    }
    free(runs);
    free(th);
    free(ts);
    return parts;
}

int or_count_two_pass(const or_frags *f, int k, int m, uint64_t x, size_t P, or_counted **out,
                      size_t *n_out, uint64_t *partition_bytes, uint64_t *kept) {
    if (k < 1 || k > 64 || m < 1 || m > k || P == 0 || x >= 255 || (k % 2) == 0) return -1;
    bbuf *parts = build_partitions(f, k, m, x, P, 1, kept);
    uint64_t pb = 0;
    or_counted *res = NULL;
    size_t r = 0, rcap = 0;
    for (size_t p = 0; p < P; p++) {
        pb += parts[p].n;
        map128 counts;
        map_init(&counts, 1024);
        count_ctx ctx = {&counts, k};
        for_each_record(&parts[p], count_record, &ctx);
        for (size_t i = 0; i < counts.cap; i++) {
            if (counts.state[i] != 1 || !(counts.vals[i] > x)) continue; /* src/lib.rs:2417 */
            if (r == rcap) {
                rcap = rcap ? rcap * 2 : 1024;
                res = (or_counted *)xrealloc(res, rcap * sizeof(or_counted));
            }
            res[r].lo = (uint64_t)counts.keys[i];
            res[r].hi = (uint64_t)(counts.keys[i] >> 64);
            res[r].count = counts.vals[i];
            r++;
        }
        map_free(&counts);
        free(parts[p].d);
    }
    free(parts);
    if (r) qsort(res, r, sizeof(or_counted), cmp_counted); /* src/lib.rs:285 */
    *out = res;
    *n_out = r;
    if (partition_bytes) *partition_bytes = pb;
    return 0;
}

/* Multithreaded two-pass count: the phases of count_inputs (src/lib.rs:
 * 920-992) with the reference's parallel structure (Rayon over input
 * batches and over partitions, ShardedMinimizerCounts shards):
 * 1. each thread counts the minimizers of its fragments into its own P
 *    shards (saturating at x + 1); shards are merged per shard, which is
 *    exact because every partial count is already clamped at x + 1;
 * 2. each thread writes the filtered super-k-mers of its fragments into its
 *    own P partition buffers;
 * 3. each partition (all threads' buffers) is counted by one thread.
 * The result is sorted as in or_count_two_pass. */
typedef struct {
    const or_frags *f;
    int k, m;
    uint64_t x;
    size_t P, lo, hi;  /* fragments [lo, hi) in phases 1-2, partitions p = t mod T in 3 */
    int t, T;
    mintab *shards;          /* phase 1: this thread's P shards */
    mintab **all_shards;     /* phase 1 merge: every thread's shards */
    mintab *merged;          /* phase 2: merged shards (read only) */
    bbuf *parts;             /* phase 2: this thread's P buffers */
    bbuf **all_parts;        /* phase 3: every thread's buffers */
    or_counted *res;         /* phase 3: survivors */
    size_t nres, capres;
} mt_job;

static void *mt_minimizers(void *arg) {
    mt_job *j = (mt_job *)arg;
    size_t mx = max_frag_len(j->f) + 1;
    run_t *runs = (run_t *)xmalloc(mx * sizeof(run_t));
    uint32_t *th = (uint32_t *)xmalloc(mx * sizeof(uint32_t)), *ts = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    for (size_t i = j->lo; i < j->hi; i++) {
        const uint8_t *c = j->f->codes + j->f->off[i];
        size_t n = j->f->off[i + 1] - j->f->off[i];
        if (n < (size_t)j->k) continue;
        size_t r = fragment_runs(c, n, j->k, j->m, runs, th, ts);
        for (size_t q = 0; q < r; q++) {
            size_t len = runs[q].end - runs[q].start;
            mintab_add(&j->shards[or_spread_hash_u32(runs[q].key) % j->P], runs[q].key,
                       (uint8_t)(len < 255 ? len : 255));
        }
    }
    free(runs);
    free(th);
    free(ts);
    return NULL;
}

static void *mt_merge(void *arg) {
    mt_job *j = (mt_job *)arg;
    for (size_t s = (size_t)j->t; s < j->P; s += (size_t)j->T) {
        mintab *dst = &j->merged[s];
        mintab_init(dst, 64, (uint8_t)(j->x + 1));
        for (int u = 0; u < j->T; u++) {
            const mintab *src = &j->all_shards[u][s];
            for (size_t i = 0; i < src->cap; i++)
                if (src->counts[i]) mintab_add(dst, src->keys[i], src->counts[i]);
        }
    }
    return NULL;
}

static void *mt_partition(void *arg) {
    mt_job *j = (mt_job *)arg;
    size_t mx = max_frag_len(j->f) + 1;
    run_t *runs = (run_t *)xmalloc(mx * sizeof(run_t));
    uint32_t *th = (uint32_t *)xmalloc(mx * sizeof(uint32_t)), *ts = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    for (size_t i = j->lo; i < j->hi; i++) {
        const uint8_t *c = j->f->codes + j->f->off[i];
        size_t n = j->f->off[i + 1] - j->f->off[i];
        if (n < (size_t)j->k) continue;
        uint8_t *packed = pack_codes(c, n);
        size_t plen = (n + 3) / 4;
        size_t r = fragment_runs(c, n, j->k, j->m, runs, th, ts);
        for (size_t q = 0; q < r; q++) {
            size_t shard = or_spread_hash_u32(runs[q].key) % j->P;
            if (!((uint64_t)mintab_get(&j->merged[shard], runs[q].key) > j->x)) continue;
            size_t send = runs[q].end + (size_t)j->k - 1;
            if (send > n) continue;
            append_superkmer(&j->parts[shard], packed, plen, runs[q].start, send - runs[q].start);
        }
        free(packed);
    }
    free(runs);
    free(th);
    free(ts);
    return NULL;
}

static void *mt_count(void *arg) {
    mt_job *j = (mt_job *)arg;
    for (size_t p = (size_t)j->t; p < j->P; p += (size_t)j->T) {
        map128 counts;
        map_init(&counts, 1024);
        count_ctx ctx = {&counts, j->k};
        for (int u = 0; u < j->T; u++) for_each_record(&j->all_parts[u][p], count_record, &ctx);
        for (size_t i = 0; i < counts.cap; i++) {
            if (counts.state[i] != 1 || !(counts.vals[i] > j->x)) continue; /* src/lib.rs:2417 */
            if (j->nres == j->capres) {
                j->capres = j->capres ? j->capres * 2 : 1024;
                j->res = (or_counted *)xrealloc(j->res, j->capres * sizeof(or_counted));
            }
            j->res[j->nres].lo = (uint64_t)counts.keys[i];
            j->res[j->nres].hi = (uint64_t)(counts.keys[i] >> 64);
            j->res[j->nres].count = counts.vals[i];
            j->nres++;
        }
        map_free(&counts);
    }
    return NULL;
}

static void mt_run(mt_job *jobs, int T, void *(*fn)(void *)) {
    pthread_t *th = (pthread_t *)xmalloc(sizeof(pthread_t) * (size_t)T);
    for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, fn, &jobs[t]);
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
    free(th);
}

int or_count_two_pass_mt(const or_frags *f, int k, int m, uint64_t x, size_t P, int threads, or_counted **out,
                         size_t *n_out) {
    if (k < 1 || k > 64 || m < 1 || m > k || P == 0 || x >= 255 || (k % 2) == 0 || threads < 1) return -1;
    const int T = threads;
    mt_job *jobs = (mt_job *)xcalloc((size_t)T, sizeof(mt_job));
    mintab **all_shards = (mintab **)xmalloc(sizeof(mintab *) * (size_t)T);
    bbuf **all_parts = (bbuf **)xmalloc(sizeof(bbuf *) * (size_t)T);
    mintab *merged = (mintab *)xcalloc(P, sizeof(mintab));
    /* fragment ranges of about equal bases */
    const uint64_t total = f->off[f->nfrag];
    size_t lo = 0;
    for (int t = 0; t < T; t++) {
        size_t hi = lo;
        const uint64_t goal = total / (uint64_t)T * (uint64_t)(t + 1);
        while (hi < f->nfrag && (t == T - 1 || f->off[hi] < goal)) hi++;
        mt_job *j = &jobs[t];
        j->f = f;
        j->k = k;
        j->m = m;
        j->x = x;
        j->P = P;
        j->lo = lo;
        j->hi = hi;
        j->t = t;
        j->T = T;
        j->shards = (mintab *)xmalloc(P * sizeof(mintab));
        for (size_t s = 0; s < P; s++) mintab_init(&j->shards[s], 64, (uint8_t)(x + 1));
        j->parts = (bbuf *)xcalloc(P, sizeof(bbuf));
        j->all_shards = all_shards;
        j->all_parts = all_parts;
        j->merged = merged;
        all_shards[t] = j->shards;
        all_parts[t] = j->parts;
        lo = hi;
    }
    mt_run(jobs, T, mt_minimizers);
    mt_run(jobs, T, mt_merge);
    mt_run(jobs, T, mt_partition);
    mt_run(jobs, T, mt_count);
    size_t r = 0;
    for (int t = 0; t < T; t++) r += jobs[t].nres;
    or_counted *res = (or_counted *)xmalloc((r ? r : 1) * sizeof(or_counted));
    r = 0;
    for (int t = 0; t < T; t++) {
        if (jobs[t].nres) memcpy(res + r, jobs[t].res, jobs[t].nres * sizeof(or_counted));
        r += jobs[t].nres;
        free(jobs[t].res);
        for (size_t s = 0; s < P; s++) {
            free(jobs[t].shards[s].keys);
            free(jobs[t].shards[s].counts);
            free(jobs[t].parts[s].d);
        }
        free(jobs[t].shards);
        free(jobs[t].parts);
    }
    for (size_t s = 0; s < P; s++) {
        free(merged[s].keys);
        free(merged[s].counts);
    }
    free(merged);
    free(all_shards);
    free(all_parts);
    free(jobs);
    if (r) qsort(res, r, sizeof(or_counted), cmp_counted); /* src/lib.rs:285 */
    *out = res;
    *n_out = r;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* KMC-style single-pass KFF u8 path (src/lib.rs:599-635, 2427-2593)        */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t **b;
    size_t *n, *cap;
    int k;
} bucket_ctx;

/* append_packed_superkmer_kmers_u64_buckets (src/lib.rs:3642-3669) */
static void bucket_record(void *vctx, const uint8_t *packed, size_t len) {
    bucket_ctx *ctx = (bucket_ctx *)vctx;
    int k = ctx->k;
    if (len < (size_t)k) return;
    uint64_t fwd = 0, rev = 0, mask = (k == 32) ? ~0ull : ((1ull << (2 * k)) - 1);
    for (size_t i = 0; i < len; i++) {
        uint64_t b = packed_base_bits(packed, i);
        fwd = ((fwd << 2) | b) & mask;
        rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
        if (i + 1 >= (size_t)k) {
            uint64_t e = fwd < rev ? fwd : rev;
            size_t bi = e & 255u;
            if (ctx->n[bi] == ctx->cap[bi]) {
                ctx->cap[bi] = ctx->cap[bi] ? ctx->cap[bi] * 2 : 64;
                ctx->b[bi] = (uint64_t *)xrealloc(ctx->b[bi], ctx->cap[bi] * sizeof(uint64_t));
            }
            ctx->b[bi][ctx->n[bi]++] = e;
        }
    }
}

static void kff_header(bbuf *o, int k, int m, uint64_t x, uint64_t records, int ordered,
                       uint64_t count_size) {
    /* write_kff_header_with_count_size (src/lib.rs:1633-1664) */
    char fb[256];
    int fl = snprintf(fb, sizeof fb,
                      "producer=MC\nk=%d\nm=%d\nthreshold=%llu\nfilter=count>threshold\n"
                      "canonical=1\n",
                      k, m, (unsigned long long)x);
    bb_push(o, "KFF", 3);
    uint8_t ver[5] = {1, 0, 0x1b, 1, 1};
    bb_push(o, ver, 5);
    bb_be32(o, (uint32_t)fl);
    bb_push(o, fb, (size_t)fl);
    /* write_kff_values (src/lib.rs:3071-3080) */
    bb_u8(o, 'v');
    bb_be64(o, 6);
    const char *names[6] = {"k", "max", "data_size", "ordered", "m", "threshold"};
    uint64_t vals[6] = {(uint64_t)k, 1, count_size, (uint64_t)ordered, (uint64_t)m, x};
    for (int i = 0; i < 6; i++) {
        bb_push(o, names[i], strlen(names[i]) + 1);
        bb_be64(o, vals[i]);
    }
    bb_u8(o, 'r');
    bb_be64(o, records);
}

static void kff_record(bbuf *o, u128 enc, int k, uint64_t count, int count_size) {
    size_t bytes = ((size_t)k + 3) / 4;
    uint8_t be[16];
    for (int i = 0; i < 16; i++) be[i] = (uint8_t)(enc >> (8 * (15 - i)));
    bb_push(o, be + 16 - bytes, bytes);
    for (int i = count_size - 1; i >= 0; i--) bb_u8(o, (uint8_t)(count >> (8 * i)));
}

int or_kff_u8_kmc(const or_frags *f, int k, int m, uint64_t x, size_t P, uint8_t **bytes,
                  size_t *n_bytes, uint64_t *records) {
    if (k < 1 || k > 32 || m < 1 || m > k || P == 0 || x >= 255 || (k % 2) == 0) return -1;
    bbuf *parts = build_partitions(f, k, m, x, P, 0, NULL);
    bbuf body = {0};
    uint64_t emitted = 0;
    uint8_t sat = (uint8_t)((x + 1) < 255 ? (x + 1) : 255);
    for (size_t p = 0; p < P; p++) {
        bucket_ctx ctx;
        ctx.b = (uint64_t **)xcalloc(256, sizeof(uint64_t *));
        ctx.n = (size_t *)xcalloc(256, sizeof(size_t));
        ctx.cap = (size_t *)xcalloc(256, sizeof(size_t));
        ctx.k = k;
        for_each_record(&parts[p], bucket_record, &ctx);
        for (size_t bi = 0; bi < 256; bi++) {
            uint64_t *v = ctx.b[bi];
            size_t n = ctx.n[bi];
            qsort(v, n, sizeof(uint64_t), cmp_u64);
            /* write_sorted_kmers_u8 (src/lib.rs:2545-2570) */
            for (size_t i = 0; i < n;) {
                uint64_t e = v[i++];
                uint8_t count = 1;
                while (i < n && v[i] == e) {
                    if (count < sat) count++;
                    i++;
                }
                if ((uint64_t)count > x) {
                    kff_record(&body, (u128)e, k, count, 1);
                    emitted++;
                }
            }
            free(v);
        }
        free(ctx.b);
        free(ctx.n);
        free(ctx.cap);
        free(parts[p].d);
    }
    free(parts);
    bbuf o = {0};
    kff_header(&o, k, m, x, emitted, 0, 1);
    if (body.n) bb_push(&o, body.d, body.n);
    bb_push(&o, "KFF", 3);
    free(body.d);
    *bytes = o.d;
    *n_bytes = o.n;
    if (records) *records = emitted;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* fof dataset presence (src/lib.rs:637-703, 4253-4370, 2184-2247)          */
/* ------------------------------------------------------------------------ */

#define SEEN_BIT 0x80000000u
#define COUNT_MASK 0x7fffffffu

/* DatasetMinimizerTable (src/lib.rs:4253-4369) */
typedef struct {
    uint32_t *keys, *states;
    size_t cap, items;
} dstab;

static void dstab_init(dstab *t, size_t cap) {
    size_t c = 64;
    while (c < cap) c <<= 1;
    t->cap = c;
    t->items = 0;
    t->keys = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    t->states = (uint32_t *)xcalloc(c, sizeof(uint32_t));
}
static void dstab_insert_state(dstab *t, uint32_t key, uint32_t state) {
    size_t mask = t->cap - 1, i = key & mask;
    while (t->states[i] != 0) i = (i + 1) & mask;
    t->keys[i] = key;
    t->states[i] = state;
    t->items++;
}
static void dstab_rebuild(dstab *t, size_t cap, int filter, size_t remaining, uint64_t x) {
    dstab old = *t;
    size_t keep = 0;
    if (filter) {
        for (size_t i = 0; i < old.cap; i++) {
            if (!old.states[i]) continue;
            uint32_t c = old.states[i] & COUNT_MASK;
            if ((uint64_t)c + remaining > x) keep++;
        }
        /* minimizer_table_capacity_for_items (src/lib.rs:4164-4169) */
        size_t c = (keep * 10 + 6) / 7;
        cap = c < 64 ? 64 : c;
    }
    dstab_init(t, cap);
    for (size_t i = 0; i < old.cap; i++) {
        if (!old.states[i]) continue;
        if (filter) {
            uint32_t c = old.states[i] & COUNT_MASK;
            if ((uint64_t)c + remaining > x) dstab_insert_state(t, old.keys[i], c);
        } else {
            dstab_insert_state(t, old.keys[i], old.states[i]);
        }
    }
    free(old.keys);
    free(old.states);
}
static void dstab_add_seen(dstab *t, uint32_t key, int accept_new) {
    if ((t->items + 1) * 10 >= t->cap * 7) dstab_rebuild(t, t->cap * 2, 0, 0, 0);
    size_t mask = t->cap - 1, i = key & mask;
    for (;;) {
        uint32_t st = t->states[i];
        if (st == 0) {
            if (accept_new) {
                t->keys[i] = key;
                t->states[i] = 1u | SEEN_BIT;
                t->items++;
            }
            return;
        }
        if (t->keys[i] == key) {
            if (!(st & SEEN_BIT)) {
                uint32_t c = (st & COUNT_MASK) + 1u;
                if (c > COUNT_MASK) c = COUNT_MASK;
                t->states[i] = c | SEEN_BIT;
            }
            return;
        }
        i = (i + 1) & mask;
    }
}
static int dstab_above(const dstab *t, uint32_t key, uint64_t x) {
    size_t mask = t->cap - 1, i = key & mask;
    for (;;) {
        uint32_t st = t->states[i];
        if (st == 0) return 0;
        if (t->keys[i] == key) return (uint64_t)(st & COUNT_MASK) > x;
        i = (i + 1) & mask;
    }
}

int or_count_datasets(const or_frags *f, size_t D, int k, int m, uint64_t x, size_t P,
                      or_counted **out, size_t *n_out, uint64_t *kept) {
    if (k < 1 || k > 32 || m < 1 || m > k || P == 0 || x >= COUNT_MASK || (k % 2) == 0 ||
        D == 0 || !f->dataset)
        return -1;
    *out = NULL;
    *n_out = 0;
    if (kept) *kept = 0;
    if (x >= D) return 0; /* src/lib.rs:644-648 */
    size_t mx = max_frag_len(f) + 1;
    run_t *runs = (run_t *)xmalloc(mx * sizeof(run_t));
    uint32_t *th = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    uint32_t *ts = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    /* Phase 1: one shard table per partition, datasets in order. */
    dstab *shards = (dstab *)xmalloc(P * sizeof(dstab));
    for (size_t s = 0; s < P; s++) dstab_init(&shards[s], 64);
    for (size_t d = 0; d < D; d++) {
        size_t remaining = D - d - 1;
        int accept_new = (1 + (uint64_t)remaining) > x; /* src/lib.rs:661 */
        for (size_t i = 0; i < f->nfrag; i++) {
            if (f->dataset[i] != d) continue;
            const uint8_t *c = f->codes + f->off[i];
            size_t n = f->off[i + 1] - f->off[i];
            if (n < (size_t)k) continue;
            size_t r = fragment_runs(c, n, k, m, runs, th, ts);
            for (size_t j = 0; j < r; j++)
                dstab_add_seen(&shards[or_spread_hash_u32(runs[j].key) % P], runs[j].key,
                               accept_new);
        }
        for (size_t s = 0; s < P; s++) dstab_rebuild(&shards[s], 0, 1, remaining, x);
    }
    uint64_t retained = 0;
    for (size_t s = 0; s < P; s++) retained += shards[s].items;
    if (kept) *kept = retained;
    /* Phase 2: per dataset, k-mers of retained runs, dedup, +1. */
    map128 pres;
    map_init(&pres, 1024);
    if (retained) {
        uint64_t *dk = NULL;
        size_t dn = 0, dcap = 0;
        u128 *cw = (u128 *)xmalloc(mx * sizeof(u128));
        for (size_t d = 0; d < D; d++) {
            dn = 0;
            for (size_t i = 0; i < f->nfrag; i++) {
                if (f->dataset[i] != d) continue;
                const uint8_t *c = f->codes + f->off[i];
                size_t n = f->off[i + 1] - f->off[i];
                if (n < (size_t)k) continue;
                size_t r = fragment_runs(c, n, k, m, runs, th, ts);
                canon_windows(c, n, k, cw); /* canonical_kmers_u64 (src/lib.rs:3671-3692) */
                for (size_t j = 0; j < r; j++) {
                    size_t shard = or_spread_hash_u32(runs[j].key) % P;
                    if (!dstab_above(&shards[shard], runs[j].key, x)) continue;
                    for (size_t w = runs[j].start; w < runs[j].end; w++) {
                        if (dn == dcap) {
                            dcap = dcap ? dcap * 2 : 1024;
                            dk = (uint64_t *)xrealloc(dk, dcap * sizeof(uint64_t));
                        }
                        dk[dn++] = (uint64_t)cw[w];
                    }
                }
            }
            /* sort + dedup per dataset (src/lib.rs:2242-2244); the per-partition
             * split by spread_hash_u64 does not change the per-key result. */
            qsort(dk, dn, sizeof(uint64_t), cmp_u64);
            for (size_t i = 0; i < dn; i++)
                if (i == 0 || dk[i] != dk[i - 1]) (*map_slot(&pres, (u128)dk[i], 1))++;
        }
        free(dk);
        free(cw);
    }
    or_counted *res = NULL;
    size_t r = 0, rcap = 0;
    for (size_t i = 0; i < pres.cap; i++) {
        if (pres.state[i] != 1 || !(pres.vals[i] > x)) continue;
        if (r == rcap) {
            rcap = rcap ? rcap * 2 : 1024;
            res = (or_counted *)xrealloc(res, rcap * sizeof(or_counted));
        }
        res[r].lo = (uint64_t)pres.keys[i];
        res[r].hi = 0;
        res[r].count = pres.vals[i];
        r++;
    }
    map_free(&pres);
    for (size_t s = 0; s < P; s++) {
        free(shards[s].keys);
        free(shards[s].states);
    }
    free(shards);
    free(runs);
    free(th);
    free(ts);
    if (r) qsort(res, r, sizeof(or_counted), cmp_counted);
    *out = res;
    *n_out = r;
    return 0;
}

/* The same dataset-presence count on `threads` host threads, as the
 * reference runs it under Rayon: datasets in order (the accept-new and
 * finish steps depend on it, src/lib.rs:659-665), the reads of a dataset in
 * parallel. Phase 1: each thread computes the run keys of its share of the
 * dataset's fragments, split by shard; thread t then applies the keys of
 * shards s = t (mod T) (SEEN-guarded, so the order of application does not
 * change a state) and finishes those shards. Phase 2: each thread collects
 * the k-mers of retained runs of its fragments into T buckets by
 * spread_hash_u64(k-mer) (the reference's partitions, src/lib.rs:2931-2933);
 * thread t sorts and deduplicates bucket t of the dataset and adds 1 to its
 * own presence map. Same output as or_count_datasets. */
typedef struct ds_job_s {
    const or_frags *f;
    int k, m;
    uint64_t x;
    size_t P;
    int t, T;
    const size_t *frag;   /* the dataset's fragments (length >= k) */
    size_t lo, hi;        /* this thread's share of them */
    int accept_new;
    size_t remaining;
    dstab *shards;
    struct ds_job_s *all;
    uint32_t **keys;      /* phase 1: per shard group u = shard % T */
    size_t *nkeys, *capkeys;
    uint64_t **bk;        /* phase 2: per bucket */
    size_t *bn, *bcap;
    map128 pres;          /* phase 2: presence of the k-mers of bucket t */
} ds_job;

static void push_u32(uint32_t **a, size_t *n, size_t *cap, uint32_t v) {
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 1024;
        *a = (uint32_t *)xrealloc(*a, *cap * sizeof(uint32_t));
    }
    (*a)[(*n)++] = v;
}
static void push_u64(uint64_t **a, size_t *n, size_t *cap, uint64_t v) {
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 1024;
        *a = (uint64_t *)xrealloc(*a, *cap * sizeof(uint64_t));
    }
    (*a)[(*n)++] = v;
}

static void *ds_runs(void *arg) {
    ds_job *j = (ds_job *)arg;
    size_t mx = max_frag_len(j->f) + 1;
    run_t *runs = (run_t *)xmalloc(mx * sizeof(run_t));
    uint32_t *th = (uint32_t *)xmalloc(mx * sizeof(uint32_t)), *ts = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    for (int u = 0; u < j->T; u++) j->nkeys[u] = 0;
    for (size_t q = j->lo; q < j->hi; q++) {
        size_t i = j->frag[q];
        const uint8_t *c = j->f->codes + j->f->off[i];
        size_t n = j->f->off[i + 1] - j->f->off[i];
        size_t r = fragment_runs(c, n, j->k, j->m, runs, th, ts);
        for (size_t w = 0; w < r; w++) {
            size_t u = (or_spread_hash_u32(runs[w].key) % j->P) % (size_t)j->T;
            push_u32(&j->keys[u], &j->nkeys[u], &j->capkeys[u], runs[w].key);
        }
    }
    free(runs);
    free(th);
    free(ts);
    return NULL;
}

static void *ds_apply(void *arg) {
    ds_job *j = (ds_job *)arg;
    const ds_job *all = (const ds_job *)j->all;
    for (int v = 0; v < j->T; v++)
        for (size_t i = 0; i < all[v].nkeys[j->t]; i++) {
            uint32_t key = all[v].keys[j->t][i];
            dstab_add_seen(&j->shards[or_spread_hash_u32(key) % j->P], key, j->accept_new);
        }
    for (size_t s = (size_t)j->t; s < j->P; s += (size_t)j->T)
        dstab_rebuild(&j->shards[s], 0, 1, j->remaining, j->x);
    return NULL;
}

static void *ds_kmers(void *arg) {
    ds_job *j = (ds_job *)arg;
    size_t mx = max_frag_len(j->f) + 1;
    run_t *runs = (run_t *)xmalloc(mx * sizeof(run_t));
    uint32_t *th = (uint32_t *)xmalloc(mx * sizeof(uint32_t)), *ts = (uint32_t *)xmalloc(mx * sizeof(uint32_t));
    u128 *cw = (u128 *)xmalloc(mx * sizeof(u128));
    for (int u = 0; u < j->T; u++) j->bn[u] = 0;
    for (size_t q = j->lo; q < j->hi; q++) {
        size_t i = j->frag[q];
        const uint8_t *c = j->f->codes + j->f->off[i];
        size_t n = j->f->off[i + 1] - j->f->off[i];
        size_t r = fragment_runs(c, n, j->k, j->m, runs, th, ts);
        canon_windows(c, n, j->k, cw);
        for (size_t w = 0; w < r; w++) {
            if (!dstab_above(&j->shards[or_spread_hash_u32(runs[w].key) % j->P], runs[w].key, j->x)) continue;
            for (size_t p = runs[w].start; p < runs[w].end; p++) {
                uint64_t km = (uint64_t)cw[p];
                size_t u = (size_t)(or_spread_hash_u64(km) % (uint64_t)j->T);
                push_u64(&j->bk[u], &j->bn[u], &j->bcap[u], km);
            }
        }
    }
    free(runs);
    free(th);
    free(ts);
    free(cw);
    return NULL;
}

static void *ds_presence(void *arg) {
    ds_job *j = (ds_job *)arg;
    const ds_job *all = (const ds_job *)j->all;
    size_t n = 0;
    for (int v = 0; v < j->T; v++) n += all[v].bn[j->t];
    if (!n) return NULL;
    uint64_t *a = (uint64_t *)xmalloc(n * sizeof(uint64_t));
    n = 0;
    for (int v = 0; v < j->T; v++) {
        memcpy(a + n, all[v].bk[j->t], all[v].bn[j->t] * sizeof(uint64_t));
        n += all[v].bn[j->t];
    }
    qsort(a, n, sizeof(uint64_t), cmp_u64);
    for (size_t i = 0; i < n; i++)
        if (i == 0 || a[i] != a[i - 1]) (*map_slot(&j->pres, (u128)a[i], 1))++;
    free(a);
    return NULL;
}

int or_count_datasets_mt(const or_frags *f, size_t D, int k, int m, uint64_t x, size_t P, int threads,
                         or_counted **out, size_t *n_out, uint64_t *kept) {
    if (k < 1 || k > 32 || m < 1 || m > k || P == 0 || x >= COUNT_MASK || (k % 2) == 0 || D == 0 ||
        !f->dataset || threads < 1)
        return -1;
    *out = NULL;
    *n_out = 0;
    if (kept) *kept = 0;
    if (x >= D) return 0; /* src/lib.rs:644-648 */
    const int T = threads;
    /* the fragments (length >= k) of each dataset, in input order */
    size_t *cnt = (size_t *)xcalloc(D + 1, sizeof(size_t));
    for (size_t i = 0; i < f->nfrag; i++)
        if (f->dataset[i] < D && f->off[i + 1] - f->off[i] >= (uint64_t)k) cnt[f->dataset[i] + 1]++;
    for (size_t d = 0; d < D; d++) cnt[d + 1] += cnt[d];
    size_t *frag = (size_t *)xmalloc((cnt[D] ? cnt[D] : 1) * sizeof(size_t));
    size_t *fill = (size_t *)xmalloc((D ? D : 1) * sizeof(size_t));
    memcpy(fill, cnt, D * sizeof(size_t));
    for (size_t i = 0; i < f->nfrag; i++)
        if (f->dataset[i] < D && f->off[i + 1] - f->off[i] >= (uint64_t)k) frag[fill[f->dataset[i]]++] = i;
    free(fill);
    dstab *shards = (dstab *)xmalloc(P * sizeof(dstab));
    for (size_t s = 0; s < P; s++) dstab_init(&shards[s], 64);
    ds_job *jobs = (ds_job *)xcalloc((size_t)T, sizeof(ds_job));
    for (int t = 0; t < T; t++) {
        ds_job *j = &jobs[t];
        j->f = f;
        j->k = k;
        j->m = m;
        j->x = x;
        j->P = P;
        j->t = t;
        j->T = T;
        j->shards = shards;
        j->all = jobs;
        j->keys = (uint32_t **)xcalloc((size_t)T, sizeof(uint32_t *));
        j->nkeys = (size_t *)xcalloc((size_t)T, sizeof(size_t));
        j->capkeys = (size_t *)xcalloc((size_t)T, sizeof(size_t));
        j->bk = (uint64_t **)xcalloc((size_t)T, sizeof(uint64_t *));
        j->bn = (size_t *)xcalloc((size_t)T, sizeof(size_t));
        j->bcap = (size_t *)xcalloc((size_t)T, sizeof(size_t));
        map_init(&j->pres, 1024);
    }
    /* a dataset's fragments split among the threads */
    #define DS_SPLIT(d)                                                          \
        for (int t = 0; t < T; t++) {                                            \
            size_t nd = cnt[(d) + 1] - cnt[(d)];                                 \
            jobs[t].frag = frag + cnt[(d)];                                      \
            jobs[t].lo = nd * (size_t)t / (size_t)T;                             \
            jobs[t].hi = nd * (size_t)(t + 1) / (size_t)T;                       \
        }
    for (size_t d = 0; d < D; d++) {
        DS_SPLIT(d)
        for (int t = 0; t < T; t++) {
            jobs[t].remaining = D - d - 1;
            jobs[t].accept_new = (1 + (uint64_t)jobs[t].remaining) > x; /* src/lib.rs:661 */
        }
        pthread_t *th = (pthread_t *)xmalloc(sizeof(pthread_t) * (size_t)T);
        for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, ds_runs, &jobs[t]);
        for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
        for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, ds_apply, &jobs[t]);
        for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
        free(th);
    }
    uint64_t retained = 0;
    for (size_t s = 0; s < P; s++) retained += shards[s].items;
    if (kept) *kept = retained;
    if (retained) {
        for (size_t d = 0; d < D; d++) {
            DS_SPLIT(d)
            pthread_t *th = (pthread_t *)xmalloc(sizeof(pthread_t) * (size_t)T);
            for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, ds_kmers, &jobs[t]);
            for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
            for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, ds_presence, &jobs[t]);
            for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
            free(th);
        }
    }
    #undef DS_SPLIT
    or_counted *res = NULL;
    size_t r = 0, rcap = 0;
    for (int t = 0; t < T; t++) {
        const map128 *pm = &jobs[t].pres;
        for (size_t i = 0; i < pm->cap; i++) {
            if (pm->state[i] != 1 || !(pm->vals[i] > x)) continue;
            if (r == rcap) {
                rcap = rcap ? rcap * 2 : 1024;
                res = (or_counted *)xrealloc(res, rcap * sizeof(or_counted));
            }
            res[r].lo = (uint64_t)pm->keys[i];
            res[r].hi = 0;
            res[r].count = pm->vals[i];
            r++;
        }
    }
    for (int t = 0; t < T; t++) {
        ds_job *j = &jobs[t];
        for (int u = 0; u < T; u++) {
            free(j->keys[u]);
            free(j->bk[u]);
        }
        free(j->keys);
        free(j->nkeys);
        free(j->capkeys);
        free(j->bk);
        free(j->bn);
        free(j->bcap);
        map_free(&j->pres);
    }
    free(jobs);
    for (size_t s = 0; s < P; s++) {
        free(shards[s].keys);
        free(shards[s].states);
    }
    free(shards);
    free(frag);
    free(cnt);
    if (r) qsort(res, r, sizeof(or_counted), cmp_counted);
    *out = res;
    *n_out = r;
    return 0;
}

int or_count_datasets_bruteforce(const or_frags *f, size_t D, int k, uint64_t x,
                                 or_counted **out, size_t *n_out) {
    map128 pres;
    map_init(&pres, 1024);
    size_t mx = max_frag_len(f) + 1;
    u128 *cw = (u128 *)xmalloc(mx * sizeof(u128));
    for (size_t d = 0; d < D; d++) {
        map128 seen;
        map_init(&seen, 1024);
        for (size_t i = 0; i < f->nfrag; i++) {
            if (f->dataset[i] != d) continue;
            size_t n = canon_windows(f->codes + f->off[i], f->off[i + 1] - f->off[i], k, cw);
            for (size_t w = 0; w < n; w++) *map_slot(&seen, cw[w], 1) = 1;
        }
        for (size_t i = 0; i < seen.cap; i++)
            if (seen.state[i] == 1) (*map_slot(&pres, seen.keys[i], 1))++;
        map_free(&seen);
    }
    free(cw);
    or_counted *res = (or_counted *)xmalloc((pres.used + 1) * sizeof(or_counted));
    size_t r = 0;
    for (size_t i = 0; i < pres.cap; i++) {
        if (pres.state[i] != 1 || !(pres.vals[i] > x)) continue;
        res[r].lo = (uint64_t)pres.keys[i];
        res[r].hi = (uint64_t)(pres.keys[i] >> 64);
        res[r].count = pres.vals[i];
        r++;
    }
    map_free(&pres);
    if (r) qsort(res, r, sizeof(or_counted), cmp_counted);
    *out = res;
    *n_out = r;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ZI: absent = {canonical K of query windows} \ {canonical K of index}     */
/* (src/zi.rs:152-175, 224-238)                                             */
/* ------------------------------------------------------------------------ */

static size_t sort_unique(u128 *v, size_t n) {
    if (!n) return 0;
    qsort(v, n, sizeof(u128), cmp_u128);
    size_t o = 1;
    for (size_t i = 1; i < n; i++)
        if (v[i] != v[o - 1]) v[o++] = v[i];
    return o;
}

int or_zi_absent(const or_frags *index, const or_frags *query, int k, or_counted **out,
                 size_t *n_out) {
    size_t ni, nq;
    u128 *iv = all_canonical(index, k, &ni);
    u128 *qv = all_canonical(query, k, &nq);
    ni = sort_unique(iv, ni);
    nq = sort_unique(qv, nq);
    or_counted *res = (or_counted *)xmalloc((nq + 1) * sizeof(or_counted));
    size_t r = 0, a = 0;
    for (size_t i = 0; i < nq; i++) {
        while (a < ni && iv[a] < qv[i]) a++;
        if (a < ni && iv[a] == qv[i]) continue;
        res[r].lo = (uint64_t)qv[i];
        res[r].hi = (uint64_t)(qv[i] >> 64);
        res[r].count = 1;
        r++;
    }
    free(iv);
    free(qv);
    *out = res;
    *n_out = r;
    return 0;
}

/* ZI's query loop on one thread (collect_query_kmers_absent_from_index,
 * src/zi.rs:152-175): the index k-mers go into a hash set, then every query
 * window's canonical k-mer is looked up and the absent ones are inserted into
 * a second set. The two phases are timed apart (*build_s, *query_s); this is
 * the CPU baseline of the served lookups, and it returns the same absent set
 * size as or_zi_absent. */
static double or_now(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}
int or_zi_lookup_timed(const or_frags *index, const or_frags *query, int k, uint64_t *lookups,
                       uint64_t *absent_unique, double *build_s, double *query_s) {
    const double t0 = or_now();
    size_t mx = max_frag_len(index);
    if (max_frag_len(query) > mx) mx = max_frag_len(query);
    u128 *cw = (u128 *)xmalloc((mx + 1) * sizeof(u128));
    map128 idx;
    map_init(&idx, 1024);
    for (size_t i = 0; i < index->nfrag; i++) {
        size_t n = canon_windows(index->codes + index->off[i], index->off[i + 1] - index->off[i], k, cw);
        for (size_t w = 0; w < n; w++) *map_slot(&idx, cw[w], 1) = 1;
    }
    const double t1 = or_now();
    map128 absent;
    map_init(&absent, 1024);
    uint64_t nl = 0;
    for (size_t i = 0; i < query->nfrag; i++) {
        size_t n = canon_windows(query->codes + query->off[i], query->off[i + 1] - query->off[i], k, cw);
        nl += n;
        for (size_t w = 0; w < n; w++)
            if (!map_slot(&idx, cw[w], 0)) *map_slot(&absent, cw[w], 1) = 1;
    }
    const double t2 = or_now();
    *lookups = nl;
    *absent_unique = absent.used;
    *build_s = t1 - t0;
    *query_s = t2 - t1;
    map_free(&idx);
    map_free(&absent);
    free(cw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Writers                                                                  */
/* ------------------------------------------------------------------------ */

static const char MC_ASCII[4] = {'A', 'C', 'G', 'T'};
static inline uint32_t ascii_bits(char b) {
    switch (b) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    default: return 3;
    }
}

/* canonical_suffix_extension (src/lib.rs:1580-1602) */
static u128 suffix_ext(const char *seq, size_t len, int k, char base) {
    u128 fwd = 0, rev = 0, mask = kmask128(k);
    for (size_t i = len - (size_t)(k - 1); i < len; i++) {
        u128 b = ascii_bits(seq[i]);
        fwd = ((fwd << 2) | b) & mask;
        rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
    }
    u128 b = ascii_bits(base);
    fwd = ((fwd << 2) | b) & mask;
    rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
    return fwd < rev ? fwd : rev;
}
static void revcomp_inplace(char *s, size_t n) {
    for (size_t i = 0; i < n / 2; i++) {
        char a = s[i], b = s[n - 1 - i];
        s[i] = MC_ASCII[3 - ascii_bits(b)];
        s[n - 1 - i] = MC_ASCII[3 - ascii_bits(a)];
    }
    if (n % 2) s[n / 2] = MC_ASCII[3 - ascii_bits(s[n / 2])];
}
static void write_lines(FILE *fp, const char *s, size_t n) {
    for (size_t i = 0; i < n; i += 80) {
        size_t c = n - i < 80 ? n - i : 80;
        fwrite(s + i, 1, c, fp);
        fputc('\n', fp);
    }
}

/* for_each_simplitig_encoded (src/lib.rs:1478-1533) */
int or_write_simplitigs(const char *path, const or_counted *kmers, size_t n, int k,
                        size_t first_index, size_t *last_index) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    map128 rem;
    map_init(&rem, n + 1);
    u128 *seeds = (u128 *)xmalloc((n + 1) * sizeof(u128));
    for (size_t i = 0; i < n; i++) {
        seeds[i] = ((u128)kmers[i].hi << 64) | kmers[i].lo;
        *map_slot(&rem, seeds[i], 1) = 1;
    }
    size_t ns = sort_unique(seeds, n);
    size_t cap = 1024, idx = first_index;
    char *seq = (char *)xmalloc(cap);
    for (size_t s = 0; s < ns; s++) {
        if (!map_remove(&rem, seeds[s])) continue;
        size_t len = (size_t)k;
        if (len + 2 > cap) {
            cap = len * 2 + 2;
            seq = (char *)xrealloc(seq, cap);
        }
        for (int i = 0; i < k; i++) seq[i] = MC_ASCII[(seeds[s] >> (2 * (k - 1 - i))) & 3u];
        for (int pass = 0; pass < 2; pass++) {
            for (;;) { /* extend_simplitig_forward (src/lib.rs:1556-1578) */
                int ext = 0;
                for (int b = 0; b < 4; b++) {
                    if (map_remove(&rem, suffix_ext(seq, len, k, MC_ASCII[b]))) {
                        if (len + 2 > cap) {
                            cap *= 2;
                            seq = (char *)xrealloc(seq, cap);
                        }
                        seq[len++] = MC_ASCII[b];
                        ext = 1;
                        break;
                    }
                }
                if (!ext) break;
            }
            revcomp_inplace(seq, len);
        }
        idx++;
        fprintf(fp, ">MC_simplitig_%zu kmers=%zu\n", idx, len - (size_t)k + 1);
        write_lines(fp, seq, len);
    }
    free(seq);
    free(seeds);
    map_free(&rem);
    if (last_index) *last_index = idx;
    return fclose(fp) == 0 ? 0 : -1;
}

/* write_kmer_fasta_record_u128 (src/lib.rs:1535-1544) */
int or_write_kmer_fasta(const char *path, const or_counted *kmers, size_t n, int k) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    char buf[65];
    for (size_t i = 0; i < n; i++) {
        u128 e = ((u128)kmers[i].hi << 64) | kmers[i].lo;
        for (int j = 0; j < k; j++) buf[j] = MC_ASCII[(e >> (2 * (k - 1 - j))) & 3u];
        fprintf(fp, ">MC_kmer_%zu count=%llu\n", i + 1, (unsigned long long)kmers[i].count);
        write_lines(fp, buf, (size_t)k);
    }
    return fclose(fp) == 0 ? 0 : -1;
}

/* write_kff (src/lib.rs:1278-1293) and the u8/u32 variants */
int or_write_kff(const char *path, const or_counted *kmers, size_t n, int k, int m,
                 uint64_t x, int count_size, int ordered) {
    bbuf o = {0};
    kff_header(&o, k, m, x, n, ordered, (uint64_t)count_size);
    for (size_t i = 0; i < n; i++)
        kff_record(&o, ((u128)kmers[i].hi << 64) | kmers[i].lo, k, kmers[i].count, count_size);
    bb_push(&o, "KFF", 3);
    FILE *fp = fopen(path, "wb");
    if (!fp) {
        free(o.d);
        return -1;
    }
    fwrite(o.d, 1, o.n, fp);
    free(o.d);
    return fclose(fp) == 0 ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* FASTA / FASTQ input: headers and qualities ignored, split at non-ACGT,   */
/* fragments shorter than min_len dropped (src/lib.rs:3264-3308).           */
/* ------------------------------------------------------------------------ */

static int base_code(int ch) {
    switch (ch) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'T': case 't': return 2;
    case 'G': case 'g': return 3;
    default: return -1;
    }
}

int or_parse_fastx(const char *path, int min_len, uint8_t **codes_out, uint64_t **off_out,
                   size_t *nfrag_out) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    size_t ccap = 1 << 16, cn = 0, ocap = 1024, on = 0;
    uint8_t *codes = (uint8_t *)xmalloc(ccap);
    uint64_t *off = (uint64_t *)xmalloc(ocap * sizeof(uint64_t));
    off[on++] = 0;
    size_t frag_start = 0;
    int fastq = -1;
    char *line = NULL;
    size_t lcap = 0;
    ssize_t ll;
    long lineno = 0;
#define CLOSE_FRAG()                                                                   \
    do {                                                                               \
        if (cn - frag_start >= (size_t)min_len && cn > frag_start) {                   \
            if (on == ocap) {                                                          \
                ocap *= 2;                                                             \
                off = (uint64_t *)xrealloc(off, ocap * sizeof(uint64_t));              \
            }                                                                          \
            off[on++] = cn;                                                            \
            frag_start = cn;                                                           \
        } else {                                                                       \
            cn = frag_start;                                                           \
        }                                                                              \
    } while (0)
    int in_seq = 0; /* fastq: line role within record */
    while ((ll = getline(&line, &lcap, fp)) > 0) {
        while (ll > 0 && (line[ll - 1] == '\n' || line[ll - 1] == '\r')) line[--ll] = 0;
        if (fastq < 0) {
            if (ll == 0) continue;
            fastq = line[0] == '@';
        }
        if (fastq) {
            int role = (int)(lineno % 4);
            lineno++;
            if (role != 1) continue;
            (void)in_seq;
        } else {
            if (ll > 0 && line[0] == '>') {
                CLOSE_FRAG();
                continue;
            }
        }
        for (ssize_t i = 0; i < ll; i++) {
            int c = base_code((unsigned char)line[i]);
            if (c < 0) {
                CLOSE_FRAG();
                continue;
            }
            if (cn == ccap) {
                ccap *= 2;
                codes = (uint8_t *)xrealloc(codes, ccap);
            }
            codes[cn++] = (uint8_t)c;
        }
        if (fastq) CLOSE_FRAG();
    }
    CLOSE_FRAG();
#undef CLOSE_FRAG
    free(line);
    fclose(fp);
    *codes_out = codes;
    *off_out = off;
    *nfrag_out = on - 1;
    return 0;
}

/* ---- full-size spot checks ---------------------------------------------- */
typedef struct {
    const uint32_t *words;
    const uint64_t *off;
    const uint32_t *dataset;
    size_t f0, f1;
    int k;
    const uint64_t *tkey;
    const int64_t *tidx;
    uint64_t tmask;
    uint64_t *counts;
    uint32_t *masks;
} spot_job;

static void *spot_worker(void *arg) {
    spot_job *j = (spot_job *)arg;
    const int k = j->k;
    const uint64_t kmask = k >= 32 ? ~0ull : ((1ull << (2 * k)) - 1ull);
    const int top = 2 * (k - 1);
    for (size_t f = j->f0; f < j->f1; f++) {
        const uint64_t s = j->off[f], e = j->off[f + 1];
        if (e - s < (uint64_t)k) continue;
        const uint32_t ds = j->dataset ? j->dataset[f] : 0;
        uint64_t fwd = 0, rev = 0;
        for (uint64_t p = s; p < e; p++) {
            const uint32_t code = (j->words[p >> 4] >> ((p & 15) * 2)) & 3u;
            const uint64_t b = code ^ (code >> 1); /* MC bits */
            fwd = ((fwd << 2) | b) & kmask;
            rev = (rev >> 2) | ((b ^ 3u) << top);
            if (p + 1 - s < (uint64_t)k) continue;
            const uint64_t key = fwd < rev ? fwd : rev;
            uint64_t h = or_spread_hash_u64(key) & j->tmask;
            while (j->tidx[h] >= 0) {
                if (j->tkey[h] == key) {
                    j->counts[j->tidx[h]]++;
                    if (j->masks) j->masks[j->tidx[h]] |= 1u << (ds & 31u);
                    break;
                }
                h = (h + 1) & j->tmask;
            }
        }
    }
    return NULL;
}

/* The same exact occurrence count for 33 <= k <= 64 (u128 canonical k-mers,
 * queries given as lo / hi words), over fragments split among threads. */
typedef struct {
    const uint32_t *words;
    const uint64_t *off;
    size_t f0, f1;
    int k;
    const u128 *tkey;
    const int64_t *tidx;
    uint64_t tmask;
    uint64_t *counts;
} spot128_job;

static void *spot128_worker(void *arg) {
    spot128_job *j = (spot128_job *)arg;
    const int k = j->k;
    const u128 mask = kmask128(k);
    for (size_t f = j->f0; f < j->f1; f++) {
        const uint64_t s = j->off[f], e = j->off[f + 1];
        if (e - s < (uint64_t)k) continue;
        u128 fwd = 0, rev = 0;
        for (uint64_t p = s; p < e; p++) {
            const uint32_t code = (j->words[p >> 4] >> ((p & 15) * 2)) & 3u;
            const u128 b = (u128)(code ^ (code >> 1));
            fwd = ((fwd << 2) | b) & mask;
            rev = (rev >> 2) | ((b ^ 3u) << (2 * (k - 1)));
            if (p + 1 - s < (uint64_t)k) continue;
            const u128 key = fwd < rev ? fwd : rev;
            uint64_t h = h128(key) & j->tmask;
            while (j->tidx[h] >= 0) {
                if (j->tkey[h] == key) {
                    j->counts[j->tidx[h]]++;
                    break;
                }
                h = (h + 1) & j->tmask;
            }
        }
    }
    return NULL;
}

int or_spot_counts128(const uint32_t *words, const uint64_t *off, size_t n_frag, int k, const uint64_t *qlo,
                      const uint64_t *qhi, size_t n_query, uint64_t *counts, int threads) {
    if (k < 1 || k > 64 || threads < 1) return 1;
    uint64_t slots = 16;
    while (slots < 4 * (uint64_t)n_query) slots <<= 1;
    u128 *tkey = (u128 *)xmalloc(slots * sizeof(u128));
    int64_t *tidx = (int64_t *)xmalloc(slots * 8);
    for (uint64_t i = 0; i < slots; i++) tidx[i] = -1;
    for (size_t q = 0; q < n_query; q++) {
        const u128 key = ((u128)qhi[q] << 64) | qlo[q];
        uint64_t h = h128(key) & (slots - 1);
        while (tidx[h] >= 0 && tkey[h] != key) h = (h + 1) & (slots - 1);
        if (tidx[h] < 0) {
            tkey[h] = key;
            tidx[h] = (int64_t)q;
        }
    }
    spot128_job *jobs = (spot128_job *)xmalloc(sizeof(spot128_job) * (size_t)threads);
    pthread_t *th = (pthread_t *)xmalloc(sizeof(pthread_t) * (size_t)threads);
    const uint64_t total = off[n_frag] - off[0];
    size_t f = 0;
    for (int t = 0; t < threads; t++) {
        spot128_job *j = &jobs[t];
        j->words = words; j->off = off; j->k = k;
        j->tkey = tkey; j->tidx = tidx; j->tmask = slots - 1;
        j->counts = (uint64_t *)calloc(n_query ? n_query : 1, 8);
        j->f0 = f;
        const uint64_t goal = off[0] + total / (uint64_t)threads * (uint64_t)(t + 1);
        while (f < n_frag && (t == threads - 1 || off[f] < goal)) f++;
        j->f1 = f;
        pthread_create(&th[t], NULL, spot128_worker, j);
    }
    for (size_t q = 0; q < n_query; q++) counts[q] = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        for (size_t q = 0; q < n_query; q++) counts[q] += jobs[t].counts[q];
        free(jobs[t].counts);
    }
    for (size_t q = 0; q < n_query; q++) {
        const u128 key = ((u128)qhi[q] << 64) | qlo[q];
        uint64_t h = h128(key) & (slots - 1);
        while (tkey[h] != key) h = (h + 1) & (slots - 1);
        counts[q] = counts[tidx[h]];
    }
    free(jobs);
    free(th);
    free(tkey);
    free(tidx);
    return 0;
}

int or_spot_counts(const uint32_t *words, const uint64_t *off, size_t n_frag, const uint32_t *dataset, int k,
                   const uint64_t *query, size_t n_query, uint64_t *counts, uint32_t *masks, int threads) {
    if (k < 1 || k > 32 || threads < 1) return 1;
    uint64_t slots = 16;
    while (slots < 4 * (uint64_t)n_query) slots <<= 1;
    uint64_t *tkey = (uint64_t *)xmalloc(slots * 8);
    int64_t *tidx = (int64_t *)xmalloc(slots * 8);
    for (uint64_t i = 0; i < slots; i++) tidx[i] = -1;
    for (size_t q = 0; q < n_query; q++) {
        uint64_t h = or_spread_hash_u64(query[q]) & (slots - 1);
        while (tidx[h] >= 0 && tkey[h] != query[q]) h = (h + 1) & (slots - 1);
        if (tidx[h] < 0) {
            tkey[h] = query[q];
            tidx[h] = (int64_t)q;
        }
    }
    spot_job *jobs = (spot_job *)xmalloc(sizeof(spot_job) * (size_t)threads);
    pthread_t *th = (pthread_t *)xmalloc(sizeof(pthread_t) * (size_t)threads);
    /* split fragments so every thread gets about the same number of bases */
    const uint64_t total = off[n_frag] - off[0];
    size_t f = 0;
    for (int t = 0; t < threads; t++) {
        spot_job *j = &jobs[t];
        j->words = words; j->off = off; j->dataset = dataset; j->k = k;
        j->tkey = tkey; j->tidx = tidx; j->tmask = slots - 1;
        j->counts = (uint64_t *)calloc(n_query ? n_query : 1, 8);
        j->masks = masks ? (uint32_t *)calloc(n_query ? n_query : 1, 4) : NULL;
        j->f0 = f;
        const uint64_t goal = off[0] + total / (uint64_t)threads * (uint64_t)(t + 1);
        while (f < n_frag && (t == threads - 1 || off[f] < goal)) f++;
        j->f1 = f;
        pthread_create(&th[t], NULL, spot_worker, j);
    }
    for (size_t q = 0; q < n_query; q++) {
        counts[q] = 0;
        if (masks) masks[q] = 0;
    }
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        for (size_t q = 0; q < n_query; q++) {
            counts[q] += jobs[t].counts[q];
            if (masks) masks[q] |= jobs[t].masks[q];
        }
        free(jobs[t].counts);
        free(jobs[t].masks);
    }
    /* duplicate queries share the first copy's totals */
    for (size_t q = 0; q < n_query; q++) {
        uint64_t h = or_spread_hash_u64(query[q]) & (slots - 1);
        while (tkey[h] != query[q]) h = (h + 1) & (slots - 1);
        counts[q] = counts[tidx[h]];
        if (masks) masks[q] = masks[tidx[h]];
    }
    free(jobs);
    free(th);
    free(tkey);
    free(tidx);
    return 0;
}
