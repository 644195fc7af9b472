// Simplitig emission: the greedy cover of for_each_simplitig_encoded
// (src/lib.rs:1478-1533), extending by canonical_suffix_extension
// (src/lib.rs:1580-1602), made fast without changing a byte of its output.
//
// The greedy visits seeds in ascending order; each seed grows forward, then
// backward, by the first of A, C, G, T whose canonical k-mer is still unused.
// Run literally that is one hash probe per tried base, in random order over
// the whole set. Here the set is first turned into its de Bruijn structure:
//  1. neighbours (parallel): for every side of every k-mer, the k-mers one
//     base away. A side with exactly one neighbour, whose facing side also
//     has exactly one, is a forced link: the greedy can only go that way.
//  2. chains (parallel): maximal runs of forced links, laid out one after
//     another, each k-mer at one position with its orientation in the run.
//  3. greedy (sequential): the same seeds and the same choices. Inside a
//     chain the next k-mer is the only candidate, so a walk takes every
//     position up to the first used one with one bit scan; only at chain ends
//     are the four bases tried against the set.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace mczi {

namespace {

typedef unsigned __int128 u128;
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint64_t LNONE = ~0ull;
// info byte per chain position: orientation, first and last base of the
// k-mer as read along the chain, chain start / end flags.
constexpr uint8_t I_REV = 1, I_START = 1 << 5, I_END = 1 << 6;

// Complement every base and reverse the order of the 32 bases in a word.
inline uint64_t revcomp_word(uint64_t x) {
    x = ~x;
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return __builtin_bswap64(x);
}

template <class T> struct Kmer;
template <> struct Kmer<uint64_t> {
    static uint64_t mask(uint32_t k) { return k >= 32 ? ~0ull : ((1ull << (2 * k)) - 1); }
    static uint64_t rc(uint64_t x, uint32_t k) { return revcomp_word(x) >> (64 - 2 * k); }
    static uint64_t of(const mczi_counted_kmer &r) { return r.lo; }
};
template <> struct Kmer<u128> {
    static u128 mask(uint32_t k) { return k >= 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1); }
    static u128 rc(u128 x, uint32_t k) {
        const u128 r = ((u128)revcomp_word((uint64_t)x) << 64) | revcomp_word((uint64_t)(x >> 64));
        return r >> (128 - 2 * k);
    }
    static u128 of(const mczi_counted_kmer &r) { return ((u128)r.hi << 64) | r.lo; }
};

double wall_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Allocator for the large arrays read at random: 2 MiB pages where the
// kernel grants them (transparent huge pages on request), so that random
// reads over gigabytes do not each miss the TLB.
template <class T> struct HugeAlloc {
    using value_type = T;
    HugeAlloc() = default;
    template <class U> HugeAlloc(const HugeAlloc<U> &) {}
    static constexpr size_t kHuge = size_t(1) << 21;
    T *allocate(size_t n) {
        const size_t bytes = n * sizeof(T);
        if (bytes < kHuge) return std::allocator<T>().allocate(n);
        void *p = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (p == MAP_FAILED) throw std::bad_alloc();
        madvise(p, bytes, MADV_HUGEPAGE);
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t n) {
        const size_t bytes = n * sizeof(T);
        if (bytes < kHuge)
            std::allocator<T>().deallocate(p, n);
        else
            munmap(p, bytes);
    }
    // resize(n) leaves new elements uninitialised: every array here is
    // written in full, or created with an explicit fill value
    template <class U> void construct(U *p) { ::new ((void *)p) U; }
    template <class U, class... A> void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
    template <class U> bool operator==(const HugeAlloc<U> &) const { return true; }
    template <class U> bool operator!=(const HugeAlloc<U> &) const { return false; }
};
template <class T> using big_vector = std::vector<T, HugeAlloc<T>>;

inline uint64_t ld(const uint64_t *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void st(uint64_t *p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
inline bool bit(const uint64_t *w, size_t i) { return (w[i >> 6] >> (i & 63)) & 1; }
inline void set_bit(uint64_t *w, size_t i) { w[i >> 6] |= 1ull << (i & 63); }

// Set bits [a, b] (inclusive).
inline void set_bits(uint64_t *w, size_t a, size_t b) {
    size_t wa = a >> 6, wb = b >> 6;
    const uint64_t ma = ~0ull << (a & 63), mb = ~0ull >> (63 - (b & 63));
    if (wa == wb) {
        w[wa] |= ma & mb;
        return;
    }
    w[wa] |= ma;
    for (size_t i = wa + 1; i < wb; i++) w[i] = ~0ull;
    w[wb] |= mb;
}

// Sorted unique k-mers with a directory on their top bits: lookup is one
// directory read and a short scan of one bucket.
template <class T> class SortedSet {
  public:
    SortedSet(const T *keys, size_t n, uint32_t k, unsigned nt) : keys_(keys), n_(n) {
        int d = 1;
        while (d < 30 && (size_t(1) << (d + 1)) <= n) d++;
        d = std::min(d, (int)(2 * k));
        shift_ = (int)(2 * k) - d;
        const size_t nb = size_t(1) << d;
        dir_.resize(nb + 1);
        // dir[b] = first index whose key has top bits >= b: entry b is
        // written by the one i with top(key[i-1]) < b <= top(key[i]).
        parallel_ranges(n + 1, 1 << 16, nt, [&](size_t b0, size_t b1) {
            for (size_t i = b0; i < b1; i++) {
                const int64_t lo = i == 0 ? -1 : (int64_t)top(keys_[i - 1]);
                const int64_t hi = i == n_ ? (int64_t)nb : (int64_t)top(keys_[i]);
                for (int64_t b = lo + 1; b <= hi; b++) dir_[(size_t)b] = (uint32_t)i;
            }
        });
    }
    uint32_t find(T key) const {
        const size_t b = (size_t)top(key);
        return find_in(key, dir_[b], dir_[b + 1]);
    }
    // The same in steps, for batches: the bucket's directory entry, then its
    // first key, so that callers can prefetch each before use.
    const uint32_t *dir_entry(T key) const { return &dir_[(size_t)top(key)]; }
    const T *key_at(uint32_t i) const { return keys_ + i; }
    uint32_t find_in(T key, uint32_t lo, uint32_t hi) const {
        if (hi - lo > 16) {
            lo = (uint32_t)(std::lower_bound(keys_ + lo, keys_ + hi, key) - keys_);
            return (lo < hi && keys_[lo] == key) ? lo : NONE;
        }
        for (; lo < hi; lo++) {
            const T v = keys_[lo];
            if (v >= key) return v == key ? lo : NONE;
        }
        return NONE;
    }

  private:
    uint64_t top(T x) const { return (uint64_t)(x >> shift_); }
    const T *keys_;
    size_t n_;
    int shift_;
    big_vector<uint32_t> dir_;
};

template <class T> class SimplitigCover {
  public:
    SimplitigCover(const mczi_counted_kmer *kmers, size_t n, uint32_t k, bool one_thread = false)
        : k_(k), mask_(Kmer<T>::mask(k)), hs_(2 * (int)k - 2), one_thread_(one_thread) {
        keys_.resize(n);
        nt_ = one_thread_ ? 1 : host_workers(n);
        parallel_ranges(n, 1 << 16, nt_, [&](size_t b0, size_t b1) {
            for (size_t i = b0; i < b1; i++) keys_[i] = Kmer<T>::of(kmers[i]) & mask_;
        });
        finish_keys();
    }
    // Keys as low / high 64-bit words, written by fill(lo, hi) (hi only for
    // k > 32).
    SimplitigCover(size_t n, uint32_t k, const std::function<void(uint64_t *, uint64_t *)> &fill)
        : k_(k), mask_(Kmer<T>::mask(k)), hs_(2 * (int)k - 2) {
        nt_ = host_workers(n);
        keys_.resize(n);
        if (sizeof(T) == 8) {
            fill(reinterpret_cast<uint64_t *>(keys_.data()), nullptr);
            parallel_ranges(n, 1 << 16, nt_, [&](size_t b0, size_t b1) {
                for (size_t i = b0; i < b1; i++) keys_[i] &= mask_;
            });
        } else {
            big_vector<uint64_t> lo(n), hi(n);
            fill(lo.data(), hi.data());
            parallel_ranges(n, 1 << 16, nt_, [&](size_t b0, size_t b1) {
                for (size_t i = b0; i < b1; i++) keys_[i] = (T)((((u128)hi[i] << 64) | lo[i]) & mask_);
            });
        }
        finish_keys();
        keys_as_given_ = n_ == n && !resorted_;
    }

  private:
    // Seeds are the keys in ascending order, each once (the set semantics of
    // the reference's HashSet).
    void finish_keys() {
        std::atomic<bool> sorted{true};
        parallel_ranges(keys_.size(), 1 << 20, nt_, [&](size_t b0, size_t b1) {
            for (size_t i = std::max<size_t>(b0, 1); i < b1; i++)
                if (keys_[i] < keys_[i - 1]) {
                    sorted = false;
                    return;
                }
        });
        resorted_ = !sorted;
        if (!sorted) std::sort(keys_.begin(), keys_.end());
        keys_.erase(std::unique(keys_.begin(), keys_.end()), keys_.end());
        n_ = keys_.size();
        require(n_ < (size_t)NONE, MCZI_E_INVALID,
                "too many k-mers for the simplitig writer (at most 2^32 - 2)");
        nt_ = one_thread_ ? 1 : host_workers(n_);
    }

  public:
    // link_fill, when given, writes the 2n link words (the device pass of
    // graph_kernels.hip) for keys that came sorted and unique.
    template <class Sink> void run(Sink &sink, const std::function<void(uint64_t *)> &link_fill = nullptr) {
        if (n_ == 0) return;
        const bool trace = getenv("MCZI_WRITER_TRACE") != nullptr;
        double t0 = wall_s();
        auto lap = [&](const char *what) {
            if (!trace) return;
            const double t = wall_s();
            fprintf(stderr, "simplitigs\t%s\t%.3f s\n", what, t - t0);
            t0 = t;
        };
        SortedSet<T> set(keys_.data(), n_, k_, nt_);
        lap("directory");
        big_vector<uint64_t> link;
        if (link_fill) {
            require(keys_as_given_, MCZI_E_INVALID, "device k-mers for the simplitig writer are not sorted and unique");
            link.resize(2 * n_);
            link_fill(link.data());
        } else {
            link = forced_links(set);
        }
        lap("links");
        lay_out_chains(link);
        big_vector<uint64_t>().swap(link);
        lap("chains");
        greedy(set, sink);
        lap("greedy");
    }

  private:
    // One link word per k-mer side, index 2i + o (o = 0: the k-mer as
    // stored, o = 1: its reverse complement). A forced link holds the
    // successor j, the base b it appends and its orientation o' as
    // j << 3 | b << 1 | o'; LNONE otherwise.
    static uint32_t l_node(uint64_t v) { return (uint32_t)(v >> 3); }
    static uint32_t l_base(uint64_t v) { return (uint32_t)(v >> 1) & 3; }
    // the side a walk leaves the successor by
    static size_t l_side(uint64_t v) { return (size_t)(v >> 3) * 2 + (size_t)(v & 1); }
    // sampled splitters: about one k-mer in 256 cuts chains into segments
    static bool sampled(uint32_t i) { return (uint32_t)(i * 2654435761u) >> 24 == 0; }

    // The successors of side o of k-mer i are canonical(suffix + b), b in
    // A, C, G, T. The side links to j when j != i is its only successor and
    // the side of j facing back has i as its only successor.
    big_vector<uint64_t> forced_links(const SortedSet<T> &set) {
        big_vector<uint64_t> link(2 * n_);
        // Batches of NB k-mers: all 8 * NB candidate lookups are issued
        // stage by stage (directory, then bucket) with prefetches, so the
        // cache misses of a batch overlap instead of following each other.
        constexpr int NB = 8, NC = 8 * NB;
        parallel_ranges(n_, 1 << 14, nt_, [&](size_t b0, size_t b1) {
            T cand[NC];
            const uint32_t *de[NC];
            uint32_t lo[NC], hi[NC];
            for (size_t i0 = b0; i0 < b1; i0 += NB) {
                const int nb_ = (int)std::min<size_t>(NB, b1 - i0);
                for (int q = 0; q < nb_; q++) {
                    const T c = keys_[i0 + q], rcc = Kmer<T>::rc(c, k_);
                    for (int o = 0; o < 2; o++) {
                        const T fb = ((o ? rcc : c) << 2) & mask_, rb = (o ? c : rcc) >> 2;
                        for (uint32_t b = 0; b < 4; b++) {
                            const T f = fb | b, r = ((T)(3 - b) << hs_) | rb;
                            const int e = q * 8 + o * 4 + (int)b;
                            cand[e] = r < f ? r : f;
                            de[e] = set.dir_entry(cand[e]);
                            __builtin_prefetch(de[e]);
                        }
                    }
                }
                for (int e = 0; e < nb_ * 8; e++) {
                    lo[e] = de[e][0];
                    hi[e] = de[e][1];
                    if (lo[e] < hi[e]) __builtin_prefetch(set.key_at(lo[e]));
                }
                for (int q = 0; q < nb_; q++) {
                    const size_t i = i0 + q;
                    const T c = keys_[i];
                    if (c == Kmer<T>::rc(c, k_)) {  // a palindrome (even k only) stays a chain of one
                        link[2 * i] = link[2 * i + 1] = LNONE;
                        continue;
                    }
                    for (int o = 0; o < 2; o++) {
                        uint32_t deg = 0;
                        uint64_t nb = LNONE;
                        for (uint32_t b = 0; b < 4; b++) {
                            const int e = q * 8 + o * 4 + (int)b;
                            const uint32_t j = set.find_in(cand[e], lo[e], hi[e]);
                            if (j != NONE) {
                                // the candidate read forward is suffix + b; it is
                                // stored reversed when its canonical form differs
                                const T f = (((o ? Kmer<T>::rc(c, k_) : c) << 2) & mask_) | b;
                                deg++;
                                nb = ((uint64_t)j << 3) | (b << 1) | (uint64_t)(cand[e] != f);
                            }
                        }
                        if (deg == 1 && (k_ & 1) == 0) {
                            const T t = keys_[l_node(nb)];
                            if (t == Kmer<T>::rc(t, k_)) deg = 0;
                        }
                        link[2 * i + o] = (deg == 1 && l_node(nb) != i) ? nb : LNONE;
                    }
                }
            }
        });
        // Keep x -> y only if y's side facing back also has one successor
        // (necessarily the reverse of x). That test is the same for x and for
        // the facing side, so clearing in place while other threads read
        // gives the same result.
        uint64_t *L = link.data();
        parallel_ranges(2 * n_, 1 << 15, nt_, [&](size_t b0, size_t b1) {
            for (size_t x = b0; x < b1; x++) {
                const uint64_t y = ld(&L[x]);
                if (y != LNONE && ld(&L[l_side(y) ^ 1]) == LNONE) st(&L[x], LNONE);
            }
        });
        return link;
    }

    uint8_t info_of(uint32_t node, uint32_t o) const {
        const T c = keys_[node];
        uint32_t f = (uint32_t)(c >> hs_) & 3, l = (uint32_t)c & 3;
        if (o) {
            const uint32_t t = 3 - l;
            l = 3 - f;
            f = t;
        }
        return (uint8_t)(o | f << 1 | l << 3);
    }

    // Chains: paths of forced links, each laid out from the end with the
    // smaller k-mer index; k-mers without links are chains of one; cycles of
    // forced links are cut at a splitter (or, with none, at their smallest
    // k-mer). List ranking over splitters keeps every walk short: splitters
    // are path ends and the sampled k-mers; (1) from each splitter side, walk
    // to the next splitter and count the k-mers between; (2) rank the chains
    // over their splitters; (3) walk each segment again, writing positions.
    void lay_out_chains(const big_vector<uint64_t> &link) {
        const uint64_t *L = link.data();
        auto split = [&](uint32_t i) {
            return L[2 * (size_t)i] == LNONE || L[2 * (size_t)i + 1] == LNONE || sampled(i);
        };
        // splitters in k-mer order, and each one's index
        big_vector<uint32_t> sid(n_, NONE);
        const size_t nblk = std::max<size_t>(1, std::min<size_t>(4 * nt_, n_ / 4096 + 1));
        const size_t bs = (n_ + nblk - 1) / nblk;
        std::vector<size_t> part(nblk + 1, 0);
        parallel_ranges(nblk, 1, nt_, [&](size_t b0, size_t b1) {
            for (size_t b = b0; b < b1; b++) {
                size_t c = 0;
                for (size_t i = b * bs; i < std::min(n_, (b + 1) * bs); i++) c += split((uint32_t)i);
                part[b + 1] = c;
            }
        });
        for (size_t b = 0; b < nblk; b++) part[b + 1] += part[b];
        const size_t ns = part[nblk];
        std::vector<uint32_t> spl(ns);
        parallel_ranges(nblk, 1, nt_, [&](size_t b0, size_t b1) {
            for (size_t b = b0; b < b1; b++) {
                size_t c = part[b];
                for (size_t i = b * bs; i < std::min(n_, (b + 1) * bs); i++)
                    if (split((uint32_t)i)) {
                        sid[i] = (uint32_t)c;
                        spl[c++] = (uint32_t)i;
                    }
            }
        });
        // (1) segments, 16 walks in flight per thread
        std::vector<uint64_t> seg_to(2 * ns, LNONE);
        std::vector<uint32_t> seg_len(2 * ns, 0);
        parallel_ranges(2 * ns, 1 << 10, nt_, [&](size_t s0, size_t s1) {
            constexpr int W = 16;
            uint64_t v[W];
            uint32_t m[W];
            size_t slot[W], next = s0;
            auto refill = [&](int w) {
                while (next < s1) {
                    const size_t sl = next++;
                    const uint64_t x = L[2 * (size_t)spl[sl >> 1] + (sl & 1)];
                    if (x == LNONE) continue;
                    v[w] = x;
                    m[w] = 0;
                    slot[w] = sl;
                    return true;
                }
                return false;
            };
            int act = 0;
            while (act < W && refill(act)) act++;
            while (act > 0) {
                for (int w = 0; w < act;) {
                    const uint32_t nd = l_node(v[w]);
                    const uint64_t a0 = L[2 * (size_t)nd], a1 = L[2 * (size_t)nd + 1];
                    if (a0 == LNONE || a1 == LNONE || sampled(nd)) {
                        seg_to[slot[w]] = ((uint64_t)sid[nd] << 1) | (v[w] & 1);
                        seg_len[slot[w]] = m[w];
                        if (!refill(w)) {
                            act--;
                            v[w] = v[act];
                            m[w] = m[act];
                            slot[w] = slot[act];
                            continue;
                        }
                    } else {
                        m[w]++;
                        v[w] = (v[w] & 1) ? a1 : a0;
                    }
                    w++;
                }
            }
        });
        big_vector<uint32_t>().swap(sid);
        // (2) chains over splitters: lengths from the path end that lays
        // each out, exclusive scan, then splitter and segment positions
        std::vector<size_t> clen(ns, 0);
        parallel_ranges(ns, 1 << 10, nt_, [&](size_t a0, size_t a1) {
            for (size_t a = a0; a < a1; a++) {
                const size_t s = spl[a];
                const bool v0 = L[2 * s] != LNONE, v1 = L[2 * s + 1] != LNONE;
                if (v0 && v1) continue;
                if (!v0 && !v1) {
                    clen[a] = 1;
                    continue;
                }
                size_t b = a, len = 1;
                uint32_t o = v0 ? 0 : 1;
                for (;;) {
                    const uint64_t to = seg_to[2 * b + o];
                    len += seg_len[2 * b + o] + 1;
                    b = (size_t)(to >> 1);
                    o = (uint32_t)(to & 1);
                    if (L[2 * (size_t)spl[b] + o] == LNONE) break;
                }
                if (spl[b] > s) clen[a] = len;
            }
        });
        size_t in_paths = 0;
        for (size_t a = 0; a < ns; a++) {
            const size_t l = clen[a];
            clen[a] = l ? in_paths : SIZE_MAX;
            in_paths += l;
        }
        pos_.assign(n_, NONE);
        info_.assign(n_, 0);
        std::vector<uint32_t> seg_pos(2 * ns, NONE);
        parallel_ranges(ns, 1 << 10, nt_, [&](size_t a0, size_t a1) {
            for (size_t a = a0; a < a1; a++) {
                if (clen[a] == SIZE_MAX) continue;
                const size_t p0 = clen[a], s = spl[a];
                size_t p = p0, b = a;
                uint32_t o = (L[2 * s] == LNONE && L[2 * s + 1] != LNONE) ? 1 : 0;
                for (;;) {
                    const uint32_t node = spl[b];
                    pos_[node] = (uint32_t)p;
                    info_[p] = info_of(node, o);
                    if (L[2 * (size_t)node + o] == LNONE) break;
                    const size_t sl = 2 * b + o;
                    seg_pos[sl] = (uint32_t)(p + 1);
                    p += seg_len[sl] + 1;
                    b = (size_t)(seg_to[sl] >> 1);
                    o = (uint32_t)(seg_to[sl] & 1);
                }
                info_[p0] |= I_START;
                info_[p] |= I_END;
            }
        });
        // splitters still unplaced lie on cycles: cut each at its first
        // splitter, walking it until a placed splitter comes back
        size_t next = in_paths;
        for (size_t a = 0; a < ns; a++) {
            if (pos_[spl[a]] != NONE) continue;
            const size_t p0 = next;
            size_t b = a;
            uint32_t o = 0;
            for (;;) {
                const uint32_t node = spl[b];
                pos_[node] = (uint32_t)next;
                info_[next] = info_of(node, o);
                const size_t sl = 2 * b + o;
                seg_pos[sl] = (uint32_t)(next + 1);
                next += seg_len[sl] + 1;
                b = (size_t)(seg_to[sl] >> 1);
                o = (uint32_t)(seg_to[sl] & 1);
                if (pos_[spl[b]] != NONE) break;
            }
            info_[p0] |= I_START;
            info_[next - 1] |= I_END;
        }
        // (3) the k-mers inside each placed segment
        parallel_ranges(2 * ns, 1 << 10, nt_, [&](size_t s0, size_t s1) {
            constexpr int W = 16;
            uint64_t v[W];
            T str[W];
            size_t p[W], next_sl = s0;
            auto refill = [&](int w) {
                while (next_sl < s1) {
                    const size_t sl = next_sl++;
                    if (seg_pos[sl] == NONE) continue;
                    const uint32_t node = spl[sl >> 1];
                    const uint32_t o = (uint32_t)(sl & 1);
                    const uint64_t x = L[2 * (size_t)node + o];
                    if (split(l_node(x))) continue;  // nothing between two splitters
                    v[w] = x;
                    str[w] = o ? Kmer<T>::rc(keys_[node], k_) : keys_[node];
                    p[w] = seg_pos[sl];
                    return true;
                }
                return false;
            };
            int act = 0;
            while (act < W && refill(act)) act++;
            while (act > 0) {
                for (int w = 0; w < act;) {
                    const uint32_t nd = l_node(v[w]), b = l_base(v[w]);
                    const T s = ((str[w] << 2) | b) & mask_;
                    pos_[nd] = (uint32_t)p[w];
                    info_[p[w]] |= (uint8_t)((v[w] & 1) | ((uint32_t)(s >> hs_) & 3) << 1 | b << 3);
                    const uint64_t x = L[l_side(v[w])];
                    if (split(l_node(x))) {
                        if (!refill(w)) {
                            act--;
                            v[w] = v[act];
                            str[w] = str[act];
                            p[w] = p[act];
                            continue;
                        }
                    } else {
                        v[w] = x;
                        str[w] = s;
                        p[w]++;
                    }
                    w++;
                }
            }
        });
        // cycles with no sampled splitter, cut before their smallest k-mer
        for (size_t i = 0; i < n_ && next < n_; i++) {
            if (pos_[i] != NONE) continue;
            const size_t p0 = next;
            size_t x = 2 * i;
            for (;;) {
                pos_[x >> 1] = (uint32_t)next;
                info_[next] = info_of((uint32_t)(x >> 1), (uint32_t)(x & 1));
                next++;
                const uint64_t y = L[x];
                if (y == LNONE || pos_[l_node(y)] != NONE) break;
                x = l_side(y);
            }
            info_[p0] |= I_START;
            info_[next - 1] |= I_END;
        }
        require(next == n_ && (info_[0] & I_START) && (info_[n_ - 1] & I_END), MCZI_E_INVALID,
                "simplitig chain layout lost k-mers");
        const size_t nw = (n_ + 63) / 64;
        cstart_.assign(nw, 0);
        cend_.assign(nw, 0);
        used_.assign(nw, 0);
        parallel_ranges(nw, 1 << 12, nt_, [&](size_t w0, size_t w1) {
            for (size_t w = w0; w < w1; w++) {
                uint64_t s = 0, e = 0;
                for (size_t j = 0; j < 64 && w * 64 + j < n_; j++) {
                    const uint8_t v = info_[w * 64 + j];
                    s |= (uint64_t)((v & I_START) != 0) << j;
                    e |= (uint64_t)((v & I_END) != 0) << j;
                }
                cstart_[w] = s;
                cend_[w] = e;
            }
        });
    }
    // First position q >= a with used or chain-end set (one exists: the
    // chain holding a - 1 ends at or after a).
    size_t scan_up(size_t a) const {
        size_t w = a >> 6;
        uint64_t m = (used_[w] | cend_[w]) & (~0ull << (a & 63));
        while (!m) {
            w++;
            m = used_[w] | cend_[w];
        }
        return (w << 6) + (size_t)__builtin_ctzll(m);
    }
    // Last position q <= a with used or chain-start set.
    size_t scan_down(size_t a) const {
        size_t w = a >> 6;
        uint64_t m = (used_[w] | cstart_[w]) & (~0ull >> (63 - (a & 63)));
        while (!m) {
            w--;
            m = used_[w] | cstart_[w];
        }
        return (w << 6) + 63 - (size_t)__builtin_clzll(m);
    }

    // Grow from position p walked in direction d (+1 along the chain, -1
    // against it); s is the current k-mer as read by the walk, r its reverse
    // complement. Appends the bases taken to out.
    void extend(const SortedSet<T> &set, size_t p, int d, T s, T r, std::vector<uint8_t> &out) {
        for (;;) {
            if (!bit(d > 0 ? cend_.data() : cstart_.data(), p)) {
                if (d > 0) {
                    const size_t q = scan_up(p + 1);
                    const bool blocked = bit(used_.data(), q);
                    const size_t last = blocked ? q - 1 : q;
                    if (last > p) {
                        set_bits(used_.data(), p + 1, last);
                        for (size_t t = p + 1; t <= last; t++) {
                            const uint32_t b = (info_[t] >> 3) & 3;
                            out.push_back((uint8_t)b);
                            s = ((s << 2) | b) & mask_;
                            r = (r >> 2) | ((T)(3 - b) << hs_);
                        }
                    }
                    if (blocked) return;
                    p = q;
                } else {
                    const size_t q = scan_down(p - 1);
                    const bool blocked = bit(used_.data(), q);
                    const size_t first = blocked ? q + 1 : q;
                    if (first < p) {
                        set_bits(used_.data(), first, p - 1);
                        for (size_t t = p; t-- > first;) {
                            const uint32_t b = 3 - ((info_[t] >> 1) & 3);
                            out.push_back((uint8_t)b);
                            s = ((s << 2) | b) & mask_;
                            r = (r >> 2) | ((T)(3 - b) << hs_);
                        }
                    }
                    if (blocked) return;
                    p = q;
                }
            }
            // chain end: the first unused of A, C, G, T
            const T fb = (s << 2) & mask_, rb = r >> 2;
            bool grown = false;
            for (uint32_t b = 0; b < 4; b++) {
                const T f = fb | b, rr = ((T)(3 - b) << hs_) | rb;
                const bool rev = rr < f;
                const uint32_t j = set.find(rev ? rr : f);
                if (j == NONE) continue;
                const uint32_t pj = pos_[j];
                if (bit(used_.data(), pj)) continue;
                set_bit(used_.data(), pj);
                out.push_back((uint8_t)b);
                s = f;
                r = rr;
                p = pj;
                d = ((uint32_t)rev == (info_[pj] & I_REV)) ? 1 : -1;
                grown = true;
                break;
            }
            if (!grown) return;
        }
    }

    // The greedy proper: each simplitig, as 2-bit codes (A0 C1 G2 T3), goes to
    // sink(seq) in the reference's order.
    template <class Sink> void greedy(const SortedSet<T> &set, Sink &sink) {
        std::vector<uint8_t> fwd, bwd, seq;
        fwd.reserve(1 << 16);
        bwd.reserve(1 << 16);
        seq.reserve(1 << 17);
        for (size_t i = 0; i < n_; i++) {
            const uint32_t p = pos_[i];
            if (bit(used_.data(), p)) continue;
            set_bit(used_.data(), p);
            const T c = keys_[i], rcc = Kmer<T>::rc(c, k_);
            const int o = info_[p] & I_REV;
            fwd.clear();
            bwd.clear();
            extend(set, p, o ? -1 : 1, c, rcc, fwd);
            extend(set, p, o ? 1 : -1, rcc, c, bwd);
            seq.resize(bwd.size() + k_ + fwd.size());
            size_t t = 0;
            for (size_t j = bwd.size(); j-- > 0;) seq[t++] = (uint8_t)(3 - bwd[j]);
            for (uint32_t j = 0; j < k_; j++) seq[t++] = (uint8_t)((uint32_t)(c >> (2 * (k_ - 1 - j))) & 3);
            if (!fwd.empty()) memcpy(&seq[t], fwd.data(), fwd.size());
            sink(seq);
        }
    }

    uint32_t k_;
    T mask_;
    int hs_;
    size_t n_ = 0;
    unsigned nt_ = 1;
    bool one_thread_ = false, resorted_ = false, keys_as_given_ = false;
    big_vector<T> keys_;
    big_vector<uint32_t> pos_;  // chain position of each k-mer (by sorted index)
    big_vector<uint8_t> info_;  // per chain position
    big_vector<uint64_t> cstart_, cend_, used_;
};

}  // namespace

// FASTA framing of write_simplitigs (src/lib.rs:1515-1533): ">MC_simplitig_i
// kmers=n", sequence in 80-column lines; text handed on in ~4 MiB chunks.
class FastaSink {
  public:
    FastaSink(uint32_t k, size_t first_index, const std::function<void(const char *, size_t)> &put)
        : k_(k), idx_(first_index), put_(put) {
        txt_.reserve((1 << 22) + (1 << 16));
    }
    void operator()(const std::vector<uint8_t> &seq) {
        static const char kAscii[4] = {'A', 'C', 'G', 'T'};
        char head[96];
        idx_++;
        const size_t len = seq.size();
        const int hl = snprintf(head, sizeof head, ">MC_simplitig_%zu kmers=%zu\n", idx_, len - k_ + 1);
        txt_.append(head, (size_t)hl);
        const size_t o0 = txt_.size();
        txt_.resize(o0 + len + (len + 79) / 80);
        char *w = &txt_[o0];
        for (size_t a = 0; a < len; a += 80) {
            const size_t m = std::min<size_t>(80, len - a);
            for (size_t j = 0; j < m; j++) w[j] = kAscii[seq[a + j]];
            w[m] = '\n';
            w += m + 1;
        }
        if (txt_.size() >= (1u << 22)) flush();
    }
    size_t finish() {
        flush();
        return idx_;
    }

  private:
    void flush() {
        if (!txt_.empty()) put_(txt_.data(), txt_.size());
        txt_.clear();
    }
    uint32_t k_;
    size_t idx_;
    const std::function<void(const char *, size_t)> &put_;
    std::string txt_;
};

// Each simplitig as an ASCII string (decode_kmer's letters, src/lib.rs:3056-3069).
struct StringSink {
    std::vector<std::string> *out;
    void operator()(const std::vector<uint8_t> &seq) {
        static const char kAscii[4] = {'A', 'C', 'G', 'T'};
        std::string t(seq.size(), 'A');
        for (size_t j = 0; j < seq.size(); j++) t[j] = kAscii[seq[j]];
        out->push_back(std::move(t));
    }
};

size_t emit_simplitigs_keys(size_t n, uint32_t k, const std::function<void(uint64_t *, uint64_t *)> &fill,
                            size_t first_index, const std::function<void(const char *, size_t)> &put,
                            const std::function<void(uint64_t *)> &link_fill) {
    require(k > 0, MCZI_E_INVALID, "k must be greater than 0");
    require(k <= 64, MCZI_E_INVALID, "k must be <= 64");
    FastaSink sink(k, first_index, put);
    if (k <= 32) {
        SimplitigCover<uint64_t> c(n, k, fill);
        c.run(sink, link_fill);
    } else {
        SimplitigCover<u128> c(n, k, fill);
        c.run(sink, link_fill);
    }
    return sink.finish();
}

size_t emit_simplitigs(const mczi_counted_kmer *kmers, size_t n, uint32_t k, size_t first_index,
                       const std::function<void(const char *, size_t)> &put) {
    require(k > 0, MCZI_E_INVALID, "k must be greater than 0");
    require(k <= 64, MCZI_E_INVALID, "k must be <= 64");
    FastaSink sink(k, first_index, put);
    if (k <= 32) {
        SimplitigCover<uint64_t> c(kmers, n, k);
        c.run(sink);
    } else {
        SimplitigCover<u128> c(kmers, n, k);
        c.run(sink);
    }
    return sink.finish();
}

namespace {

// minimizer_hash_u64 / minimizer_hash_u128 (src/lib.rs:2979-2994)
uint32_t minimizer_hash_u64(uint64_t v) {
    const uint64_t folded = v ^ ((v >> 17) | (v << 47)) ^ (v >> 32);
    return (uint32_t)folded * 0x9E3779B1u;
}
uint32_t minimizer_hash_u128(u128 v) {
    const uint64_t h = (uint64_t)(v >> 64);
    return minimizer_hash_u64((uint64_t)v ^ ((h << 23) | (h >> 41)));
}

// simplitig_partition (src/lib.rs:2935-2956): the smallest canonical m-mer
// of the k-mer, hashed, masked to a power-of-two partition count.
uint32_t simplitig_partition(u128 enc, uint32_t k, uint32_t m, uint32_t parts) {
    const u128 mask = Kmer<u128>::mask(m);
    u128 best = ~(u128)0;
    for (uint32_t start = 0; start + m <= k; start++) {
        const u128 v = (enc >> (2 * (k - start - m))) & mask;
        const u128 r = Kmer<u128>::rc(v, m);
        best = std::min(best, std::min(v, r));
    }
    return spread_hash_u32(minimizer_hash_u128(best)) & (parts - 1);
}

// simplitig_partition_count (src/lib.rs:2967-2977), with the host worker
// count in place of Rayon's thread count.
uint32_t simplitig_partition_count(uint32_t requested) {
    size_t want = requested;
    if (want == 0) {
        want = (size_t)host_workers(size_t(1) << 20) * 16;
        if (const char *e = getenv("MC_SIMPLITIG_PARTITIONS")) {
            char *end = nullptr;
            const unsigned long long v = strtoull(e, &end, 10);
            if (end != e && *end == 0) want = (size_t)v;
        }
    }
    want = std::min<size_t>(std::max<size_t>(want, 1), 1024);
    size_t p = 1;
    while (p < want) p <<= 1;
    return (uint32_t)std::min<size_t>(p, 1024);
}

}  // namespace

std::vector<std::string> simplitig_sequences_parallel(const mczi_counted_kmer *kmers, size_t n, uint32_t k,
                                                      uint32_t minimizer, uint32_t partition_count) {
    require(k > 0, MCZI_E_INVALID, "k must be greater than 0");
    require(k <= 64, MCZI_E_INVALID, "k must be <= 64");
    require(minimizer > 0 && minimizer <= k, MCZI_E_INVALID, "minimizer size must be > 0 and <= k");
    require(minimizer <= 64, MCZI_E_INVALID, "minimizer size must be <= 64 because MC stores minimizers in u128");
    const uint32_t parts = simplitig_partition_count(partition_count);
    const unsigned nt = host_workers(n);
    std::vector<uint32_t> part(n);
    parallel_ranges(n, 1 << 14, nt, [&](size_t b0, size_t b1) {
        for (size_t i = b0; i < b1; i++)
            part[i] = simplitig_partition(((u128)kmers[i].hi << 64) | kmers[i].lo, k, minimizer, parts);
    });
    std::vector<size_t> start(parts + 1, 0);
    for (size_t i = 0; i < n; i++) start[part[i] + 1]++;
    for (uint32_t q = 0; q < parts; q++) start[q + 1] += start[q];
    std::vector<mczi_counted_kmer> by_part(n);
    {
        std::vector<size_t> at(start.begin(), start.end() - 1);
        for (size_t i = 0; i < n; i++) by_part[at[part[i]]++] = kmers[i];
    }
    std::vector<std::vector<std::string>> per(parts);
    parallel_ranges(parts, 1, nt, [&](size_t q0, size_t q1) {
        for (size_t q = q0; q < q1; q++) {
            StringSink sink{&per[q]};
            const mczi_counted_kmer *b = by_part.data() + start[q];
            const size_t m = start[q + 1] - start[q];
            if (k <= 32) {
                SimplitigCover<uint64_t> c(b, m, k, true);
                c.run(sink);
            } else {
                SimplitigCover<u128> c(b, m, k, true);
                c.run(sink);
            }
        }
    });
    std::vector<std::string> out;
    for (auto &v : per)
        for (auto &t : v) out.push_back(std::move(t));
    return out;
}

std::vector<std::string> simplitig_sequences(const mczi_counted_kmer *kmers, size_t n, uint32_t k) {
    require(k > 0, MCZI_E_INVALID, "k must be greater than 0");
    require(k <= 64, MCZI_E_INVALID, "k must be <= 64");
    std::vector<std::string> out;
    StringSink sink{&out};
    if (k <= 32) {
        SimplitigCover<uint64_t> c(kmers, n, k);
        c.run(sink);
    } else {
        SimplitigCover<u128> c(kmers, n, k);
        c.run(sink);
    }
    return out;
}

}  // namespace mczi
