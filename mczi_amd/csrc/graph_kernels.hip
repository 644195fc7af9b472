// De Bruijn neighbours of the emitted k-mer set, on the device, for the
// simplitig writer (the greedy of src/lib.rs:1478-1533 walks exactly these
// edges: canonical_suffix_extension, src/lib.rs:1580-1602).
//
// Input: the sorted, unique canonical k-mers the count left in HBM. For every
// side of every k-mer (2i: the k-mer as stored, 2i + 1: its reverse
// complement) the four one-base extensions are looked up through a directory
// on the top bits of the key. A side is a forced link when it has exactly one
// successor j != i and j's side facing back has exactly one (necessarily i).
// The link word is j << 3 | appended base << 1 | orientation of j, or all
// ones: the same words the host pass in simplitigs.cpp computes.
//
// Both kernels are latency-bound gathers (about 3 dependent reads per lookup,
// 8 lookups per k-mer) over arrays of a few GB; one thread per k-mer keeps
// enough of them in flight.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace mczi {

namespace {

typedef unsigned __int128 u128;
constexpr unsigned long long LNONE = ~0ull;
constexpr uint32_t NOT_FOUND = 0xFFFFFFFFu;

// complement every base and reverse the order of the 32 bases in a word
__device__ __forceinline__ unsigned long long revcomp_word(unsigned long long x) {
    x = __builtin_bitreverse64(~x);
    return ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
}

template <bool WIDE> struct Key;
template <> struct Key<false> {
    typedef unsigned long long T;
    __device__ static T at(const unsigned long long *lo, const unsigned long long *, uint64_t i) { return lo[i]; }
    __device__ static T rc(T x, uint32_t k) { return revcomp_word(x) >> (64 - 2 * k); }
};
template <> struct Key<true> {
    typedef u128 T;
    __device__ static T at(const unsigned long long *lo, const unsigned long long *hi, uint64_t i) {
        return ((u128)hi[i] << 64) | lo[i];
    }
    __device__ static T rc(T x, uint32_t k) {
        const u128 r = ((u128)revcomp_word((unsigned long long)x) << 64) |
                       revcomp_word((unsigned long long)(x >> 64));
        return r >> (128 - 2 * k);
    }
};

// dir[b] = first index whose key has top bits >= b (b in [0, nb]); entry b is
// written by the one i with top(key[i-1]) < b <= top(key[i]).
template <bool WIDE>
__global__ void k_directory(const unsigned long long *__restrict__ lo, const unsigned long long *__restrict__ hi,
                            uint64_t n, int shift, uint64_t nb, uint32_t *__restrict__ dir) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t a = i == 0 ? -1 : (int64_t)(uint64_t)(Key<WIDE>::at(lo, hi, i - 1) >> shift);
        int64_t e = i == n ? (int64_t)nb : (int64_t)(uint64_t)(Key<WIDE>::at(lo, hi, i) >> shift);
        if (e > (int64_t)nb) e = (int64_t)nb;  // keys are < 4^k; never write past dir[nb]
        for (int64_t b = a + 1; b <= e; b++) dir[b] = (uint32_t)i;
    }
}

template <bool WIDE>
__device__ __forceinline__ uint32_t find_key(typename Key<WIDE>::T key, const unsigned long long *lo,
                                             const unsigned long long *hi, const uint32_t *dir, int shift) {
    typedef typename Key<WIDE>::T T;
    const uint64_t b = (uint64_t)(key >> shift);
    uint32_t a = dir[b];
    const uint32_t end = dir[b + 1];
    uint32_t e = end;
    while (a < e) {
        const uint32_t mid = a + ((e - a) >> 1);
        const T v = Key<WIDE>::at(lo, hi, mid);
        if (v < key)
            a = mid + 1;
        else
            e = mid;
    }
    return (a < end && Key<WIDE>::at(lo, hi, a) == key) ? a : NOT_FOUND;
}

template <bool WIDE>
__global__ void k_links(const unsigned long long *__restrict__ lo, const unsigned long long *__restrict__ hi,
                        uint64_t n, uint32_t k, const uint32_t *__restrict__ dir, int shift,
                        unsigned long long *__restrict__ link) {
    typedef typename Key<WIDE>::T T;
    const T mask = 2 * k >= 8 * sizeof(T) ? ~(T)0 : (((T)1 << (2 * k)) - 1);
    const int hs = 2 * (int)k - 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const T c = Key<WIDE>::at(lo, hi, i), rcc = Key<WIDE>::rc(c, k);
        if (c == rcc) {  // a palindrome (even k only) stays a chain of one
            link[2 * i] = link[2 * i + 1] = LNONE;
            continue;
        }
        for (int o = 0; o < 2; o++) {
            const T fb = ((o ? rcc : c) << 2) & mask, rb = (o ? c : rcc) >> 2;
            uint32_t deg = 0;
            unsigned long long nb = LNONE;
            for (uint32_t b = 0; b < 4; b++) {
                const T f = fb | b, r = ((T)(3 - b) << hs) | rb;
                const bool rev = r < f;
                const uint32_t j = find_key<WIDE>(rev ? r : f, lo, hi, dir, shift);
                if (j != NOT_FOUND) {
                    deg++;
                    nb = ((unsigned long long)j << 3) | (b << 1) | (unsigned long long)rev;
                }
            }
            if (deg == 1 && (k & 1) == 0) {
                const T t = Key<WIDE>::at(lo, hi, nb >> 3);
                if (t == Key<WIDE>::rc(t, k)) deg = 0;
            }
            link[2 * i + o] = (deg == 1 && (nb >> 3) != i) ? nb : LNONE;
        }
    }
}

// Keep x -> y only where y's side facing back (y's other side) is single too.
__global__ void k_links_facing(const unsigned long long *__restrict__ in, unsigned long long *__restrict__ out,
                               uint64_t m) {
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < m;
         x += (uint64_t)gridDim.x * blockDim.x) {
        unsigned long long y = in[x];
        if (y != LNONE && in[(((y >> 3) << 1) | (y & 1)) ^ 1] == LNONE) y = LNONE;
        out[x] = y;
    }
}

unsigned blocks_for(uint64_t n, unsigned threads) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + threads - 1) / threads, 1u << 20));
}

}  // namespace

void device_forced_links(const unsigned long long *lo, const unsigned long long *hi, uint64_t n, uint32_t k,
                         uint64_t *link_host) {
    require(k > 0 && k <= 64, MCZI_E_INVALID, "k must be greater than 0 and <= 64");
    require(n < NOT_FOUND, MCZI_E_INVALID, "too many k-mers for the simplitig writer (at most 2^32 - 2)");
    require(k <= 32 || hi != nullptr, MCZI_E_INVALID, "k > 32 needs the high k-mer words");
    if (n == 0) return;
    Device &d = Device::get();
    hipStream_t s = d.stream();
    int dbits = 1;
    while (dbits < 30 && (uint64_t(1) << (dbits + 1)) <= n) dbits++;
    dbits = std::min(dbits, (int)(2 * k));
    const int shift = (int)(2 * k) - dbits;
    const uint64_t nb = uint64_t(1) << dbits;
    DevBuf dir((nb + 1) * 4), la(2 * n * 8), lb(2 * n * 8);
    const bool wide = k > 32;
    const unsigned T = 256;
    if (wide) {
        hipLaunchKernelGGL(k_directory<true>, dim3(blocks_for(n + 1, T)), dim3(T), 0, s, lo, hi, n, shift, nb,
                           dir.as<uint32_t>());
        hipLaunchKernelGGL(k_links<true>, dim3(blocks_for(n, T)), dim3(T), 0, s, lo, hi, n, k, dir.as<uint32_t>(),
                           shift, la.as<unsigned long long>());
    } else {
        hipLaunchKernelGGL(k_directory<false>, dim3(blocks_for(n + 1, T)), dim3(T), 0, s, lo, hi, n, shift, nb,
                           dir.as<uint32_t>());
        hipLaunchKernelGGL(k_links<false>, dim3(blocks_for(n, T)), dim3(T), 0, s, lo, hi, n, k, dir.as<uint32_t>(),
                           shift, la.as<unsigned long long>());
    }
    hipLaunchKernelGGL(k_links_facing, dim3(blocks_for(2 * n, T)), dim3(T), 0, s, la.as<unsigned long long>(),
                       lb.as<unsigned long long>(), 2 * n);
    MCZI_HIP(hipGetLastError());
    d.download(link_host, lb.as<void>(), 2 * n * 8, s);
}

}  // namespace mczi
