// Device-side helpers shared by the MC and ZI kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mczi {

// ntHash seed values (low 32 bits), indexed by packed-seq code A0 C1 T2 G3;
// the complement of a code is code ^ 2. Rotation 7 per base, canonical hash
// = forward + reverse-complement (wrapping). Restates seq-hash's NtHasher<true>
// as used by vendor/simd-minimizers-2.3.1/src/minimizers.rs:258-291.
constexpr uint32_t NT_SEED0 = 0x95c60474u;
constexpr uint32_t NT_SEED1 = 0x62a02b4cu;
constexpr uint32_t NT_SEED2 = 0x82572324u;
constexpr uint32_t NT_SEED3 = 0x4be24456u;
constexpr uint32_t NT_ROT = 7u;

__host__ __device__ __forceinline__ uint32_t nt_seed(uint32_t c) {
    return (c & 2u) ? ((c & 1u) ? NT_SEED3 : NT_SEED2) : ((c & 1u) ? NT_SEED1 : NT_SEED0);
}

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, uint32_t r) {
    r &= 31u;
    return r ? (x << r) | (x >> (32u - r)) : x;
}
__host__ __device__ __forceinline__ uint32_t rotr32(uint32_t x, uint32_t r) {
    r &= 31u;
    return r ? (x >> r) | (x << (32u - r)) : x;
}

// spread_hash_u32 (src/lib.rs:2984-2990)
__host__ __device__ __forceinline__ uint32_t spread_hash_u32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    return v ^ (v >> 16);
}

// spread_hash_u64 (src/lib.rs:2996-3002); also used to place k-mers in tables.
__host__ __device__ __forceinline__ uint64_t spread_hash_u64(uint64_t v) {
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdull;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ull;
    return v ^ (v >> 33);
}

// 32 bases starting at base index `pos` of a packed word stream (16 bases per
// u32, base i at bits 2*(i%16)); base pos+j lands at bits 2j of the result.
// Reads words[pos/16 .. pos/16+2]; callers keep 3 words of slack.
template <typename W>
__device__ __forceinline__ uint64_t extract32_bases(const W *words, uint64_t pos) {
    uint64_t w = pos >> 4;
    uint32_t sh = (uint32_t)(pos & 15u) * 2u;
    uint64_t lo = (uint64_t)words[w] | ((uint64_t)words[w + 1] << 32);
    if (sh) lo = (lo >> sh) | ((uint64_t)words[w + 2] << (64u - sh));
    return lo;
}

__device__ __forceinline__ uint32_t base_code(const uint32_t *words, uint64_t pos) {
    return (words[pos >> 4] >> ((pos & 15u) * 2u)) & 3u;
}

// packed-seq code -> MC 2-bit encoding (A0 C1 G2 T3), src/lib.rs:3192-3197.
__host__ __device__ __forceinline__ uint32_t mc_bits(uint32_t code) { return code ^ (code >> 1); }

constexpr unsigned long long EMPTY_KEY = ~0ull;

// Finishing pass of a sort by the top key bits: the keys (with VALS, and
// their values) are already ordered by key >> shift, so only runs of equal
// prefixes (a handful of keys each) are left to order. Each block takes a
// chunk of FIN_CHUNK positions widened to whole runs (a run belongs to the
// chunk where it starts), stages it in LDS, and moves every key to its run's
// start plus its stable rank in the run (smaller keys, then equal keys that
// come before it), then writes the chunk back. Run bounds depend only on the prefix, which sorting inside a run never
// changes and which lies in the high 32 bits (shift >= 32), so reading a
// neighbour's elements while it rewrites them still sees the same bounds. A
// run longer than FINISH_MAX sets *flag; the caller then sorts in full.
constexpr int FINISH_MAX = 64;
constexpr int FIN_CHUNK = 2048;
template <bool VALS>
__global__ __launch_bounds__(256) void k_finish_runs(unsigned long long *keys, uint32_t *vals, uint64_t n,
                                                     int shift, unsigned long long *flag) {
    __shared__ unsigned long long sk[FIN_CHUNK + FINISH_MAX + 1];
    __shared__ uint32_t sv[VALS ? FIN_CHUNK + FINISH_MAX + 1 : 1];
    __shared__ uint64_t s_a, s_b;
    for (uint64_t c0 = (uint64_t)blockIdx.x * FIN_CHUNK; c0 < n; c0 += (uint64_t)gridDim.x * FIN_CHUNK) {
        const uint64_t c1 = c0 + FIN_CHUNK < n ? c0 + FIN_CHUNK : n;
        __syncthreads();  // the previous chunk is written back
        if (threadIdx.x < 64) {
            // [a, b): from the first run starting in [c0, c1) to the end of
            // the run holding c1 - 1; the first wave tests 64 positions at
            // once on each side (FINISH_MAX = 64)
            const int lane = threadIdx.x;
            uint64_t a = c0;
            if (c0 > 0) {
                const unsigned long long pp = keys[c0 - 1] >> shift;
                const uint64_t i = c0 + (uint64_t)lane;
                const bool other = i >= c1 || (keys[i] >> shift) != pp;
                const unsigned long long m = __ballot(other);
                a = m ? c0 + (uint64_t)__ffsll((long long)m) - 1 : c0 + 64;
                if (!m) a = c0 + FINISH_MAX + 1;  // the run goes on too far
            }
            uint64_t b = a;
            if (a < c1) {
                const unsigned long long pl = keys[c1 - 1] >> shift;
                const uint64_t i = c1 + (uint64_t)lane;
                const bool other = i >= n || (keys[i] >> shift) != pl;
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
            sk[i] = keys[a + i];
            if constexpr (VALS) sv[i] = vals[a + i];
        }
        __syncthreads();
        // every element goes to its run's start plus its stable rank
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            const unsigned long long key = sk[i], pfx = key >> shift;
            int rs = i;
            while (rs > 0 && (sk[rs - 1] >> shift) == pfx && i - rs < FINISH_MAX) rs--;
            int rank = 0, e = rs;
            while (e < len && (sk[e] >> shift) == pfx && e - rs <= FINISH_MAX) {
                rank += sk[e] < key || (sk[e] == key && e < i);
                e++;
            }
            if (e - rs > FINISH_MAX) {
                atomicOr(flag, 1ull);
                continue;
            }
            keys[a + rs + rank] = key;
            if constexpr (VALS) vals[a + rs + rank] = sv[i];
        }
    }
}

}  // namespace mczi
