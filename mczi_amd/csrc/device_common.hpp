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

}  // namespace mczi
