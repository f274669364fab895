// tg_device.h — device-side helpers shared by the gfx950 kernels.
//
// Encodings (reference vicLeva/tuna v1.4.0):
//   kache 2-bit code  A0 C1 G2 T3 = ((c>>2)^(c>>1))&3   src/partition_hash.hpp:69-70
//   nt 2-bit code     A0 C1 T2 G3 = kache ^ (kache>>1)  include/minimizer_window.hpp:138-140
//   packed stream     u32 word j holds bases 16j..16j+15, first base in bits 31:30
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

constexpr uint64_t EMPTY64 = ~0ull;

__device__ __forceinline__ uint32_t kache_code(uint32_t c) { return ((c >> 2) ^ (c >> 1)) & 3u; }

__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, uint32_t s)
{
    s &= 63u;
    return s ? (x << s) | (x >> (64u - s)) : x;
}
// One-bit rotations as two funnel shifts of the 32-bit halves (alignbit(a, b, s)
// is the low word of the 64-bit a:b shifted right by s).
__device__ __forceinline__ uint64_t rotl1(uint64_t x)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return ((uint64_t)__builtin_amdgcn_alignbit(hi, lo, 31u) << 32) | __builtin_amdgcn_alignbit(lo, hi, 31u);
}
__device__ __forceinline__ uint64_t rotr1(uint64_t x)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return ((uint64_t)__builtin_amdgcn_alignbit(lo, hi, 1u) << 32) | __builtin_amdgcn_alignbit(hi, lo, 1u);
}

// Published ntHash seeds for nt codes A, C, T, G (Mohamadi et al. 2016),
// the values of include/nt_hash.hpp:26-31.  The reverse strand uses seed[b^2].
__host__ __device__ __forceinline__ uint64_t nt_seed(uint32_t b)
{
    const uint64_t s0 = 0x3c8bfbb395c60474ull, s1 = 0x3193c18562a02b4cull;
    const uint64_t s2 = 0x295549f54be24456ull, s3 = 0x20323ed082572324ull;
    return b == 0 ? s0 : (b == 1 ? s1 : (b == 2 ? s2 : s3));
}

// 64-bit finaliser (MurmurHash3 fmix64) used for table slots only; slot
// placement never affects results.
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Reverse the order of the 2-bit groups of a 64-bit word.
__device__ __forceinline__ uint64_t rev_pairs64(uint64_t x)
{
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
}

// 2k-bit k-mer (k <= 32) whose first base is at bit offset 2*g of the packed
// stream (bit 0 = MSB of word 0).  Reads words g/16 .. g/16+2.
__device__ __forceinline__ uint64_t load_kmer64(const uint32_t* __restrict__ packed,
                                                uint64_t g, uint32_t k)
{
    const uint64_t w = g >> 4;
    const uint32_t sh = (uint32_t)(g & 15u) * 2u;
    const uint64_t hi = ((uint64_t)packed[w] << 32) | packed[w + 1];
    const uint32_t w2 = packed[w + 2];
    const uint64_t top = sh ? (hi << sh) | ((uint64_t)w2 >> (32u - sh)) : hi;
    return top >> (64u - 2u * k);
}

// 2k-bit k-mer, 32 < k <= 64, as (hi, lo) with lo = last 32 bases.
__device__ __forceinline__ void load_kmer128(const uint32_t* __restrict__ packed, uint64_t g,
                                             uint32_t k, uint64_t& hi, uint64_t& lo)
{
    const uint64_t w = g >> 4;
    const uint32_t sh = (uint32_t)(g & 15u) * 2u;
    uint64_t a = ((uint64_t)packed[w] << 32) | packed[w + 1];
    uint64_t b = ((uint64_t)packed[w + 2] << 32) | packed[w + 3];
    const uint32_t c = packed[w + 4];
    if (sh) {
        a = (a << sh) | (b >> (64u - sh));
        b = (b << sh) | ((uint64_t)c >> (32u - sh));
    }
    // a:b is a 128-bit left-aligned window; right-align 2k bits.
    const uint32_t r = 128u - 2u * k;  // 0..62
    if (r == 0) { hi = a; lo = b; }
    else { lo = (b >> r) | (a << (64u - r)); hi = a >> r; }
}

// Canonical form, k <= 32: min(fwd, reverse complement), complement = 3-b.
__device__ __forceinline__ uint64_t canon64(uint64_t x, uint32_t k)
{
    const uint64_t rc = (~rev_pairs64(x)) >> (64u - 2u * k);
    return x < rc ? x : rc;
}

__device__ __forceinline__ void canon128(uint64_t& hi, uint64_t& lo, uint32_t k)
{
    // reverse complement of the 128-bit value (hi:lo), then right-align 2k bits
    uint64_t rh = ~rev_pairs64(lo), rl = ~rev_pairs64(hi);
    const uint32_t r = 128u - 2u * k;
    if (r) { rl = (rl >> r) | (rh << (64u - r)); rh >>= r; }
    const bool fwd_smaller = hi < rh || (hi == rh && lo <= rl);
    if (!fwd_smaller) { hi = rh; lo = rl; }
}

}  // namespace tg
