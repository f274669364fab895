// merge.hip — the device merge table of a job counted in slices.
//
// A job too large for one pass of the engine (its record and index space is
// 32-bit) is cut at chunk boundaries into slices; k-mers never span chunks, so
// every k-mer of the job is counted in exactly one slice.  Each slice is
// counted without the ci/cx filter and its (key, count) pairs are added into
// this table; the filter is applied once, when the table is drained.  The
// reference reaches inputs larger than memory the same way in spirit: its disk
// pipeline writes every partition first and counts them afterwards
// (src/pipeline.hpp:186-263); here the unit of a pass is a slice of the input
// and the per-key sums make the passes exact.
//
// Layout: slot i holds nw key words at keys[nw*i] (word 0, the low 64 bits of
// the key, first, as in the result arrays) and a u32 count at cnt[i].
// k <= 32: a slot is claimed by a CAS of its key word; EMPTY64 (~0) is never a
// canonical key (the reverse complement of ~0 is 0, the smaller).  k > 32: the
// count word is the slot state (0 empty, LOCKED being written, else the
// count).  The host keeps the table at most half full, so every probe sequence
// ends at a free or matching slot.
#include "tg_device.h"
#include "tg_kernels.h"

namespace tg {

namespace {

constexpr uint32_t LOCKED = 0xFFFFFFFFu;
constexpr uint32_t MT = 256;  // threads per block

__device__ __forceinline__ uint64_t merge_slot(uint64_t hi, uint64_t lo, uint64_t mask)
{
    return mix64(lo ^ mix64(hi ^ 0x2545F4914F6CDD1Dull)) & mask;
}

// Adds c to the key's slot (claiming a free one on first sight); returns 1 when
// a slot was claimed.
template <int NW>
__device__ uint32_t merge_add(const MergeTab& t, uint64_t hi, uint64_t lo, uint32_t c)
{
    uint64_t s = merge_slot(NW == 2 ? hi : 0ull, lo, t.mask);
    for (uint64_t probes = 0; probes <= t.mask;) {
        if (NW == 1) {
            unsigned long long* kp = (unsigned long long*)&t.keys[s];
            uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t claimed = 0;
            if (cur == EMPTY64) {
                cur = atomicCAS(kp, (unsigned long long)EMPTY64, (unsigned long long)lo);
                claimed = cur == EMPTY64;
                if (claimed) cur = lo;
            }
            if (cur == lo) {
                atomicAdd(&t.cnt[s], c);
                return claimed;
            }
        } else {
            uint32_t* cp = &t.cnt[s];
            const uint32_t st = __hip_atomic_load(cp, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (st == 0u) {
                if (atomicCAS(cp, 0u, LOCKED) == 0u) {
                    __hip_atomic_store(&t.keys[2 * s], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&t.keys[2 * s + 1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(cp, c, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    return 1u;
                }
                continue;  // lost the claim: read the slot again
            }
            if (st == LOCKED) continue;  // being written by its claimer
            if (__hip_atomic_load(&t.keys[2 * s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == lo &&
                __hip_atomic_load(&t.keys[2 * s + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == hi) {
                atomicAdd(cp, c);
                return 0u;
            }
        }
        s = (s + 1) & t.mask;
        ++probes;
    }
    atomicOr(t.err, 1u);  // full table: the host sizing was violated
    return 0u;
}

// Sum of v over the block (lanes of wave 0 read the partial sums).
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* sh)
{
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    if (l == 0) sh[w] = v;
    __syncthreads();
    uint32_t tot = 0;
    for (uint32_t i = 0; i < MT / 64; ++i) tot += sh[i];
    __syncthreads();
    return tot;
}

template <int NW>
__global__ __launch_bounds__(MT) void k_merge_insert(MergeTab t, const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ cnt, uint64_t n)
{
    __shared__ uint32_t s_red[MT / 64];
    uint32_t claimed = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * MT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * MT) {
        const uint64_t lo = NW == 1 ? keys[i] : keys[2 * i];
        const uint64_t hi = NW == 1 ? 0ull : keys[2 * i + 1];
        claimed += merge_add<NW>(t, hi, lo, cnt[i]);
    }
    const uint32_t tot = block_sum(claimed, s_red);
    if (threadIdx.x == 0 && tot) atomicAdd(t.claimed, (unsigned long long)tot);
}

// Every occupied slot of src added into dst (a larger table).
template <int NW>
__global__ __launch_bounds__(MT) void k_merge_rehash(MergeTab dst, MergeTab src)
{
    __shared__ uint32_t s_red[MT / 64];
    uint32_t claimed = 0;
    for (uint64_t s = (uint64_t)blockIdx.x * MT + threadIdx.x; s <= src.mask; s += (uint64_t)gridDim.x * MT) {
        const uint32_t c = src.cnt[s];
        if (NW == 1) {
            const uint64_t key = src.keys[s];
            if (key != EMPTY64) claimed += merge_add<1>(dst, 0ull, key, c);
        } else if (c) {
            claimed += merge_add<2>(dst, src.keys[2 * s + 1], src.keys[2 * s], c);
        }
    }
    const uint32_t tot = block_sum(claimed, s_red);
    if (threadIdx.x == 0 && tot) atomicAdd(dst.claimed, (unsigned long long)tot);
}

// Occupied slots whose count c satisfies ci <= c <= cx, compacted into out
// (unspecified order); ctr[0] receives the number written.
template <int NW>
__global__ __launch_bounds__(MT) void k_merge_drain(MergeTab t, uint32_t ci, uint64_t cx,
                                                    uint64_t* __restrict__ out_keys,
                                                    uint32_t* __restrict__ out_cnt,
                                                    unsigned long long* ctr)
{
    __shared__ uint32_t s_wave[MT / 64];
    __shared__ unsigned long long s_base;
    constexpr uint32_t PER = 16;  // slots per thread
    const uint64_t b0 = (uint64_t)blockIdx.x * MT * PER;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t mine = 0;
    for (uint32_t j = 0; j < PER; ++j) {
        const uint64_t s = b0 + (uint64_t)j * MT + threadIdx.x;
        if (s > t.mask) break;
        const uint32_t c = t.cnt[s];
        const bool used = NW == 1 ? t.keys[s] != EMPTY64 : c != 0u;
        mine += used && c >= ci && (uint64_t)c <= cx;
    }
    // exclusive prefix of `mine` over the block
    uint32_t incl = mine;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if ((int)lane >= d) incl += y;
    }
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t i = 0; i < MT / 64; ++i) {
        if (i < wv) before += s_wave[i];
        total += s_wave[i];
    }
    if (threadIdx.x == 0) s_base = total ? atomicAdd(ctr, (unsigned long long)total) : 0ull;
    __syncthreads();
    uint64_t pos = s_base + before + incl - mine;
    for (uint32_t j = 0; j < PER; ++j) {
        const uint64_t s = b0 + (uint64_t)j * MT + threadIdx.x;
        if (s > t.mask) break;
        const uint32_t c = t.cnt[s];
        const bool used = NW == 1 ? t.keys[s] != EMPTY64 : c != 0u;
        if (used && c >= ci && (uint64_t)c <= cx) {
            if (NW == 1) {
                out_keys[pos] = t.keys[s];
            } else {
                out_keys[2 * pos] = t.keys[2 * s];
                out_keys[2 * pos + 1] = t.keys[2 * s + 1];
            }
            out_cnt[pos] = c;
            ++pos;
        }
    }
}

__global__ __launch_bounds__(MT) void k_merge_clear(MergeTab t, uint32_t nw)
{
    for (uint64_t s = (uint64_t)blockIdx.x * MT + threadIdx.x; s <= t.mask; s += (uint64_t)gridDim.x * MT) {
        if (nw == 1) t.keys[s] = EMPTY64;
        t.cnt[s] = 0u;
    }
}

unsigned grid_for(uint64_t n)
{
    const uint64_t b = (n + MT - 1) / MT;
    return (unsigned)std::min<uint64_t>(std::max<uint64_t>(b, 1), 8192);
}

}  // namespace

void launch_merge_clear(hipStream_t st, const MergeTab& t, uint32_t nw)
{
    k_merge_clear<<<grid_for(t.mask + 1), MT, 0, st>>>(t, nw);
}

void launch_merge_insert(hipStream_t st, const MergeTab& t, uint32_t nw, const uint64_t* keys,
                         const uint32_t* cnt, uint64_t n)
{
    if (!n) return;
    if (nw == 1) k_merge_insert<1><<<grid_for(n), MT, 0, st>>>(t, keys, cnt, n);
    else k_merge_insert<2><<<grid_for(n), MT, 0, st>>>(t, keys, cnt, n);
}

void launch_merge_rehash(hipStream_t st, const MergeTab& dst, const MergeTab& src, uint32_t nw)
{
    if (nw == 1) k_merge_rehash<1><<<grid_for(src.mask + 1), MT, 0, st>>>(dst, src);
    else k_merge_rehash<2><<<grid_for(src.mask + 1), MT, 0, st>>>(dst, src);
}

void launch_merge_drain(hipStream_t st, const MergeTab& t, uint32_t nw, uint32_t ci, uint64_t cx,
                        uint64_t* out_keys, uint32_t* out_cnt, unsigned long long* ctr)
{
    const uint64_t blocks = (t.mask + 1 + MT * 16 - 1) / (MT * 16);
    if (nw == 1) k_merge_drain<1><<<(unsigned)blocks, MT, 0, st>>>(t, ci, cx, out_keys, out_cnt, ctr);
    else k_merge_drain<2><<<(unsigned)blocks, MT, 0, st>>>(t, ci, cx, out_keys, out_cnt, ctr);
}

}  // namespace tg
