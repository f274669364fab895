// kernels.hip — gfx950 kernels of the two-phase k-mer counting path.
//
// Phase 1 (reference src/partition_hash.hpp:48-101, include/minimizer_window.hpp,
// include/nt_hash.hpp): pack ASCII to 2 bits, split chunks into strips, one lane
// per strip rolls canonical ntHash and the block-anchored window minimum and
// cuts superkmers; records go to strip-private slots.
// Routing (replaces the per-partition writers, src/superkmer_io.hpp:154-190):
// per-tile LDS histograms, column scan, LDS-cursor scatter into partition order.
// Phase 2 (src/count.hpp:71-175, Streaming_Kmer_Hash_Table upsert): one
// workgroup per work item keeps an open-addressed LDS table, spills to a
// global table when a probe window is full, then drains with the ci/cx filter
// (src/count.hpp:180-216).
#include "tg_device.h"
#include "tg_kernels.h"

#include <algorithm>
#include <mutex>
#include <type_traits>

namespace tg {

// ============================================================== pack ====
__device__ __forceinline__ uint32_t pack4(uint32_t x)
{
    const uint32_t t = ((x >> 2) ^ (x >> 1)) & 0x03030303u;
    return ((t & 3u) << 6) | (((t >> 8) & 3u) << 4) | (((t >> 16) & 3u) << 2) | (t >> 24);
}

__device__ __forceinline__ uint32_t pack16(uint4 v)
{
    return (pack4(v.x) << 24) | (pack4(v.y) << 16) | (pack4(v.z) << 8) | pack4(v.w);
}

// The 16 bytes from byte 4Q + r of the 32 bytes lo:hi (little-endian words).
template <int Q>
__device__ __forceinline__ uint4 bytes_at(uint4 lo, uint4 hi, uint32_t r)
{
    const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    return make_uint4(__builtin_amdgcn_alignbyte(w[Q + 1], w[Q], r), __builtin_amdgcn_alignbyte(w[Q + 2], w[Q + 1], r),
                      __builtin_amdgcn_alignbyte(w[Q + 3], w[Q + 2], r), __builtin_amdgcn_alignbyte(w[Q + 4], w[Q + 3], r));
}

// Word i of the packed stream: bases 16i..16i+15 (zero past n).  mis = the
// base pointer's offset from 16-byte alignment (the same for every thread): a
// misaligned input is read as the two aligned 16-byte words that cover the 16
// bases, while the second lies inside the input (the first lies in the input's
// aligned block, so in the same page); the last words go base by base.
template <int Q>
__device__ __forceinline__ uint32_t pack_word_q(const uint8_t* __restrict__ b, uint64_t n, uint64_t i, uint32_t mis)
{
    const uint64_t b0 = i * 16u;
    if (Q == 0 && mis == 0) {
        if (b0 + 16u <= n) return pack16(*reinterpret_cast<const uint4*>(b + b0));
    } else if (b0 + 32u <= n + mis) {
        const uint4* a = reinterpret_cast<const uint4*>(b - mis) + i;
        return pack16(bytes_at<Q>(a[0], a[1], mis & 3u));
    }
    uint32_t w = 0;
    for (uint32_t j = 0; j < 16u; ++j)
        if (b0 + j < n) w |= kache_code(b[b0 + j]) << (30u - 2u * j);
    return w;
}

// Four words per thread per trip with their loads issued together, while the
// four lie where every load is in range (word i reads at most bytes
// 16i - mis .. 16i - mis + 31 < n when 16i + 32 <= n); the rest word by word.
template <int Q>
__device__ __forceinline__ void pack_range(const uint8_t* __restrict__ b, uint64_t n, uint32_t* __restrict__ out,
                                           uint64_t n_words, uint32_t mis)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t fast_end = min(n_words, n >= 32u ? (n - 32u) / 16u + 1u : 0ull);
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    constexpr bool ALIGNED = Q == 0;
    if (ALIGNED && mis == 0) {
        const uint4* a = reinterpret_cast<const uint4*>(b);
        for (; i + 3u * stride < fast_end; i += 4u * stride) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = a[i + u * stride];
#pragma unroll
            for (int u = 0; u < 4; ++u) out[i + u * stride] = pack16(v[u]);
        }
    } else {
        const uint4* a = reinterpret_cast<const uint4*>(b - mis);
        for (; i + 3u * stride < fast_end; i += 4u * stride) {
            uint4 lo[4], hi[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                lo[u] = a[i + u * stride];
                hi[u] = a[i + u * stride + 1u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) out[i + u * stride] = pack16(bytes_at<Q>(lo[u], hi[u], mis & 3u));
        }
    }
    for (; i < n_words; i += stride) out[i] = pack_word_q<Q>(b, n, i, mis);
}

__device__ __forceinline__ void pack_any(const uint8_t* __restrict__ b, uint64_t n, uint32_t* __restrict__ out,
                                         uint64_t n_words)
{
    const uint32_t mis = (uint32_t)((uintptr_t)b & 15u);
    switch (mis >> 2) {
    case 0: pack_range<0>(b, n, out, n_words, mis); break;
    case 1: pack_range<1>(b, n, out, n_words, mis); break;
    case 2: pack_range<2>(b, n, out, n_words, mis); break;
    default: pack_range<3>(b, n, out, n_words, mis); break;
    }
}

__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ b, uint64_t n,
                                              uint32_t* __restrict__ out, uint64_t n_words)
{
    pack_any(b, n, out, n_words);
}

// Batch variant: bases and their count come from the batch descriptor.
__global__ __launch_bounds__(256) void k_pack_desc(const BatchDesc* __restrict__ d,
                                                   uint32_t* __restrict__ out, uint64_t cap_words)
{
    const uint64_t n = d->n_bases;
    pack_any(d->bases, n, out, min(cap_words, (n + 15) / 16 + 8));
}

void launch_pack_desc(hipStream_t st, const BatchDesc* d, uint32_t* out, uint64_t cap_words)
{
    uint64_t blocks = (cap_words + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    k_pack_desc<<<(unsigned)blocks, 256, 0, st>>>(d, out, cap_words);
}

void launch_pack(hipStream_t st, const uint8_t* bases, uint64_t n, uint32_t* out, uint64_t n_words)
{
    uint64_t blocks = (n_words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    k_pack<<<(unsigned)blocks, 256, 0, st>>>(bases, n, out, n_words);
}

// ============================================================ strips ====
__global__ void k_chunk_counts(const uint64_t* __restrict__ off, uint64_t nch, uint32_t k,
                               uint32_t S, uint64_t* __restrict__ nstr, uint64_t* __restrict__ nkm,
                               unsigned long long* __restrict__ long_list)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) long_list[0] = 0;  // the strip fill's piece count
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nch; c += stride) {
        // a decreasing offset pair is an empty chunk (keeps every later index in bounds)
        const uint64_t L = off[c + 1] > off[c] ? off[c + 1] - off[c] : 0;
        const uint64_t K = L >= k ? L - k + 1 : 0;
        nkm[c] = K;
        nstr[c] = (K + S - 1) / S;
    }
}

void launch_chunk_counts(hipStream_t st, const uint64_t* off, uint64_t n_chunks, uint32_t k,
                         uint32_t S, uint64_t* nstr, uint64_t* nkm, uint64_t* long_list)
{
    uint64_t blocks = (n_chunks + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    k_chunk_counts<<<(unsigned)blocks, 256, 0, st>>>(off, n_chunks, k, S, nstr, nkm,
                                                      reinterpret_cast<unsigned long long*>(long_list));
}

// Strip table: a thread per chunk writes the strips of a chunk of at most
// FILL_SHORT strips; a longer chunk is listed in pieces of FILL_PIECE strips
// (long_list[1..], their number in long_list[0]) and every piece is written by
// a workgroup of k_fill_strips_long, so one long chunk (a genome) is not
// written by a single thread.  Pieces per strip: at most 1/33 + 1/4096 < 1/32.
constexpr uint32_t FILL_SHORT = 32, FILL_PIECE = 4096;

uint64_t fill_list_words(uint64_t cap_strips) { return cap_strips / FILL_SHORT + 2; }

__global__ void k_fill_strips(uint64_t nch, const uint64_t* __restrict__ nstr_excl,
                              const uint64_t* __restrict__ nkm_excl,
                              const uint64_t* __restrict__ nkm, uint32_t S,
                              uint32_t* __restrict__ strip_chunk, uint32_t* __restrict__ strip_tb,
                              uint64_t* __restrict__ strip_slot, unsigned long long* __restrict__ long_list,
                              uint64_t long_cap)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nch; c += stride) {
        const uint64_t K = nkm[c];
        const uint64_t base = nstr_excl[c], kb = nkm_excl[c];
        const uint64_t ns = (K + S - 1) / S;
        if (ns > FILL_SHORT) {
            const uint64_t pieces = (ns + FILL_PIECE - 1) / FILL_PIECE;
            const uint64_t j = atomicAdd(&long_list[0], (unsigned long long)pieces);
            for (uint64_t q = 0; q < pieces && j + q < long_cap; ++q) long_list[1 + j + q] = c | (q << 32);
            continue;
        }
        for (uint64_t i = 0; i < ns; ++i) {
            strip_chunk[base + i] = (uint32_t)c;
            strip_tb[base + i] = (uint32_t)(i * S);
            strip_slot[base + i] = kb + i * S;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_strips_long(const uint64_t* __restrict__ nstr_excl,
                                                          const uint64_t* __restrict__ nkm_excl,
                                                          const uint64_t* __restrict__ nkm, uint32_t S,
                                                          uint32_t* __restrict__ strip_chunk,
                                                          uint32_t* __restrict__ strip_tb,
                                                          uint64_t* __restrict__ strip_slot,
                                                          const unsigned long long* __restrict__ long_list,
                                                          uint64_t long_cap)
{
    const uint64_t n = min((uint64_t)long_list[0], long_cap);
    for (uint64_t e = blockIdx.x; e < n; e += gridDim.x) {
        const uint64_t v = long_list[1 + e];
        const uint64_t c = v & 0xFFFFFFFFull, q = v >> 32;
        const uint64_t ns = (nkm[c] + S - 1) / S;
        const uint64_t base = nstr_excl[c], kb = nkm_excl[c];
        const uint64_t i1 = min(ns, (q + 1) * FILL_PIECE);
        for (uint64_t i = q * FILL_PIECE + threadIdx.x; i < i1; i += 256u) {
            strip_chunk[base + i] = (uint32_t)c;
            strip_tb[base + i] = (uint32_t)(i * S);
            strip_slot[base + i] = kb + i * S;
        }
    }
}

void launch_fill_strips(hipStream_t st, uint64_t n_chunks, const uint64_t* nstr_excl,
                        const uint64_t* nkm_excl, const uint64_t* nkm, uint32_t S,
                        uint32_t* strip_chunk, uint32_t* strip_tb, uint64_t* strip_slot,
                        uint64_t* long_list, uint64_t list_words)
{
    uint64_t blocks = (n_chunks + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    auto* ll = reinterpret_cast<unsigned long long*>(long_list);
    const uint64_t cap = list_words - 1;
    k_fill_strips<<<(unsigned)blocks, 256, 0, st>>>(n_chunks, nstr_excl, nkm_excl, nkm, S, strip_chunk,
                                                     strip_tb, strip_slot, ll, cap);
    const unsigned lb = (unsigned)std::min<uint64_t>(std::max<uint64_t>(cap, 1u), 1024u);
    k_fill_strips_long<<<lb, 256, 0, st>>>(nstr_excl, nkm_excl, nkm, S, strip_chunk, strip_tb, strip_slot, ll,
                                            cap);
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n);

// ============================================================== scan ====
// Exclusive scan of u64: blocks of 256 threads x 8 items.
constexpr uint32_t SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* sh, uint64_t& total)
{
    // sh: SCAN_T/64 entries
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    uint64_t wbase = 0, tot = 0;
    for (uint32_t i = 0; i < blockDim.x / 64u; ++i) {
        if (i < wv) wbase += sh[i];
        tot += sh[i];
    }
    total = tot;
    __syncthreads();
    return wbase + x - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const uint64_t* __restrict__ in, uint64_t n,
                                                      uint64_t* __restrict__ sums)
{
    __shared__ uint64_t sh[SCAN_T / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_I;
    uint64_t s = 0;
    if (base + SCAN_I <= n && (reinterpret_cast<uintptr_t>(in) & 15u) == 0) {
        // whole run of this thread: 16-byte loads
        const ulonglong2* p = reinterpret_cast<const ulonglong2*>(in + base);
#pragma unroll
        for (uint32_t i = 0; i < SCAN_I / 2; ++i) {
            const ulonglong2 v = p[i];
            s += v.x + v.y;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < SCAN_I; ++i)
            if (base + i < n) s += in[base + i];
    }
    uint64_t tot;
    block_excl_scan_u64(s, sh, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const uint64_t* __restrict__ in, uint64_t n,
                                                     const uint64_t* __restrict__ boff,
                                                     uint64_t* __restrict__ out)
{
    __shared__ uint64_t sh[SCAN_T / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_I;
    uint64_t v[SCAN_I];
    uint64_t s = 0;
    // whole run of this thread with both arrays 16-byte aligned: 16-byte
    // loads and stores
    const bool vec = base + SCAN_I <= n &&
                     ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    if (vec) {
        const ulonglong2* p = reinterpret_cast<const ulonglong2*>(in + base);
#pragma unroll
        for (uint32_t i = 0; i < SCAN_I / 2; ++i) {
            const ulonglong2 w = p[i];
            v[2 * i] = w.x;
            v[2 * i + 1] = w.y;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < SCAN_I; ++i) v[i] = base + i < n ? in[base + i] : 0;
    }
#pragma unroll
    for (uint32_t i = 0; i < SCAN_I; ++i) s += v[i];
    uint64_t tot;
    uint64_t run = block_excl_scan_u64(s, sh, tot) + (boff ? boff[blockIdx.x] : 0);
    if (vec) {
        ulonglong2* q = reinterpret_cast<ulonglong2*>(out + base);
#pragma unroll
        for (uint32_t i = 0; i < SCAN_I / 2; ++i) {
            ulonglong2 w;
            w.x = run;
            run += v[2 * i];
            w.y = run;
            run += v[2 * i + 1];
            q[i] = w;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < SCAN_I; ++i) {
            if (base + i < n) out[base + i] = run;
            run += v[i];
        }
    }
    if (base < n && base + SCAN_I >= n) out[n] = run;  // thread owning the last item
}

uint64_t scan_tmp_words(uint64_t n)
{
    uint64_t words = 0;
    while (n > SCAN_TILE) {
        const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
        words += 2 * nb + 2;  // sums + scanned sums (+ total)
        n = nb;
    }
    return words + 2;
}

void launch_scan_u64(hipStream_t st, const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* tmp)
{
    if (n == 0) {
        k_fill_u64<<<1, 64, 0, st>>>(out, 0ull, 1);
        return;
    }
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 1) {
        k_scan_apply<<<1, SCAN_T, 0, st>>>(in, n, nullptr, out);
        return;
    }
    uint64_t* sums = tmp;
    uint64_t* soff = tmp + nb;
    k_scan_reduce<<<(unsigned)nb, SCAN_T, 0, st>>>(in, n, sums);
    launch_scan_u64(st, sums, nb, soff, tmp + 2 * nb + 2);
    k_scan_apply<<<(unsigned)nb, SCAN_T, 0, st>>>(in, n, soff, out);
}

__global__ void k_u32_to_u64(const uint32_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

void launch_u32_to_u64(hipStream_t st, const uint32_t* in, uint64_t n, uint64_t* out)
{
    if (!n) return;
    uint64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    k_u32_to_u64<<<(unsigned)b, 256, 0, st>>>(in, n, out);
}

// ======================================================== superkmers ====
// Lane state for one strip: rolling ntHash over the packed stream plus the
// superkmer cut state.  The cut rule follows src/partition_hash.hpp:78-93: a
// superkmer starts at t = 0, where the window minimum hash changes, and where
// the current superkmer already holds capk = 256-k k-mers (the 255-base cap of
// the one-byte record header, :61-63,81).
struct SkLane {
    // packed-stream reader + rolling hash (include/nt_hash.hpp:114-118).  The
    // next bases sit top-aligned in buf; every 16 rolls (the same rolls for
    // all lanes of the wave still rolling) the word loaded 16 rolls earlier is
    // appended at bit 2*phase and the next one is loaded, so the refill is a
    // wave-uniform branch and its load has 16 rolls to arrive.
    const uint32_t* packed;
    uint64_t widx;              // index of the word in nxt
    uint64_t buf;
    uint32_t nxt, phase2, u;    // prefetched word, 2 * (start base & 15), rolls so far
    uint64_t f, rv, lmer;
    uint32_t osh;
    // cut state
    uint32_t ts, tb, te, K;
    uint32_t t0, t0p, P0;       // superkmer start, current piece start, minimizer position
    uint64_t prevH;
    uint32_t hsel;
    uint32_t rbase, rl;         // bucket base and bucket bits of the strip's job
    uint32_t job, nsk;          // job of the strip, superkmers emitted
    bool have, done, flagged, tail;
    uint64_t cb;
};

// tab: the 16 forward entries, then the 16 reverse ones, indexed by
// (outgoing << 2 | incoming) in kache codes
__device__ __forceinline__ void sk_roll(SkLane& L, const uint64_t* tab)
{
    // the lanes that roll agree on u
    const uint32_t u = __builtin_amdgcn_readfirstlane(L.u);
    if (u && (u & 15u) == 0) {
        L.buf |= (uint64_t)L.nxt << L.phase2;
        L.nxt = L.packed[++L.widx];  // the stream keeps 8 spare words
    }
    L.u = u + 1u;
    // byte offset of entry (o, in) = 8 (o << 2 | in), in one funnel shift:
    // the outgoing base's code in the low bits of the m-mer shifted down,
    // the incoming one in the stream's top bits
    const uint32_t hi = (uint32_t)(L.buf >> 32);
    const uint32_t kb = hi >> 30;
    const uint32_t off = __builtin_amdgcn_alignbit((uint32_t)(L.lmer >> L.osh), hi, 27u) & 0x78u;
    L.buf <<= 2;
    L.lmer = (L.lmer << 2) | kb;
    const uint64_t* e = reinterpret_cast<const uint64_t*>(reinterpret_cast<const char*>(tab) + off);
    L.f = rotl1(L.f) ^ e[0];
    L.rv = rotr1(L.rv) ^ e[16];
}

// Routing field of a record (18 bits of the minimizer hash H): the bucket
// base + (H & (2^L - 1)) on top, then the next XB bits of H.  A single job
// has base 0 and L = log2(n), so the reference partition H & (n-1)
// (src/partition_hash.hpp:549-553) is the top of the field and the bits below
// it pick the device sub-partition and split large ones in Phase 2.  In a batch
// every job owns its own bucket range, so no table ever mixes two jobs.
__device__ __forceinline__ uint32_t route_hsel(uint64_t H, uint32_t base, uint32_t L, uint32_t xb)
{
    const uint32_t h = (uint32_t)H;
    return ((base + (h & ((1u << L) - 1u))) << xb) | ((h >> L) & ((1u << xb) - 1u));
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A record of RW words, moved with vector loads and stores.
template <int RW>
struct Rec {
    uint32_t w[RW];
};

template <int RW>
__device__ __forceinline__ Rec<RW> load_rec(const uint32_t* base, uint64_t i)
{
    Rec<RW> r;
    if (RW == 1) {
        r.w[0] = base[i];
    } else if (RW == 2) {
        const uint2 v = reinterpret_cast<const uint2*>(base)[i];
        r.w[0] = v.x;
        r.w[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < RW / 4; ++q) {
            const uint4 v = reinterpret_cast<const uint4*>(base)[i * (RW / 4) + q];
            r.w[4 * q] = v.x;
            r.w[4 * q + 1] = v.y;
            r.w[4 * q + 2] = v.z;
            r.w[4 * q + 3] = v.w;
        }
    }
    return r;
}

template <int RW>
__device__ __forceinline__ void store_rec(uint32_t* base, uint64_t i, const Rec<RW>& r)
{
    if (RW == 2) {
        reinterpret_cast<uint2*>(base)[i] = make_uint2(r.w[0], r.w[1]);
    } else {
#pragma unroll
        for (int q = 0; q < RW / 4; ++q)
            reinterpret_cast<uint4*>(base)[i * (RW / 4) + q] =
                make_uint4(r.w[4 * q], r.w[4 * q + 1], r.w[4 * q + 2], r.w[4 * q + 3]);
    }
}

template <int RW>
__device__ __forceinline__ uint32_t rec_meta(const Rec<RW>& r) { return RW == 2 ? r.w[1] : r.w[0]; }

// the meta word alone (RW = 1 views the meta array)
template <>
__device__ __forceinline__ uint32_t rec_meta<1>(const Rec<1>& r) { return r.w[0]; }

template <int RW>
__device__ __forceinline__ uint32_t rec_len(uint32_t meta)
{
    return RW == 2 ? (meta & REC_LEN_MASK) : (meta & REC_FAT_LEN_MASK);
}

// Fat record of the piece starting at global base `start` with meta `meta`,
// from the packed words in[] = packed[start / 16 ...]: its bases realigned so
// the first is in bits 31:30 of word 1, bases past the piece cleared (the
// packed stream keeps 8 spare words, so the reads stay in bounds).
template <int RW>
__device__ __forceinline__ Rec<RW> fat_build(const uint32_t (&in)[RW], uint32_t start, uint32_t meta)
{
    Rec<RW> r;
    const uint32_t len = meta & REC_FAT_LEN_MASK;
    const uint32_t sh = 2u * (start & 15u);
    r.w[0] = meta;
#pragma unroll
    for (int q = 0; q < RW - 1; ++q) {
        uint32_t v = sh ? (in[q] << sh) | (in[q + 1] >> (32u - sh)) : in[q];
        const int keep = (int)len - 16 * q;
        v = keep <= 0 ? 0u : (keep >= 16 ? v : v & ~(0xFFFFFFFFu >> (2 * keep)));
        r.w[q + 1] = v;
    }
    return r;
}

// Per-wave staging of records: the active lanes append (start, meta) with a
// ballot, and when the buffer nears full they write it to the workgroup's run
// (space taken with one LDS atomic) with every lane active: that is where a
// fat record gathers its bases, outside the divergent per-k-mer loop.
struct SkStage {
    uint2* buf;        // this wave's buffer (LDS)
    uint32_t* wg_n;    // workgroup run fill (LDS)
    uint64_t run;      // workgroup run: first record index
    // The wave's fill is pushed - flushed, kept in registers.  Both counters
    // only grow, and every lane still in the loop has taken part in every
    // push and flush so far, so the active lanes agree; a lane that left the
    // loop keeps older values, and the wave's counts are then the maxima.
    uint32_t pushed, flushed;
};

// The record writes of a flush, out of line (a flush comes every few dozen
// k-mers and is inlined at every k-mer of the unrolled window loop otherwise);
// its arguments are plain values so that nothing is copied to the stack.
// TWO: two records per lane per round (their packed words loaded together);
// the inlined flush of the small windows keeps one, as a second register set
// would spill there.
template <int RW, bool TWO>
__device__ __forceinline__ void sk_write_recs(const uint2* buf, uint32_t cnt, uint32_t rank, uint32_t nact,
                                              uint64_t at, uint32_t* recs, uint32_t* metas,
                                              const uint32_t* packed)
{
    if (RW == 2) {
        for (uint32_t i = rank; i < cnt; i += nact) {
            const uint2 e = buf[i];
            Rec<RW> r;
            r.w[0] = e.x;
            r.w[1] = e.y;
            store_rec<RW>(recs, at + i, r);
            metas[at + i] = e.y;
        }
        return;
    }
    if (!TWO) {
        for (uint32_t i = rank; i < cnt; i += nact) {
            const uint2 e = buf[i];
            const uint32_t* s1 = packed + (e.x >> 4);
            uint32_t in1[RW];
#pragma unroll
            for (int q = 0; q < RW; ++q) in1[q] = s1[q];
            store_rec<RW>(recs, at + i, fat_build<RW>(in1, e.x, e.y));
            metas[at + i] = e.y;
        }
        return;
    }
    // a lane without a second record loads its first one's words again
    for (uint32_t i = rank; i < cnt; i += 2u * nact) {
        const uint32_t i2 = i + nact;
        const bool two = i2 < cnt;
        const uint2 e = buf[i], e2 = buf[two ? i2 : i];
        const uint32_t* s1 = packed + (e.x >> 4);
        const uint32_t* s2 = packed + (e2.x >> 4);
        uint32_t in1[RW], in2[RW];
#pragma unroll
        for (int q = 0; q < RW; ++q) {
            in1[q] = s1[q];
            in2[q] = s2[q];
        }
        store_rec<RW>(recs, at + i, fat_build<RW>(in1, e.x, e.y));
        metas[at + i] = e.y;
        if (two) {
            store_rec<RW>(recs, at + i2, fat_build<RW>(in2, e2.x, e2.y));
            metas[at + i2] = e2.y;
        }
    }
}

template <int RW>
__device__ __noinline__ void sk_write_out(const uint2* buf, uint32_t cnt, uint32_t rank, uint32_t nact,
                                          uint64_t at, uint32_t* recs, uint32_t* metas,
                                          const uint32_t* packed)
{
    sk_write_recs<RW, true>(buf, cnt, rank, nact, at, recs, metas, packed);
}

// OL: the record writes out of line (the large windows, whose unrolled loop
// has many k-mers and registers to spare for a call; the small ones would
// spill registers around it)
template <int RW, bool OL>
__device__ __forceinline__ void sk_flush(SkStage& S, const SkArgs& a, uint32_t cnt)
{
    const uint64_t act = __ballot(1);
    const uint32_t nact = (uint32_t)__popcll(act), rank = lane_rank(act);
    uint32_t base = 0;
    if (rank == 0) base = atomicAdd(S.wg_n, cnt);
    base = __builtin_amdgcn_readfirstlane(base);
    if (OL) sk_write_out<RW>(S.buf, cnt, rank, nact, S.run + base, a.recs, a.metas, a.packed);
    else sk_write_recs<RW, false>(S.buf, cnt, rank, nact, S.run + base, a.recs, a.metas, a.packed);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    S.flushed = S.pushed;
}

template <int RW, bool OL>
__device__ __forceinline__ void sk_push(SkStage& S, const SkArgs& a, bool e, uint2 rec)
{
    const uint64_t mask = __ballot(e);
    if (!mask) return;
    const uint32_t base = S.pushed - S.flushed;
    // lanes that do not emit write the scratch slot past the buffer end
    const uint32_t slot = e ? base + lane_rank(mask) : SK_WAVE_BUF + (threadIdx.x & 63u);
    S.buf[slot] = rec;
    S.pushed += (uint32_t)__popcll(mask);
    const uint32_t cnt = S.pushed - S.flushed;
    if (cnt > SK_WAVE_BUF - 64u) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        sk_flush<RW, OL>(S, a, cnt);
    }
}

// Staged entry of the piece [p0, t_end) of the current superkmer: global start
// and meta.  Slim records keep the superkmer's minpos; a fat piece after the
// first sets REC_CONT instead.
template <int RW>
__device__ __forceinline__ uint2 sk_entry(const SkLane& L, const SkArgs& a, uint32_t p0, uint32_t t_end)
{
    const uint32_t len = t_end - p0 + a.k - 1u;
    const bool first = p0 == L.t0;
    const uint32_t meta = (RW != 2 && !first) ? (len | REC_CONT | (L.hsel << REC_HSEL_SHIFT))
                                              : (len | ((L.P0 - L.t0) << REC_MINPOS_SHIFT) |
                                                 (L.hsel << REC_HSEL_SHIFT));
    return make_uint2((uint32_t)L.cb + p0, meta);
}

// Decision for k-mer t with window minimum (Hc, Pc), written without branches
// so the active lanes stay converged for the one ballot in sk_push.  With fat
// records a piece ends after a.piece k-mers (a cut that starts no new
// superkmer).
// PUSH false: the superkmers are only counted (nsk), no record is staged.
template <int RW, bool OL, bool PUSH = true>
__device__ __forceinline__ void sk_kmer(SkLane& L, const SkArgs& a, SkStage& S, uint32_t t,
                                        uint64_t Hc, uint32_t Pc)
{
    const bool first = t == L.ts;
    const bool cut = !first && ((Hc != L.prevH) || (L.have && (t - L.t0 >= a.capk)));
    const bool pcut = RW != 2 && !first && !cut && L.have && (t - L.t0p >= a.piece);
    const bool emit = (cut || pcut) && L.have && L.t0 >= L.tb;
    const bool flag = !first && !cut && t == L.tb && !L.have;
    const uint2 rec = sk_entry<RW>(L, a, L.t0p, t);
    L.nsk += (emit && L.t0p == L.t0) ? 1u : 0u;
    const bool stop = (cut && t >= L.te) || flag;
    const bool start = (first && L.ts == 0) || (cut && t < L.te);
    L.t0 = start ? t : L.t0;
    L.t0p = (start || pcut) ? t : L.t0p;
    L.P0 = start ? Pc : L.P0;
    L.hsel = start ? route_hsel(Hc, L.rbase, L.rl, a.xb) : L.hsel;
    L.have = L.have || start;
    L.prevH = Hc;
    const bool last = !stop && t == L.K - 1u;
    L.tail = last && L.have && L.t0 >= L.tb;   // [t0p, K) is emitted after the loop
    L.flagged = L.flagged || flag;
    L.done = stop || last;
    if (PUSH) sk_push<RW, OL>(S, a, emit, rec);
}

// The same decision for a head strip (the strip starts at the chunk's first
// k-mer: ts = tb = 0), which needs neither look-back nor the flag: a cut
// (minimizer change, or the cap: t >= tcap) starts a superkmer below te and
// stops the lane at or past te; a piece ends at t >= tpiece.  Every threshold
// is set at the cut or piece end that starts a piece, so a k-mer costs three
// compares and the selects of the new piece's fields.  tcap starts at 0, so
// k-mer 0 starts the first superkmer (nothing is emitted before it: have).
// nsk counts superkmer starts: each started superkmer's first piece is
// emitted exactly once (at the next cut or piece end, at the stop, or as the
// chunk-end record).
// The head state lives in registers of the general one (a head strip has
// ts = tb = 0 and needs neither t0 nor P0 otherwise): tcap in t0, tpiece in
// t0p (set to ~0 when the head loop starts), the current piece's global start
// in P0, k - 1 - its start in ts, the current piece's meta bits in hsel and a
// continuation piece's in tb.
template <int RW, bool OL>
__device__ __forceinline__ void sk_kmer_head(SkLane& L, const SkArgs& a, SkStage& S, uint32_t t,
                                             uint64_t Hc, uint32_t Pc)
{
    uint32_t& tcap = L.t0;
    uint32_t& tpiece = L.t0p;
    uint32_t& sx = L.P0;
    uint32_t& lb = L.ts;
    uint32_t& mcur = L.hsel;
    uint32_t& mcont = L.tb;
    const bool cut = Hc != L.prevH || t >= tcap;
    const bool ev = cut || t >= tpiece;
    const bool emit = ev && L.have;
    const uint2 rec = make_uint2(sx, (t + lb) | mcur);
    const bool stop = cut && t >= L.te;
    const bool start = cut && !stop;
    const uint32_t hs = route_hsel(Hc, L.rbase, L.rl, a.xb) << REC_HSEL_SHIFT;
    const uint32_t mhead = ((Pc - t) << REC_MINPOS_SHIFT) | hs;
    mcur = start ? mhead : (ev ? mcont : mcur);
    mcont = start ? (hs | REC_CONT) : mcont;
    sx = ev ? (uint32_t)L.cb + t : sx;
    lb = ev ? a.k - 1u - t : lb;
    tpiece = RW != 2 && ev ? t + a.piece : tpiece;
    tcap = start ? t + a.capk : tcap;
    L.nsk += start ? 1u : 0u;
    L.have = true;
    L.prevH = Hc;
    const bool last = !stop && t == L.K - 1u;
    L.tail = last;
    L.done = stop || last;
    sk_push<RW, OL>(S, a, emit, rec);
}

__device__ __forceinline__ void sk_init(SkLane& L, const SkArgs& a, uint32_t s, const uint64_t* s_init)
{
    const uint32_t c = a.strip_chunk[s];
    if (a.job_route) {
        L.job = a.chunk_job[c];
        const uint32_t r = a.job_route[L.job];
        L.rbase = r & 0xFFFFu;
        L.rl = r >> 16;
    } else {
        L.job = 0;
        L.rbase = 0;
        L.rl = a.log2n;
    }
    L.tb = a.strip_tb[s];
    L.cb = a.off[c];
    const uint32_t len = (uint32_t)(a.off[c + 1] - L.cb);
    L.K = len - a.k + 1u;
    L.te = min(L.tb + a.S, L.K);
    L.ts = 0;
    if (a.lookback != LOOKBACK_FULL && L.tb > a.lookback) L.ts = ((L.tb - a.lookback) / a.W) * a.W;
    L.packed = a.packed;
    const uint64_t gq = L.cb + L.ts;
    const uint64_t w0 = gq >> 4;
    L.phase2 = 2u * (uint32_t)(gq & 15u);
    L.buf = (((uint64_t)a.packed[w0] << 32) | a.packed[w0 + 1]) << L.phase2;
    L.widx = w0 + 2;
    L.nxt = a.packed[w0 + 2];
    L.u = 0;
    L.f = s_init[0];
    L.rv = s_init[1];
    L.lmer = 0;
    L.osh = 2u * (a.m - 1u);
    L.t0 = 0; L.t0p = 0; L.P0 = 0; L.prevH = 0;
    L.hsel = 0;
    L.nsk = 0;
    L.have = false; L.done = false; L.flagged = false; L.tail = false;
}

__device__ __forceinline__ void sk_tables(const SkArgs& a, uint64_t* tf, uint64_t* tr, uint64_t* s_init)
{
    const uint32_t m = a.m;
    if (threadIdx.x < 16) {
        // entry (out, in) in kache codes; the seeds take nt codes (c ^ c >> 1)
        const uint32_t ko = threadIdx.x >> 2, ki = threadIdx.x & 3u;
        const uint32_t o = ko ^ (ko >> 1), i = ki ^ (ki >> 1);
        // include/nt_hash.hpp:80-93
        tf[threadIdx.x] = rotl64(nt_seed(o), m) ^ nt_seed(i);
        tr[threadIdx.x] = rotl64(nt_seed(o ^ 2u), 63u) ^ rotl64(nt_seed(i ^ 2u), m - 1u);
    }
    if (threadIdx.x == 16) {
        // hashes of the virtual m-mer A^m: rolling m real bases into it gives
        // the exact hash of the first real m-mer (the A contributions cancel).
        uint64_t f = 0, r = 0;
        for (uint32_t i = 0; i < m; ++i) {
            f ^= rotl64(nt_seed(0), m - 1u - i);
            r ^= rotl64(nt_seed(2), i);
        }
        s_init[0] = f;
        s_init[1] = r;
    }
}

// Register-resident window for W <= WMAX.  Per block of W m-mers: sv[r] holds
// the suffix minimum of the previous ("older") block from r on, with its
// rightmost argmin sp[r]; the running prefix minimum of the current block is
// (pv, pp), leftmost argmin.  This is the two-stack of
// include/minimizer_window.hpp:66-99 with its tie rule (strict '<' on both
// sides, the older side wins only when strictly smaller, :176-199).
template <int WMAX>
__device__ __forceinline__ void sk_backward(uint64_t (&sv)[WMAX], uint8_t (&sp)[WMAX], uint32_t W)
{
    uint64_t cur = 0;
    uint32_t cp = 0;
#pragma unroll
    for (int q = WMAX - 1; q >= 0; --q) {
        if ((uint32_t)q < W) {
            if ((uint32_t)q == W - 1u || sv[q] < cur) {
                cur = sv[q];
                cp = (uint32_t)q;
            }
            sv[q] = cur;
            sp[q] = (uint8_t)cp;
        }
    }
}

// MODE 0: every block, general decision (the fix-up pass); MODE 1: only the
// blocks whose strips all start at their chunk's first k-mer (every block of
// a read input), with the head decision; MODE 2: only the other blocks,
// general decision.  The main pass launches MODE 1 and MODE 2 over the same
// grid, so each block is counted by exactly one of them.
template <int WMAX, int RW, int MODE>
__global__ __launch_bounds__(SK_BLOCK) __attribute__((amdgpu_waves_per_eu(WMAX <= 12 ? 5 : 1))) void k_superkmers(SkArgs a)
{
    // one LDS object, so the hash tables sit at address 0 and a table read
    // needs no base added to its offset
    struct SkShared {
        uint64_t tab[32];  // forward [0, 16), reverse [16, 32)
        uint64_t init[2];
        uint2 buf[SK_BLOCK / 64][SK_WAVE_BUF + 64];
        uint32_t wgn;
    };
    __shared__ SkShared sh;
    const uint64_t* tab = sh.tab;
    const uint64_t* s_init = sh.init;
    auto& s_buf = sh.buf;
    uint32_t& s_wgn = sh.wgn;
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t idx = blockIdx.x * SK_BLOCK + threadIdx.x;
    // strips (or flagged strips) of this launch; with device-side sizes the
    // launch covers a capacity and the live count is read here
    uint32_t n = a.n;
    if (a.n_dev) n = (uint32_t)min((uint64_t)n, *a.n_dev);
    if (a.fix_dev) {
        const uint32_t nf = *a.flag_count;
        if (nf > n && blockIdx.x == 0 && threadIdx.x == 0 && MODE != 2) atomicOr(a.err, 1u);
        n = min(n, nf);
    }
    if (MODE != 0) {
        // a block without strips counts as a head block
        const bool hd = idx >= n || a.strip_tb[a.list ? a.list[idx] : idx] == 0u;
        if ((__syncthreads_and(hd) != 0) != (MODE == 1)) return;
    }
    sk_tables(a, sh.tab, sh.tab + 16, sh.init);
    if (threadIdx.x == 0) s_wgn = 0;
    __syncthreads();
    SkStage S;
    S.buf = s_buf[wv];
    S.wg_n = &s_wgn;
    S.pushed = 0;
    S.flushed = 0;
    // main pass: the run is the slot region of this block's strips; fix-up:
    // a private region of SK_BLOCK*S records per block
    const uint32_t first = blockIdx.x * SK_BLOCK;
    S.run = a.list ? (uint64_t)a.fix_base + (uint64_t)blockIdx.x * SK_BLOCK * a.S
                   : (first < n ? a.strip_slot[first] : 0ull);
    bool tail = false;
    uint2 tail_rec = make_uint2(0u, 0u);
    uint32_t nsk = 0, job = 0;
    if (idx < n) {
        const uint32_t s = a.list ? a.list[idx] : idx;
        SkLane L;
        sk_init(L, a, s, s_init);
        const uint32_t W = a.W;
        for (uint32_t i = 1; i < a.m; ++i) sk_roll(L, tab);
        uint64_t sv[WMAX];
        uint8_t sp[WMAX];
#pragma unroll
        for (int r = 0; r < WMAX; ++r) {
            if ((uint32_t)r < W) {
                sk_roll(L, tab);
                sv[r] = L.f ^ L.rv;
            }
        }
        sk_backward<WMAX>(sv, sp, W);
        uint32_t bs = L.ts;
        uint64_t pv = 0;
        uint32_t pp = 0;
        constexpr bool head = MODE == 1;
        if (head) {
            // head state (sk_kmer_head): tcap = t0 = 0 (k-mer 0 cuts), no
            // piece end yet; the other fields are set at that first cut
            L.t0 = 0;
            L.t0p = 0xFFFFFFFFu;
        }
        while (!L.done) {
#pragma unroll
            for (int r = 0; r < WMAX; ++r) {
                if ((uint32_t)r < W && !L.done) {
                    const uint32_t t = bs + (uint32_t)r;
                    uint64_t Hc;
                    uint32_t Pc;
                    if (r == 0) {
                        Hc = sv[0];
                        Pc = bs + sp[0];
                    } else if (sv[r] < pv) {
                        Hc = sv[r];
                        Pc = bs + sp[r];
                    } else {
                        Hc = pv;
                        Pc = bs + W + pp;
                    }
                    if constexpr (head) sk_kmer_head<RW, (WMAX > 16)>(L, a, S, t, Hc, Pc);
                    else sk_kmer<RW, (WMAX > 16)>(L, a, S, t, Hc, Pc);
                    if (!L.done) {
                        sk_roll(L, tab);
                        const uint64_t h = L.f ^ L.rv;
                        sv[r] = h;
                        if (r == 0 || h < pv) {
                            pv = h;
                            pp = (uint32_t)r;
                        }
                    }
                }
            }
            if (!L.done) {
                sk_backward<WMAX>(sv, sp, W);
                bs += W;
            }
        }
        if (L.flagged) {
            const uint32_t slot = atomicAdd(a.flag_count, 1u);
            a.flag_list[slot] = s;
        }
        tail = L.tail;
        if (head) {
            tail_rec = make_uint2(L.P0, (L.K + L.ts) | L.hsel);
            nsk = L.nsk;
        } else {
            tail_rec = sk_entry<RW>(L, a, L.t0p, L.K);
            nsk = L.nsk + ((tail && L.t0p == L.t0) ? 1u : 0u);
        }
        job = L.job;
    }
    // every lane of the wave is here again: the wave's counts, the chunk-end
    // records, then flush
#pragma unroll
    for (uint32_t d = 32; d; d >>= 1) {
        S.pushed = max(S.pushed, (uint32_t)__shfl_xor((int)S.pushed, d, 64));
        S.flushed = max(S.flushed, (uint32_t)__shfl_xor((int)S.flushed, d, 64));
    }
    sk_push<RW, (WMAX > 16)>(S, a, tail, tail_rec);
    const uint32_t left = S.pushed - S.flushed;
    if (left) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        sk_flush<RW, (WMAX > 16)>(S, a, left);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.run_start[a.run_base + blockIdx.x] = (uint32_t)S.run;
        a.run_cnt[a.run_base + blockIdx.x] = s_wgn;
    }
    // superkmers per job: one atomic per wave when the wave's strips share a job
    const uint32_t j0 = __builtin_amdgcn_readfirstlane(job);
    if (__ballot(job != j0 && nsk) == 0) {
        uint32_t sum = nsk;
#pragma unroll
        for (uint32_t d = 32; d; d >>= 1) sum += __shfl_xor(sum, d, 64);
        if ((threadIdx.x & 63u) == 0 && sum) atomicAdd(&a.sk_count[j0], (unsigned long long)sum);
    } else if (nsk) {
        atomicAdd(&a.sk_count[job], (unsigned long long)nsk);
    }
}

void launch_superkmers(hipStream_t st, const SkArgs& a, uint32_t rw)
{
    if (a.n == 0) return;
    const unsigned blocks = (a.n + SK_BLOCK - 1) / SK_BLOCK;
    // the fix-up pass (a list of flagged strips) has no head strip
#define TG_SK(WM, RWV)                                                    \
    do {                                                                   \
        if (a.list) {                                                      \
            k_superkmers<WM, RWV, 0><<<blocks, SK_BLOCK, 0, st>>>(a);      \
        } else {                                                           \
            k_superkmers<WM, RWV, 1><<<blocks, SK_BLOCK, 0, st>>>(a);      \
            k_superkmers<WM, RWV, 2><<<blocks, SK_BLOCK, 0, st>>>(a);      \
        }                                                                  \
    } while (0)
#define TG_SK_RW(WM) \
    do { if (rw == 2) TG_SK(WM, 2); else if (rw == 4) TG_SK(WM, 4); else TG_SK(WM, 8); } while (0)
    if (a.W <= 8) TG_SK_RW(8);
    else if (a.W <= 12) TG_SK_RW(12);
    else if (a.W <= 16) TG_SK_RW(16);
    else if (a.W <= 32) TG_SK_RW(32);
    else if (a.W <= 48) TG_SK_RW(48);
    else TG_SK_RW(64);
#undef TG_SK_RW
#undef TG_SK
}

// Count word of the HBM spill and merge tables as slot state: being written.
constexpr uint32_t LOCKED32 = 0xFFFFFFFFu;

// ============================================================ k > 64 ====
// Keys of more than 64 bases take a path of their own: the superkmers are
// counted (the `superkmers:` total) without records, and every k-mer goes
// straight into one open-addressed table in HBM that stores each key's
// position in the packed stream (no routing: a key of 4 or 8 words would leave
// room for few keys in an LDS table).

// Superkmers of k > 64, counted only: the pass over the strips the main pass
// (k_superkmers_wdq below, which keeps a deque instead of the window) lists,
// those whose deque overflowed (same look-back) and those whose look-back held
// no cut (from their chunk start).  The window holds W <= 256 m-mer hashes per
// lane in LDS ([position][lane], one wave per workgroup) and keeps their values
// only: a cut depends on the value of the window minimum, never on its
// position.  The decision is the general one (look-back, flags, cap) with the
// cap of the reference's header type sk_hdr_t (src/superkmer_io.hpp:45-48:
// 2k - m <= 255 gives a one-byte length, 255 bases, else 65535).
__device__ __forceinline__ void wide_backward(uint64_t* sv, uint32_t W, uint32_t lane)
{
    uint64_t cur = 0;
    for (int q = (int)W - 1; q >= 0; --q) {
        const uint64_t v = sv[q * 64 + lane];
        if ((uint32_t)q == W - 1u || v < cur) cur = v;
        sv[q * 64 + lane] = cur;
    }
}

__global__ __launch_bounds__(64) void k_superkmers_wide(SkArgs a)
{
    __shared__ uint64_t s_tab[34];  // forward [0, 16), reverse [16, 32), init [32, 34)
    extern __shared__ uint64_t s_sv[];  // [W][64]
    sk_tables(a, s_tab, s_tab + 16, s_tab + 32);
    __syncthreads();
    const uint32_t lane = threadIdx.x;
    const uint32_t idx = blockIdx.x * 64u + lane;
    uint32_t n = a.n;
    if (a.n_dev) n = (uint32_t)min((uint64_t)n, *a.n_dev);
    if (a.fix_dev) {
        const uint32_t nf = *a.flag_count;
        if (nf > n && idx == 0) atomicOr(a.err, 1u);
        n = min(n, nf);
    }
    if (idx >= n) return;
    const uint32_t s = a.list ? a.list[idx] : idx;
    SkLane L;
    sk_init(L, a, s, s_tab + 32);
    SkStage S{};
    const uint32_t W = a.W;
    for (uint32_t i = 1; i < a.m; ++i) sk_roll(L, s_tab);
    for (uint32_t r = 0; r < W; ++r) {
        sk_roll(L, s_tab);
        s_sv[r * 64 + lane] = L.f ^ L.rv;
    }
    wide_backward(s_sv, W, lane);
    uint32_t bs = L.ts;
    uint64_t pv = 0;
    while (!L.done) {
        for (uint32_t r = 0; r < W && !L.done; ++r) {
            const uint64_t v = s_sv[r * 64 + lane];
            const uint64_t Hc = (r == 0 || v < pv) ? v : pv;
            sk_kmer<2, false, false>(L, a, S, bs + r, Hc, 0u);
            if (!L.done) {
                sk_roll(L, s_tab);
                const uint64_t h = L.f ^ L.rv;
                s_sv[r * 64 + lane] = h;
                if (r == 0 || h < pv) pv = h;
            }
        }
        if (!L.done) {
            wide_backward(s_sv, W, lane);
            bs += W;
        }
    }
    if (L.flagged) {
        const uint32_t slot = atomicAdd(a.flag_count, 1u);
        a.flag_list[slot] = s;
    }
    const uint32_t nsk = L.nsk + ((L.tail && L.t0p == L.t0) ? 1u : 0u);
    if (nsk) atomicAdd(&a.sk_count[L.job], (unsigned long long)nsk);
}

// The main pass of the same count with the window minimum kept by a monotone
// deque instead of the whole window: the m-mers of the window that no later
// one undercuts, in increasing value (an older m-mer of equal value is dropped:
// the cut rule reads the minimum's value only).  For hash values in random
// order it holds about ln W entries; a lane keeps up to WDQ of them in LDS
// ([entry][thread], 12 bytes each), so a CU runs 12 waves where the window
// kernel runs 2.  A lane whose deque would grow past a.dq_cap (m-mers whose
// values ascend, about one position in 10^6 for W = 107 and random sequence)
// stops with nothing counted and lists its strip in ovf_list; the window
// kernel counts the listed strips again with the same look-back.
constexpr uint32_t WDQ = WDQ_MAX, WDQ_BLOCK = 256;

__global__ __launch_bounds__(WDQ_BLOCK) void k_superkmers_wdq(SkArgs a)
{
    __shared__ uint64_t s_tab[34];  // forward [0, 16), reverse [16, 32), init [32, 34)
    __shared__ uint64_t s_dv[WDQ * WDQ_BLOCK];
    __shared__ uint32_t s_dp[WDQ * WDQ_BLOCK];
    sk_tables(a, s_tab, s_tab + 16, s_tab + 32);
    __syncthreads();
    const uint32_t tid = threadIdx.x;
    const uint32_t s = blockIdx.x * WDQ_BLOCK + tid;
    uint32_t n = a.n;
    if (a.n_dev) n = (uint32_t)min((uint64_t)n, *a.n_dev);
    uint32_t nsk = 0;
    if (s < n) {
        SkLane L;
        sk_init(L, a, s, s_tab + 32);
        SkStage S{};
        const uint32_t W = a.W;
        for (uint32_t i = 1; i < a.m; ++i) sk_roll(L, s_tab);
        // deque: entries [head, head + sz) mod WDQ; its front cached in (fv, fp)
        uint32_t head = 0, sz = 0, fp = 0;
        const uint32_t cap = min(max(a.dq_cap, 1u), WDQ);
        uint64_t fv = 0;
        bool over = false;
        auto push = [&](uint64_t h, uint32_t pos) {
            while (sz > 0u && s_dv[((head + sz - 1u) & (WDQ - 1u)) * WDQ_BLOCK + tid] >= h) --sz;
            if (sz == cap) {
                over = true;
                return;
            }
            const uint32_t e = ((head + sz) & (WDQ - 1u)) * WDQ_BLOCK + tid;
            s_dv[e] = h;
            s_dp[e] = pos;
            if (sz == 0u) {
                fv = h;
                fp = pos;
            }
            ++sz;
        };
        for (uint32_t r = 0; r < W && !over; ++r) {
            sk_roll(L, s_tab);
            push(L.f ^ L.rv, r);
        }
        // k-mer t of the walk (chunk k-mer L.ts + t) has the m-mers [t, t + W)
        for (uint32_t t = 0; !over;) {
            sk_kmer<2, false, false>(L, a, S, L.ts + t, fv, 0u);
            if (L.done) break;
            sk_roll(L, s_tab);
            push(L.f ^ L.rv, t + W);
            ++t;
            if (fp < t) {  // the front left the window (at most one per step)
                head = (head + 1u) & (WDQ - 1u);
                --sz;
                fv = s_dv[head * WDQ_BLOCK + tid];
                fp = s_dp[head * WDQ_BLOCK + tid];
            }
        }
        if (over) {
            // counted again with the same look-back by the window kernel
            const uint32_t slot = atomicAdd(a.ovf_count, 1u);
            a.ovf_list[slot] = s;
        } else {
            if (L.flagged) {
                const uint32_t slot = atomicAdd(a.flag_count, 1u);
                a.flag_list[slot] = s;
            }
            nsk = L.nsk + ((L.tail && L.t0p == L.t0) ? 1u : 0u);
        }
    }
    // one atomic per wave (a single job's total is one counter)
    for (uint32_t d = 32; d; d >>= 1) nsk += __shfl_xor(nsk, d, 64);
    if ((tid & 63u) == 0u && nsk) atomicAdd(&a.sk_count[0], (unsigned long long)nsk);
}

void launch_superkmers_wide(hipStream_t st, const SkArgs& a)
{
    if (a.n == 0) return;
    if (!a.list) {
        k_superkmers_wdq<<<(a.n + WDQ_BLOCK - 1u) / WDQ_BLOCK, WDQ_BLOCK, 0, st>>>(a);
        return;
    }
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void*)k_superkmers_wide, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(WIDE_WMAX * 64u * sizeof(uint64_t)));
    });
    const unsigned blocks = (a.n + 63u) / 64u;
    k_superkmers_wide<<<blocks, 64, (size_t)a.W * 64u * sizeof(uint64_t), st>>>(a);
}

// Base j of the packed stream sits at bits 31-2(j%16) of word j/16.
__device__ __forceinline__ uint32_t packed_base(const uint32_t* __restrict__ packed, uint64_t j)
{
    return (packed[j >> 4] >> (30u - 2u * (uint32_t)(j & 15u))) & 3u;
}

// The 32 bases from base s of the packed stream, base s in bits 63:62.  Reads
// words s/16 .. s/16+2 (the stream carries 8 words of padding past its end).
__device__ __forceinline__ uint64_t packed32(const uint32_t* __restrict__ packed, uint64_t s)
{
    const uint64_t q = s >> 4;
    const uint32_t sh = 2u * (uint32_t)(s & 15u);
    const uint64_t hi = ((uint64_t)packed[q] << 32) | packed[q + 1];
    const uint64_t lo = packed[q + 2];
    return (hi << sh) | ((lo << sh) >> 32);
}

// Multi-word keys (word 0 = the last 32 bases, Kmer<k>'s layout; the key ends
// in word tw = (2k-1)/64, whose low topb = 2k - 64 tw bits it uses, and the
// words above tw stay zero): push a base into the forward k-mer on the right,
// its complement into the reverse complement on the left at bit rtop =
// 2(k-1) - 64 tw of word tw.
template <int NW>
__device__ __forceinline__ void wide_roll(uint64_t (&f)[NW], uint64_t (&r)[NW], uint32_t b, uint32_t tw,
                                          uint64_t topmask, uint32_t rtop)
{
#pragma unroll
    for (int i = NW - 1; i > 0; --i) f[i] = (f[i] << 2) | (f[i - 1] >> 62);
    f[0] = (f[0] << 2) | b;
#pragma unroll
    for (int i = 0; i < NW; ++i) f[i] = (uint32_t)i > tw ? 0ull : ((uint32_t)i == tw ? f[i] & topmask : f[i]);
#pragma unroll
    for (int i = 0; i < NW - 1; ++i) r[i] = (r[i] >> 2) | (r[i + 1] << 62);
    r[NW - 1] >>= 2;
#pragma unroll
    for (int i = 0; i < NW; ++i)
        if ((uint32_t)i == tw) r[i] |= (uint64_t)(b ^ 3u) << rtop;
}

// The forward k-mer whose first base is base p of the packed stream, in the
// key layout: word w < tw holds the 32 bases from p + k - 32 - 32w, word tw
// the first k - 32 tw bases, right-aligned.
template <int NW>
__device__ __forceinline__ void wide_key_at(const uint32_t* __restrict__ packed, uint64_t p, uint32_t k,
                                            uint32_t tw, uint64_t (&y)[NW])
{
    const uint32_t nb = k - 32u * tw;  // 1..32
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if ((uint32_t)w < tw) y[w] = packed32(packed, p + k - 32u - 32u * (uint32_t)w);
        else if ((uint32_t)w == tw) y[w] = packed32(packed, p) >> (64u - 2u * nb);
        else y[w] = 0;
    }
}

// Reverse complement of a key: the 2-bit digits of the 64 NW-bit value are
// mirrored (words swapped, digits reversed inside each word), the value moved
// down by 64 NW - 2k bits and the k digits complemented.
template <int NW>
__device__ __forceinline__ void wide_rc(const uint64_t (&x)[NW], uint32_t k, uint32_t tw, uint64_t topmask,
                                        uint64_t (&y)[NW])
{
    uint64_t t[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        uint64_t v = __builtin_bitreverse64(x[w]);
        v = ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
        t[NW - 1 - w] = v;
    }
    const uint32_t D = 64u * NW - 2u * k, dw = D >> 6, db = D & 63u;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        uint64_t a = 0, b = 0;
#pragma unroll
        for (int j = 0; j < NW; ++j) {
            if ((uint32_t)j == (uint32_t)w + dw) a = t[j];
            if ((uint32_t)j == (uint32_t)w + dw + 1u) b = t[j];
        }
        const uint64_t v = (a >> db) | (db ? (b << (64u - db)) : 0ull);
        y[w] = (uint32_t)w < tw ? ~v : ((uint32_t)w == tw ? (~v & topmask) : 0ull);
    }
}

template <int NW>
__device__ __forceinline__ bool wide_le(const uint64_t (&x)[NW], const uint64_t (&y)[NW])
{
#pragma unroll
    for (int i = NW - 1; i >= 0; --i)
        if (x[i] != y[i]) return x[i] < y[i];
    return true;
}

template <int NW>
__device__ __forceinline__ bool wide_eq(const uint64_t (&x)[NW], const uint64_t (&y)[NW])
{
    bool e = true;
#pragma unroll
    for (int i = 0; i < NW; ++i) e = e && x[i] == y[i];
    return e;
}

// The k-mer table holds where a key is, not the key: slot = {tag, count - 1},
// two words.  tag = fingerprint (bits 63:41, the top of the key's hash) | bit
// 40 (the canonical key is the reverse complement of the bases at the
// position) | position + 1 (bits 39:0, base index of the k-mer's first base in
// the packed stream).  A tag is claimed with one 64-bit compare-and-swap and
// never changes afterwards; a key is compared with a slot's by reading the
// slot's bases from the packed stream, which no kernel writes while the table
// is filled, and every later occurrence adds 1 with a relaxed atomic (the
// claim itself adds nothing: the count word holds count - 1).  No slot state
// is published in parts, so neither side needs an acquire or release.
constexpr uint32_t WIDE_POS_BITS = 40;
constexpr uint64_t WIDE_POS_MASK = (1ull << WIDE_POS_BITS) - 1ull;
constexpr uint64_t WIDE_FP_MASK = ~((1ull << (WIDE_POS_BITS + 1)) - 1ull);
// A key that finds no slot within WIDE_PROBES sets *err (the host then counts
// again with a table twice as large).
constexpr uint32_t WIDE_PROBES = 4096;

struct __align__(16) WideSlot {
    uint64_t tag;
    uint64_t cnt;  // low 32 bits: the count - 1 (of an occupied slot)
};

template <int NW>
__device__ __forceinline__ bool wide_insert(const WideArgs& w, const uint64_t (&c)[NW], const uint64_t (&o)[NW],
                                            uint64_t pos, uint32_t orient, uint32_t tw)
{
    uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
    for (int i = 0; i < NW; ++i) h = mix64(h ^ c[i]);
    const uint64_t fp = h & WIDE_FP_MASK;
    const uint64_t mine = fp | ((uint64_t)orient << WIDE_POS_BITS) | (pos + 1u);
    uint64_t slot = h & w.mask;
    for (uint32_t p = 0; p <= w.mask && p < WIDE_PROBES; ++p) {
        uint64_t* tp = w.tab + 2u * slot;
        uint64_t v = __hip_atomic_load(tp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v == 0u) {
            v = atomicCAS((unsigned long long*)tp, 0ull, (unsigned long long)mine);
            if (v == 0u) return true;
        }
        if (((v ^ fp) & WIDE_FP_MASK) == 0u) {
            uint64_t y[NW];
            wide_key_at<NW>(w.packed, (v & WIDE_POS_MASK) - 1u, w.k, tw, y);
            if ((v >> WIDE_POS_BITS) & 1u ? wide_eq<NW>(y, o) : wide_eq<NW>(y, c)) {
                atomicAdd(reinterpret_cast<uint32_t*>(tp + 1), 1u);
                return true;
            }
        }
        slot = (slot + 1u) & w.mask;
    }
    atomicOr(w.err, 1u);
    return false;
}

// One lane per strip: the strip's k-mers [tb, te) of its chunk, the first read
// from the packed stream (its reverse complement computed), the others rolled
// one base at a time.
template <int NW>
__global__ __launch_bounds__(256) void k_wide_insert(WideArgs w)
{
    const uint64_t n = min((uint64_t)w.n_strips, *w.n_dev);
    const uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = w.strip_chunk[s];
    const uint64_t cb = w.off[c];
    const uint64_t K = w.off[c + 1] - cb - w.k + 1u;
    const uint64_t tb = w.strip_tb[s];
    const uint64_t te = min(tb + w.S, K);
    const uint32_t tw = (2u * w.k - 1u) / 64u;
    const uint32_t topb = 2u * w.k - 64u * tw;  // 1..64
    const uint64_t topmask = topb >= 64u ? ~0ull : ((1ull << topb) - 1u);
    const uint32_t rtop = topb - 2u;
    uint64_t f[NW], r[NW];
    uint64_t g = cb + tb;  // first base of k-mer t
    wide_key_at<NW>(w.packed, g, w.k, tw, f);
    wide_rc<NW>(f, w.k, tw, topmask, r);
    for (uint64_t t = tb; t < te; ++t, ++g) {
        if (t != tb) wide_roll<NW>(f, r, packed_base(w.packed, g + w.k - 1u), tw, topmask, rtop);
        if (((t - tb) & 31u) == 0u && __hip_atomic_load(w.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        const bool fwd = wide_le<NW>(f, r);
        const bool ok = fwd ? wide_insert<NW>(w, f, r, g, 0u, tw) : wide_insert<NW>(w, r, f, g, 1u, tw);
        if (!ok) return;
    }
}

// Drain, two passes over equal block ranges of the table (slot order, so the
// result order is fixed): per-block totals of written (ci <= count <= cx) and
// distinct keys, then, at the scanned block offsets, every written key with
// onw = ceil(k/32) words (the reference Kmer<k> layout) and its count.
__global__ __launch_bounds__(256) void k_wide_drain_count(WideArgs w, uint64_t per_block, uint32_t ci, uint64_t cx,
                                                          uint64_t* __restrict__ blk_keep,
                                                          uint64_t* __restrict__ blk_occ)
{
    __shared__ uint32_t s_k[4], s_o[4];
    const WideSlot* tab = reinterpret_cast<const WideSlot*>(w.tab);
    const uint64_t b0 = blockIdx.x * per_block;
    uint32_t keep = 0, occ = 0;
    for (uint64_t i = b0 + threadIdx.x; i < b0 + per_block; i += 256u) {
        const WideSlot e = tab[i];
        const uint32_t c = (uint32_t)e.cnt + 1u;
        occ += e.tag != 0u ? 1u : 0u;
        keep += (e.tag != 0u && c >= ci && (uint64_t)c <= cx) ? 1u : 0u;
    }
    for (uint32_t d = 32; d; d >>= 1) {
        keep += __shfl_xor(keep, d, 64);
        occ += __shfl_xor(occ, d, 64);
    }
    if ((threadIdx.x & 63u) == 0u) {
        s_k[threadIdx.x >> 6] = keep;
        s_o[threadIdx.x >> 6] = occ;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_keep[blockIdx.x] = (uint64_t)s_k[0] + s_k[1] + s_k[2] + s_k[3];
        blk_occ[blockIdx.x] = (uint64_t)s_o[0] + s_o[1] + s_o[2] + s_o[3];
    }
}

// Each thread takes WIDE_PT consecutive slots per step; one block scan of the
// per-thread totals places them, the written slots of the step are listed in
// LDS in slot order, and the block rebuilds their keys one per thread, so the
// base reads of a step are in flight together.
constexpr uint32_t WIDE_PT = 8;

__device__ __forceinline__ uint32_t block_excl_u32(uint32_t v, uint32_t* sh, uint32_t& total);

template <int NW>
__global__ __launch_bounds__(256) void k_wide_drain_write(WideArgs w, uint64_t per_block, uint32_t ci, uint64_t cx,
                                                          uint32_t onw, const uint64_t* __restrict__ blk_off,
                                                          uint64_t* __restrict__ out_keys,
                                                          uint32_t* __restrict__ out_cnt)
{
    __shared__ uint32_t s_red[4];
    __shared__ uint64_t s_tag[256 * WIDE_PT];
    __shared__ uint32_t s_cnt[256 * WIDE_PT];
    const WideSlot* tab = reinterpret_cast<const WideSlot*>(w.tab);
    const uint32_t tw = (2u * w.k - 1u) / 64u;
    const uint32_t topb = 2u * w.k - 64u * tw;
    const uint64_t topmask = topb >= 64u ? ~0ull : ((1ull << topb) - 1u);
    uint64_t base = blk_off[blockIdx.x];
    const uint64_t b0 = blockIdx.x * per_block;
    for (uint64_t i0 = b0; i0 < b0 + per_block; i0 += 256u * WIDE_PT) {
        const uint64_t s0 = i0 + (uint64_t)threadIdx.x * WIDE_PT;
        WideSlot e[WIDE_PT];
#pragma unroll
        for (uint32_t j = 0; j < WIDE_PT; ++j) e[j] = tab[s0 + j];
        uint32_t keep = 0, nk = 0;
#pragma unroll
        for (uint32_t j = 0; j < WIDE_PT; ++j) {
            const uint32_t c = (uint32_t)e[j].cnt + 1u;
            const bool kp = e[j].tag != 0u && c >= ci && (uint64_t)c <= cx;
            keep |= (kp ? 1u : 0u) << j;
            nk += kp ? 1u : 0u;
        }
        uint32_t tot;
        uint32_t at = block_excl_u32(nk, s_red, tot);  // ends with a barrier
#pragma unroll
        for (uint32_t j = 0; j < WIDE_PT; ++j) {
            if ((keep >> j) & 1u) {
                s_tag[at] = e[j].tag;
                s_cnt[at] = (uint32_t)e[j].cnt + 1u;
                ++at;
            }
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < tot; q += 256u) {
            const uint64_t tag = s_tag[q];
            uint64_t y[NW], z[NW];
            wide_key_at<NW>(w.packed, (tag & WIDE_POS_MASK) - 1u, w.k, tw, y);
            wide_rc<NW>(y, w.k, tw, topmask, z);
            const bool rcf = (tag >> WIDE_POS_BITS) & 1u;
            const uint64_t o = base + q;
#pragma unroll
            for (int r = 0; r < NW; ++r)
                if ((uint32_t)r < onw) out_keys[onw * o + r] = rcf ? z[r] : y[r];
            out_cnt[o] = s_cnt[q];
        }
        base += tot;
        __syncthreads();  // the list is rewritten by the next step
    }
}

void launch_wide_insert(hipStream_t st, const WideArgs& w, uint32_t nw)
{
    if (w.n_strips == 0) return;
    const unsigned blocks = (unsigned)((w.n_strips + 255u) / 256u);
    if (nw == 4) k_wide_insert<4><<<blocks, 256, 0, st>>>(w);
    else k_wide_insert<8><<<blocks, 256, 0, st>>>(w);
}

uint32_t wide_drain_blocks(uint64_t slots)
{
    // slots is a power of two >= 2^16: every block gets a multiple of
    // 256 WIDE_PT slots
    return (uint32_t)std::min<uint64_t>(slots / (256u * WIDE_PT), 4096u);
}

void launch_wide_drain_count(hipStream_t st, const WideArgs& w, uint32_t ci, uint64_t cx, uint64_t* blk_keep,
                             uint64_t* blk_occ)
{
    const uint32_t nb = wide_drain_blocks(w.mask + 1u);
    k_wide_drain_count<<<nb, 256, 0, st>>>(w, (w.mask + 1u) / nb, ci, cx, blk_keep, blk_occ);
}

void launch_wide_drain_write(hipStream_t st, const WideArgs& w, uint32_t nw, uint32_t ci, uint64_t cx, uint32_t onw,
                             const uint64_t* blk_off, uint64_t* out_keys, uint32_t* out_cnt)
{
    const uint32_t nb = wide_drain_blocks(w.mask + 1u);
    const uint64_t per = (w.mask + 1u) / nb;
    if (nw == 4) k_wide_drain_write<4><<<nb, 256, 0, st>>>(w, per, ci, cx, onw, blk_off, out_keys, out_cnt);
    else k_wide_drain_write<8><<<nb, 256, 0, st>>>(w, per, ci, cx, onw, blk_off, out_keys, out_cnt);
}

// =========================================================== routing ====
// Records are read run by run; a tile is a fixed range of the concatenated
// run space.  Histograms live in LDS; every tile writes one row of the
// [tile][sub-partition] matrix.
constexpr uint32_t ROUTE_T = 1024;

// Run descriptors cached per tile (LDS: 16 KB).  With 8192 sub-partitions the
// histogram block (64 KiB of counters + this cache) must stay within half of
// the CU's 160 KiB, so that two blocks share a CU.
constexpr uint32_t ROUTE_RUNS = 1008;

__device__ __forceinline__ uint32_t block_excl_u32(uint32_t v, uint32_t* sh, uint32_t& total);

// Tile geometry: records per tile and live tiles.  With a device-side record
// total the tile is the smallest power of two >= 16384 that keeps the tile
// count within max_tiles (the rule the host applies for exact launches).
struct RouteGeom {
    uint64_t n_rec;
    uint32_t tile, ntiles;
};

__device__ __forceinline__ RouteGeom route_geom(const RouteArgs& a)
{
    RouteGeom g;
    if (a.n_rec_dev) {
        g.n_rec = *a.n_rec_dev;
        if (g.n_rec > a.rec_cap) {
            // more records than the partition buffers hold: route nothing
            if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(a.err, 2u);
            g.n_rec = 0;
        }
        uint64_t t = 16384;
        while ((g.n_rec + t - 1) / t > a.max_tiles) t <<= 1;
        g.tile = (uint32_t)t;
        g.ntiles = (uint32_t)((g.n_rec + t - 1) / t);
    } else {
        g.n_rec = a.n_rec;
        g.tile = a.tile;
        g.ntiles = gridDim.x;
    }
    return g;
}

// First r in [l, h) with run_off[r + d] >= x, or h when there is none
// (run_off does not decrease).  Wave 0 searches 64-ary: every round probes 64
// evenly spaced indices at once, so a few thousand runs take three rounds of
// loads instead of a dozen dependent ones.
__device__ __forceinline__ uint32_t wave_run_search(const uint64_t* run_off, uint32_t d, uint64_t x,
                                                    uint32_t l, uint32_t h)
{
    const uint32_t lane = threadIdx.x & 63u;
    while (l < h) {
        const uint64_t span = h - l;
        const uint32_t p = l + (uint32_t)(((uint64_t)(lane + 1u) * span) / 65u);
        const uint64_t hit = __ballot(p < h && run_off[p + d] >= x);
        if (!hit) {
            l = __builtin_amdgcn_readlane(p, 63) + 1u;
        } else {
            const uint32_t j = (uint32_t)__builtin_ctzll(hit);
            h = __builtin_amdgcn_readlane(p, j);
            if (j) l = __builtin_amdgcn_readlane(p, j - 1u) + 1u;
        }
    }
    return l;
}

// The runs of the tile [lo, hi): s_meta[0] = the first run that ends after lo,
// s_meta[1] = the runs cached from there (at most cap - 1), s_meta[2] = 1 when
// the tile has more runs than that.  Ends with a barrier.
__device__ __forceinline__ void tile_runs(const RouteArgs& a, uint64_t lo, uint64_t hi, uint32_t cap,
                                          uint32_t* s_meta)
{
    if (threadIdx.x < 64) {
        const uint32_t l = wave_run_search(a.run_off, 1u, lo + 1u, 0u, a.n_runs);
        const uint32_t ef = wave_run_search(a.run_off, 0u, hi, l, a.n_runs);
        const uint32_t e = min(ef, l + cap - 1u);
        if (threadIdx.x == 0) {
            s_meta[0] = l;
            s_meta[1] = e - l;
            s_meta[2] = ef > e ? 1u : 0u;
        }
    }
    __syncthreads();
}

// Calls f(record) for every record of this block's tile.  The tile's run
// descriptors (record index of the run's first record minus its offset in the
// run space) are cached in LDS; each wave walks 64*U consecutive records at a
// time and every lane finds its run by a short forward scan.
template <int RW, class F>
__device__ __forceinline__ void for_tile_records(const RouteArgs& a, const RouteGeom& g,
                                                 uint64_t* s_off, int64_t* s_base,
                                                 uint32_t* s_meta, F&& f)
{
    const uint64_t lo = (uint64_t)blockIdx.x * g.tile;
    const uint64_t hi = min(lo + g.tile, g.n_rec);
    tile_runs(a, lo, hi, ROUTE_RUNS, s_meta);
    const uint32_t r0 = s_meta[0], nr = s_meta[1];
    for (uint32_t i = threadIdx.x; i <= nr; i += blockDim.x) {
        s_off[i] = a.run_off[r0 + i];
        if (i < nr) s_base[i] = (int64_t)a.run_start[r0 + i] - (int64_t)a.run_off[r0 + i];
    }
    __syncthreads();
    if (s_meta[2]) {
        // many tiny runs: walk them one by one
        for (uint32_t r = r0; r < a.n_runs && a.run_off[r] < hi; ++r) {
            const uint64_t b = max(a.run_off[r], lo), e = min(a.run_off[r + 1], hi);
            const int64_t base = (int64_t)a.run_start[r] - (int64_t)a.run_off[r];
            for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x)
                f(load_rec<RW>(RW == 1 ? a.metas : a.recs, (uint64_t)(base + (int64_t)i)));
        }
        return;
    }
    constexpr uint32_t U = RW == 8 ? 4 : (RW == 1 ? 16 : 8);  // loads in flight per lane
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    uint32_t rr = 0;  // run of this lane's last record (non-decreasing)
    // slots first (run lookups), then all U loads back to back: lanes past the
    // end load slot 0, which always exists
    for (uint64_t base = lo + (uint64_t)wv * 64u * U; base < hi; base += (uint64_t)nwv * 64u * U) {
        uint64_t at[U];
        bool ok[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint64_t i = base + u * 64u + lane;
            ok[u] = i < hi;
            at[u] = 0;
            if (ok[u]) {
                while (s_off[rr + 1] <= i) ++rr;
                at[u] = (uint64_t)(s_base[rr] + (int64_t)i);
            }
        }
        Rec<RW> v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) v[u] = load_rec<RW>(RW == 1 ? a.metas : a.recs, at[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (ok[u]) f(v[u]);
    }
}

// RW is the record kind (2 slim, 4/8 fat): it fixes the length field's width;
// the histogram reads only the meta array.
template <int RW>
__global__ __launch_bounds__(ROUTE_T) void k_route_hist(RouteArgs a)
{
    // one 64-bit counter per sub-partition: superkmers (low half) and k-mers
    // (high half), so a record costs one LDS atomic
    extern __shared__ unsigned long long sh64[];
    __shared__ uint64_t s_off[ROUTE_RUNS + 1];
    __shared__ int64_t s_base[ROUTE_RUNS];
    __shared__ uint32_t s_meta[4];
    const RouteGeom g = route_geom(a);
    if (blockIdx.x >= g.ntiles) return;
    const uint32_t nsp = a.nsp;
    for (uint32_t i = threadIdx.x; i < nsp; i += blockDim.x) sh64[i] = 0ull;
    __syncthreads();
    const uint32_t shift = REC_HSEL_SHIFT + a.sb;
    const uint32_t k1 = a.k - 1u;
    for_tile_records<1>(a, g, s_off, s_base, s_meta, [&](const Rec<1>& v) {
        const uint32_t meta = v.w[0];
        atomicAdd(&sh64[meta >> shift], 1ull | ((unsigned long long)(rec_len<RW>(meta) - k1) << 32));
    });
    __syncthreads();
    uint32_t* hs = a.hsk + (uint64_t)blockIdx.x * nsp;
    uint32_t* hk = a.hkm + (uint64_t)blockIdx.x * nsp;
    for (uint32_t i = threadIdx.x; i < nsp; i += blockDim.x) {
        const unsigned long long v = sh64[i];
        hs[i] = (uint32_t)v;
        hk[i] = (uint32_t)(v >> 32);
    }
}

template <int RW>
__global__ __launch_bounds__(ROUTE_T) void k_route_coarse(RouteArgs a, uint32_t log2f, uint32_t* tmp);
template <int RW>
__global__ __launch_bounds__(ROUTE_T) void k_route_fine(RouteArgs a, uint32_t log2f, const uint32_t* tmp,
                                                      uint32_t nr);

// Dynamic-LDS limits of the routing kernels, set once per process at their
// largest sizes, so no attribute call happens while a launch sequence is
// being captured: the histogram at 2^14 sub-partitions, the fine scatter at
// 3 arrays of F = 2^HSEL_BITS / 256 per-sub-partition words.
constexpr uint32_t FINE_F_MAX = 1u << (HSEL_BITS - 8);

template <int RW>
void set_route_attributes()
{
    const int hist = (int)(2u * (1u << 14) * sizeof(uint32_t));
    const int cur = (int)(3u * FINE_F_MAX * sizeof(uint32_t));
    (void)hipFuncSetAttribute((const void*)k_route_hist<RW>, hipFuncAttributeMaxDynamicSharedMemorySize, hist);
    (void)hipFuncSetAttribute((const void*)k_route_fine<RW>, hipFuncAttributeMaxDynamicSharedMemorySize, cur);
}

void set_kernel_attributes()
{
    static std::once_flag once;
    std::call_once(once, [] {
        set_route_attributes<2>();
        set_route_attributes<4>();
        set_route_attributes<8>();
    });
}

void launch_route_hist(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint32_t rw)
{
    if (ntiles == 0) return;
    set_kernel_attributes();
    const size_t lds = 2u * a.nsp * sizeof(uint32_t);
    if (rw == 2) k_route_hist<2><<<ntiles, ROUTE_T, lds, st>>>(a);
    else if (rw == 4) k_route_hist<4><<<ntiles, ROUTE_T, lds, st>>>(a);
    else k_route_hist<8><<<ntiles, ROUTE_T, lds, st>>>(a);
}

// Per sub-partition: tile-exclusive offsets (in place) and totals.  Eight
// tiles per step keep eight loads in flight per thread.
__global__ __launch_bounds__(256) void k_route_colscan(RouteArgs a, uint32_t ntiles,
                                                     uint64_t* __restrict__ part_sk,
                                                     uint64_t* __restrict__ part_km)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nsp) return;
    if (a.n_rec_dev) ntiles = route_geom(a).ntiles;
    uint32_t* __restrict__ hs = a.hsk;
    const uint32_t* __restrict__ hk = a.hkm;
    uint64_t run = 0, km = 0;
    uint32_t t = 0;
    for (; t + 8 <= ntiles; t += 8) {
        uint32_t v[8], w[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v[j] = hs[(uint64_t)(t + j) * a.nsp + p];
            w[j] = hk[(uint64_t)(t + j) * a.nsp + p];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            hs[(uint64_t)(t + j) * a.nsp + p] = (uint32_t)run;
            run += v[j];
            km += w[j];
        }
    }
    for (; t < ntiles; ++t) {
        const uint64_t i = (uint64_t)t * a.nsp + p;
        const uint32_t v = hs[i];
        hs[i] = (uint32_t)run;
        run += v;
        km += hk[i];
    }
    part_sk[p] = run;
    part_km[p] = km;
}

void launch_route_colscan(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint64_t* part_sk,
                          uint64_t* part_km)
{
    k_route_colscan<<<(a.nsp + 255u) / 256u, 256, 0, st>>>(a, ntiles, part_sk, part_km);
}

// Scatter, level 1: every tile writes its records into coarse buckets of F
// consecutive sub-partitions (at most 256 buckets).  The tile goes through its
// records in chunks that fit in LDS (128 KiB of fat records, one workgroup per
// CU; 64 KiB of slim ones); a chunk is grouped by bucket in LDS (histogram,
// scan, place) and then written out bucket by bucket, so consecutive threads
// write consecutive addresses of one bucket.  Longer runs per bucket write
// faster: runs of 512 B move 16 B records at about 3.5 TB/s, runs of 256 B
// at about 3.2 TB/s.
constexpr uint32_t COARSE_RUNS = 512;  // run descriptors cached per tile

template <int RW>
__global__ __launch_bounds__(ROUTE_T) void k_route_coarse(RouteArgs a, uint32_t log2f, uint32_t* tmp)
{
    constexpr uint32_t CH = (RW == 2 ? 16384u : 32768u) / RW;  // records per chunk
    constexpr uint32_t PER = CH / ROUTE_T;
    __shared__ Rec<RW> s_rec[CH];
    __shared__ uint32_t s_hist[256], s_start[256], s_cur[256], s_ord[256];
    __shared__ uint64_t s_off[COARSE_RUNS + 1];
    __shared__ int64_t s_base[COARSE_RUNS];
    __shared__ uint32_t s_meta[4];
    __shared__ uint32_t s_red[ROUTE_T / 64];
    const RouteGeom g = route_geom(a);
    if (blockIdx.x >= g.ntiles) return;
    const uint32_t nc = a.nsp >> log2f, F = 1u << log2f;
    const uint32_t tid = threadIdx.x;
    const uint32_t* hs = a.hsk + (uint64_t)blockIdx.x * a.nsp;
    for (uint32_t c = tid; c < nc; c += blockDim.x) {
        uint32_t o = (uint32_t)a.part_off[c * F];
        for (uint32_t f = 0; f < F; ++f) o += hs[c * F + f];
        s_cur[c] = o;
        s_ord[c] = a.split_flags ? a.split_flags[c] : 0u;
    }
    const uint64_t lo = (uint64_t)blockIdx.x * g.tile;
    const uint64_t hi = min(lo + g.tile, g.n_rec);
    tile_runs(a, lo, hi, COARSE_RUNS, s_meta);
    const uint32_t r0 = s_meta[0], nr = s_meta[1];
    const bool cached = s_meta[2] == 0;
    for (uint32_t i = tid; i <= nr; i += blockDim.x) {
        s_off[i] = a.run_off[r0 + i];
        if (i < nr) s_base[i] = (int64_t)a.run_start[r0 + i] - (int64_t)a.run_off[r0 + i];
    }
    __syncthreads();
    const uint32_t shift = REC_HSEL_SHIFT + a.sb + log2f;
    uint32_t rr = 0;  // cached run of this thread's last record (indices only grow)
    for (uint64_t cb = lo; cb < hi; cb += CH) {
        for (uint32_t c = tid; c < nc; c += blockDim.x) s_hist[c] = 0;
        __syncthreads();
        // record slots first (run lookups), then all PER loads back to back
        // (lanes past the end load record 0, which always exists), then the
        // histogram
        uint64_t at[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint64_t i = cb + j * ROUTE_T + tid;
            at[j] = 0;
            if (i < hi) {
                int64_t base;
                if (cached) {
                    while (s_off[rr + 1] <= i) ++rr;
                    base = s_base[rr];
                } else {
                    uint32_t l = 0, h = a.n_runs;
                    while (l < h) {
                        const uint32_t mid = (l + h) >> 1;
                        if (a.run_off[mid + 1] > i) h = mid;
                        else l = mid + 1;
                    }
                    base = (int64_t)a.run_start[l] - (int64_t)a.run_off[l];
                }
                at[j] = (uint64_t)(base + (int64_t)i);
            }
        }
        Rec<RW> v[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) v[j] = load_rec<RW>(a.recs, at[j]);
        uint32_t bk[PER], pos[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            bk[j] = 0xFFFFFFFFu;
            if (cb + j * ROUTE_T + tid < hi) {
                bk[j] = rec_meta<RW>(v[j]) >> shift;
                pos[j] = atomicAdd(&s_hist[bk[j]], 1u);
            }
        }
        __syncthreads();
        uint32_t tot;
        const uint32_t ex = block_excl_u32(tid < nc ? s_hist[tid] : 0u, s_red, tot);
        if (tid < nc) s_start[tid] = ex;
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j)
            if (bk[j] != 0xFFFFFFFFu) s_rec[s_start[bk[j]] + pos[j]] = v[j];
        __syncthreads();
        const uint32_t nch = (uint32_t)min((uint64_t)CH, hi - cb);
        for (uint32_t i = tid; i < nch; i += ROUTE_T) {
            const Rec<RW> r = s_rec[i];
            const uint32_t meta = rec_meta<RW>(r);
            const uint32_t b = meta >> shift;
            const uint64_t at = (uint64_t)s_cur[b] + (i - s_start[b]);
            store_rec<RW>(tmp, at, r);
            if (s_ord[b]) a.tmp_meta[at] = meta;  // the fine pass sizes these buckets' groups
        }
        __syncthreads();
        if (tid < nc) s_cur[tid] += s_hist[tid];
        __syncthreads();
    }
}

// Group counts of the record ranges of the split-ordered coarse buckets, from
// their coarse-ordered meta words: workgroup (c, r) counts range r of bucket c
// (the bucket cut into FINE_RANGES near-equal ranges, as the fine scatter cuts
// it) into fine_hist[((c << gbits) + g) * FINE_RANGES + r].
__global__ __launch_bounds__(ROUTE_T) void k_fine_hist(RouteArgs a, uint32_t log2f)
{
    constexpr uint32_t R = FINE_RANGES;
    __shared__ uint32_t hist[FINE_F_MAX];
    // range-major: consecutive workgroups (spread over the XCDs) take
    // different buckets
    const uint32_t nc = a.nsp >> log2f;
    const uint32_t c = blockIdx.x % nc, r = blockIdx.x / nc, tid = threadIdx.x;
    if (!a.split_flags[c] || !a.split_bits) return;
    const uint32_t gbits = log2f + a.split_bits, G = 1u << gbits, gm = G - 1u;
    const uint32_t shift = REC_HSEL_SHIFT + a.sb - a.split_bits;
    const uint64_t b = a.part_off[c << log2f], e = a.part_off[(c + 1) << log2f];
    const uint64_t rs = (e - b + R - 1) / R;
    const uint64_t lo = min(e, b + r * rs), hi = min(e, lo + rs);
    for (uint32_t g = tid; g < G; g += blockDim.x) hist[g] = 0;
    __syncthreads();
    constexpr uint32_t U = 8;  // loads in flight per thread
    for (uint64_t i0 = lo + tid; i0 < hi; i0 += U * ROUTE_T) {
        // unconditional loads (past the end: the range's first word)
        uint32_t mv[U];
#pragma unroll
        for (uint32_t j = 0; j < U; ++j) {
            const uint64_t i = i0 + (uint64_t)j * ROUTE_T;
            mv[j] = a.tmp_meta[i < hi ? i : lo];
        }
#pragma unroll
        for (uint32_t j = 0; j < U; ++j)
            if (i0 + (uint64_t)j * ROUTE_T < hi) atomicAdd(&hist[(mv[j] >> shift) & gm], 1u);
    }
    __syncthreads();
    uint32_t* out = a.fine_hist + ((uint64_t)c << gbits) * R + r;
    for (uint32_t g = tid; g < G; g += blockDim.x) out[(uint64_t)g * R] = hist[g];
}

// Scatter, level 2: the records of a coarse bucket are put in groups: the
// sub-partition, then the top split_bits of the split field (G = F <<
// split_bits <= 512 groups).  The group sizes are known before any record
// moves, so every group is written contiguously at its final place: a bucket
// that is not split-ordered has the sub-partitions' starts; a split-ordered
// one counts its meta words, either per record range in k_fine_hist (nr > 1:
// workgroup (c, r) moves range r of bucket c, its group cursors being the
// group starts plus the group's counts in the ranges before r) or, with one
// workgroup per bucket (nr == 1), in a first pass of its own.  Records are
// then read in chunks that fit in LDS (128 KiB), grouped there (histogram,
// scan, place) and written out as runs, so consecutive threads write
// consecutive addresses.  A bucket's records are thus ordered by split group,
// and a work item that counts 1/2^i of a bucket reads one contiguous range.
template <int RW>
__global__ __launch_bounds__(ROUTE_T) void k_route_fine(RouteArgs a, uint32_t log2f, const uint32_t* tmp,
                                                      uint32_t nr)
{
    constexpr uint32_t CH = 32768u / RW;  // records per chunk (128 KiB): >= 16 per group
    constexpr uint32_t PER = CH / ROUTE_T;
    __shared__ Rec<RW> s_rec[CH];
    __shared__ uint32_t s_red[ROUTE_T / 64];
    extern __shared__ uint32_t s_dyn[];  // G cursors, G counts, G group starts
    // range-major: the first nc workgroups take range 0 of every bucket (a
    // bucket that is not split-ordered is all range 0), so consecutive
    // workgroups, which the dispatcher spreads over the XCDs, take different
    // buckets
    const uint32_t nc = a.nsp >> log2f;
    const uint32_t F = 1u << log2f, c = blockIdx.x % nc, r = blockIdx.x / nc, tid = threadIdx.x;
    // split groups only in coarse buckets holding a bucket the work items
    // will split (split_flags); elsewhere the groups are the sub-partitions,
    // whose starts are known, and one workgroup moves the bucket
    const uint32_t sbits = a.split_flags && a.split_flags[c] ? a.split_bits : 0u;
    if (r && !sbits) return;
    const uint32_t gbits = log2f + sbits, G = 1u << gbits;
    uint32_t* cur = s_dyn;
    uint32_t* hist = s_dyn + G;
    uint32_t* gst = s_dyn + 2 * G;
    const uint64_t b = a.part_off[c << log2f], e = a.part_off[(c + 1) << log2f];
    const uint32_t shift = REC_HSEL_SHIFT + a.sb - sbits;
    const uint32_t gm = G - 1u;
    uint64_t lo = b, hi = e;  // the records this workgroup moves
    if (!sbits) {
        for (uint32_t f = tid; f < F; f += blockDim.x) cur[f] = (uint32_t)a.part_off[(c << log2f) + f];
    } else if (nr > 1) {
        const uint64_t rs = (e - b + nr - 1) / nr;
        lo = min(e, b + r * rs);
        hi = min(e, lo + rs);
        const uint32_t* h = a.fine_hist + ((uint64_t)c << gbits) * nr;
        uint32_t carry = 0;
        for (uint32_t g0 = 0; g0 < G; g0 += blockDim.x) {
            const uint32_t g = g0 + tid;
            uint32_t all = 0, before = 0;
            if (g < G)
                for (uint32_t q = 0; q < nr; ++q) {
                    const uint32_t v = h[(uint64_t)g * nr + q];
                    all += v;
                    before += q < r ? v : 0u;
                }
            uint32_t tot;
            const uint32_t ex = block_excl_u32(g < G ? all : 0u, s_red, tot);
            if (g < G) {
                cur[g] = (uint32_t)b + carry + ex + before;
                if (r == 0) a.split_off[((uint64_t)c << gbits) + g] = b + carry + ex;
            }
            carry += tot;
            __syncthreads();
        }
        if (r == 0 && tid == 0) a.split_off[(uint64_t)(c + 1) << gbits] = e;
    } else {
        // group sizes of the whole region (from the meta words), then their
        // absolute starts; the end of the region closes its last group
        for (uint32_t g = tid; g < G; g += blockDim.x) hist[g] = 0;
        __syncthreads();
        constexpr uint32_t U = 8;  // loads in flight per thread
        for (uint64_t i0 = b + tid; i0 < e; i0 += U * ROUTE_T) {
            // unconditional loads (past the end: the region's first word), so
            // that all U are in flight together
            uint32_t mv[U];
#pragma unroll
            for (uint32_t j = 0; j < U; ++j) {
                const uint64_t i = i0 + (uint64_t)j * ROUTE_T;
                mv[j] = a.tmp_meta[i < e ? i : b];
            }
#pragma unroll
            for (uint32_t j = 0; j < U; ++j)
                if (i0 + (uint64_t)j * ROUTE_T < e) atomicAdd(&hist[(mv[j] >> shift) & gm], 1u);
        }
        __syncthreads();
        uint32_t carry = 0;
        for (uint32_t g0 = 0; g0 < G; g0 += blockDim.x) {
            const uint32_t g = g0 + tid;
            uint32_t tot;
            const uint32_t ex = block_excl_u32(g < G ? hist[g] : 0u, s_red, tot);
            if (g < G) {
                cur[g] = (uint32_t)b + carry + ex;
                a.split_off[((uint64_t)c << gbits) + g] = b + carry + ex;
            }
            carry += tot;
            __syncthreads();
        }
        if (tid == 0) a.split_off[(uint64_t)(c + 1) << gbits] = e;
    }
    __syncthreads();
    for (uint64_t cb = lo; cb < hi; cb += CH) {
        for (uint32_t g = tid; g < G; g += blockDim.x) hist[g] = 0;
        __syncthreads();
        // all PER loads back to back (lanes past the end load the chunk's
        // first record), then the histogram
        Rec<RW> v[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint64_t i = cb + j * ROUTE_T + tid;
            v[j] = load_rec<RW>(tmp, i < hi ? i : cb);
        }
        uint32_t fk[PER], pos[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            fk[j] = 0xFFFFFFFFu;
            if (cb + j * ROUTE_T + tid < hi) {
                fk[j] = (rec_meta<RW>(v[j]) >> shift) & gm;
                pos[j] = atomicAdd(&hist[fk[j]], 1u);
            }
        }
        __syncthreads();
        // exclusive scan of the G counts (G <= 1024: in steps of blockDim.x)
        uint32_t carry = 0;
        for (uint32_t g0 = 0; g0 < G; g0 += blockDim.x) {
            const uint32_t g = g0 + tid;
            uint32_t tot;
            const uint32_t ex = block_excl_u32(g < G ? hist[g] : 0u, s_red, tot);
            if (g < G) gst[g] = carry + ex;
            carry += tot;
            __syncthreads();
        }
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j)
            if (fk[j] != 0xFFFFFFFFu) s_rec[gst[fk[j]] + pos[j]] = v[j];
        __syncthreads();
        const uint32_t nch = (uint32_t)min((uint64_t)CH, hi - cb);
        for (uint32_t i = tid; i < nch; i += ROUTE_T) {
            const Rec<RW> r = s_rec[i];
            const uint32_t g = (rec_meta<RW>(r) >> shift) & gm;
            store_rec<RW>(a.dst, (uint64_t)cur[g] + (i - gst[g]), r);
        }
        __syncthreads();
        for (uint32_t g = tid; g < G; g += blockDim.x) cur[g] += hist[g];
        __syncthreads();
    }
}

uint32_t route_split_bits(uint32_t nsp)
{
    uint32_t log2f = 0;
    while ((nsp >> log2f) > 256u) ++log2f;
    uint32_t lnsp = 0;
    while ((1u << lnsp) < nsp) ++lnsp;
    const uint32_t xb = HSEL_BITS - lnsp;
    // at most 8 split groups per bucket (items of up to 8 splits read one
    // range; more splits filter) and 512 groups: a 128 KiB chunk then writes
    // runs of 16 records or more per group on average (coalesced stores)
    return min(min(xb, 3u), 9u - min(log2f, 9u));
}

// Coarse buckets to split-order: those holding a bucket whose estimated
// distinct k-mers exceed one table (k-mers x scale > 1, the rule of
// bucket_splits), so the fine pass lays out its split groups.
__global__ __launch_bounds__(256) void k_route_flags(RouteArgs a, uint32_t log2f)
{
    const uint32_t c = threadIdx.x;
    if (c >= (a.nsp >> log2f)) return;
    const float scale = *a.split_scale;
    uint32_t any = 0;
    for (uint32_t f = 0; f < (1u << log2f); ++f)
        any |= (float)a.split_km[(c << log2f) + f] * scale > 1.0f ? 1u : 0u;
    a.split_flags[c] = any;
}

void launch_route_scatter(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint32_t* tmp, uint32_t rw)
{
    if (ntiles == 0) return;
    // coarse buckets: up to 256, F = nsp / nc sub-partitions each
    uint32_t log2f = 0;
    while ((a.nsp >> log2f) > 256u) ++log2f;
    const uint32_t nc = a.nsp >> log2f;
    // groups of the fine scatter: F = nsp / nc sub-partitions of 2^split_bits
    // split groups each, G <= FINE_F_MAX (its LDS)
    set_kernel_attributes();
    if (a.split_flags) k_route_flags<<<1, 256, 0, st>>>(a, log2f);
    const size_t lc = 0, lf = 3ull * (1u << (log2f + a.split_bits)) * sizeof(uint32_t);
    // split-ordered buckets are moved in FINE_RANGES ranges when their
    // group counts per range have somewhere to go
    const uint32_t nr = a.split_flags && a.fine_hist && a.split_bits ? FINE_RANGES : 1u;
    if (rw == 2) k_route_coarse<2><<<ntiles, ROUTE_T, lc, st>>>(a, log2f, tmp);
    else if (rw == 4) k_route_coarse<4><<<ntiles, ROUTE_T, lc, st>>>(a, log2f, tmp);
    else k_route_coarse<8><<<ntiles, ROUTE_T, lc, st>>>(a, log2f, tmp);
    if (nr > 1) k_fine_hist<<<nc * nr, ROUTE_T, 0, st>>>(a, log2f);
    if (rw == 2) k_route_fine<2><<<nc * nr, ROUTE_T, lf, st>>>(a, log2f, tmp, nr);
    else if (rw == 4) k_route_fine<4><<<nc * nr, ROUTE_T, lf, st>>>(a, log2f, tmp, nr);
    else k_route_fine<8><<<nc * nr, ROUTE_T, lf, st>>>(a, log2f, tmp, nr);
}

// ============================================================== fill ====
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint64_t n)
{
    if (!n) return;
    uint64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    k_fill_u32<<<(unsigned)b, 256, 0, st>>>(p, v, n);
}
void launch_fill_u64(hipStream_t st, uint64_t* p, uint64_t v, uint64_t n)
{
    if (!n) return;
    uint64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    k_fill_u64<<<(unsigned)b, 256, 0, st>>>(p, v, n);
}

__global__ __launch_bounds__(256) void k_zero_regions(ZeroRegions z)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint32_t q = 0; q < z.count; ++q) {
        if (z.wide[q]) {
            uint64_t* p = static_cast<uint64_t*>(z.p[q]);
            for (uint64_t i = t0; i < z.n[q]; i += stride) p[i] = 0ull;
        } else {
            uint32_t* p = static_cast<uint32_t*>(z.p[q]);
            for (uint64_t i = t0; i < z.n[q]; i += stride) p[i] = 0u;
        }
    }
}

void launch_zero_regions(hipStream_t st, const ZeroRegions& z)
{
    uint64_t mx = 0;
    for (uint32_t q = 0; q < z.count; ++q) mx = std::max(mx, z.n[q]);
    if (!mx) return;
    uint64_t b = (mx + 255) / 256;
    if (b > 4096) b = 4096;
    k_zero_regions<<<(unsigned)b, 256, 0, st>>>(z);
}

// ========================================================== counting ====
constexpr uint32_t SPILL_PROBES = 4096;

// Spill-table view passed by value to the out-of-line insert paths (a
// reference to the kernel's argument struct would force a per-thread copy of
// it into scratch memory).
//
// Overflow: a key that finds no slot within SPILL_PROBES sets *err and is
// dropped; the host then counts the whole job again with a larger table, so
// once *err is set every later insert returns at once instead of scanning a
// full table.
struct SpillRef {
    uint64_t* keys;
    uint32_t* cnt;
    uint64_t mask;
    uint32_t* err;
};

__device__ __forceinline__ SpillRef spill_ref(const CountArgs& a)
{
    return SpillRef{a.sp_keys, a.sp_cnt, a.sp_mask, a.sp_err};
}

// Global spill table, one 64-bit word per key: claim by CAS on the key.
__device__ __noinline__ bool spill_insert1(SpillRef t, uint64_t x, uint32_t w = 1u)
{
    if (__hip_atomic_load(t.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    uint64_t slot = mix64(x ^ 0x9e3779b97f4a7c15ull) & t.mask;
    for (uint32_t p = 0; p <= t.mask && p < SPILL_PROBES; ++p) {
        unsigned long long* kp = (unsigned long long*)&t.keys[slot];
        const uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == x) {
            atomicAdd(&t.cnt[slot], w);
            return true;
        }
        if (cur == EMPTY64) {
            const uint64_t old = atomicCAS(kp, (unsigned long long)EMPTY64, (unsigned long long)x);
            if (old == EMPTY64 || old == x) {
                atomicAdd(&t.cnt[slot], w);
                return true;
            }
        }
        slot = (slot + 1) & t.mask;
    }
    atomicOr(t.err, 1u);
    return false;
}

// Global spill table, two words per key: the count word is the slot state
// (0 empty, LOCKED32 being written, otherwise published count).
__device__ __noinline__ bool spill_insert2(SpillRef t, uint64_t hi, uint64_t lo)
{
    if (__hip_atomic_load(t.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    uint64_t slot = mix64(lo ^ mix64(hi ^ 0x9e3779b97f4a7c15ull)) & t.mask;
    for (uint32_t p = 0; p <= t.mask && p < SPILL_PROBES;) {
        uint32_t* cp = &t.cnt[slot];
        const uint32_t c = __hip_atomic_load(cp, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (c == 0u) {
            if (atomicCAS(cp, 0u, LOCKED32) == 0u) {
                __hip_atomic_store(&t.keys[2 * slot], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&t.keys[2 * slot + 1], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(cp, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                return true;
            }
            continue;
        }
        if (c == LOCKED32) continue;
        const uint64_t kh = __hip_atomic_load(&t.keys[2 * slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t kl = __hip_atomic_load(&t.keys[2 * slot + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (kh == hi && kl == lo) {
            atomicAdd(cp, 1u);
            return true;
        }
        slot = (slot + 1) & t.mask;
        ++p;
    }
    atomicOr(t.err, 1u);
    return false;
}

// Global spill table of a batch: SW words per key (the k-mer words and the
// job tag), count word as slot state as in spill_insert2.
struct SpKey {
    uint64_t w[3];
};

template <int SW>
__device__ __noinline__ bool spill_insert_w(SpillRef t, SpKey key, uint32_t w = 1u)
{
    if (__hip_atomic_load(t.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
    for (int i = 0; i < SW; ++i) h = mix64(h ^ key.w[i]);
    uint64_t slot = h & t.mask;
    for (uint32_t p = 0; p <= t.mask && p < SPILL_PROBES;) {
        uint32_t* cp = &t.cnt[slot];
        const uint32_t c = __hip_atomic_load(cp, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (c == 0u) {
            if (atomicCAS(cp, 0u, LOCKED32) == 0u) {
#pragma unroll
                for (int i = 0; i < SW; ++i)
                    __hip_atomic_store(&t.keys[SW * slot + i], key.w[i], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(cp, w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                return true;
            }
            continue;
        }
        if (c == LOCKED32) continue;
        bool eq = true;
#pragma unroll
        for (int i = 0; i < SW; ++i)
            eq = eq && __hip_atomic_load(&t.keys[SW * slot + i], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) == key.w[i];
        if (eq) {
            atomicAdd(cp, w);
            return true;
        }
        slot = (slot + 1) & t.mask;
        ++p;
    }
    atomicOr(t.err, 1u);
    return false;
}

// Slot of a key in an LDS table of CAP slots: the top bits of a
// multiplicative hash of the folded key (slot placement never affects results).
template <int NW, uint32_t CAP>
__device__ __forceinline__ uint32_t slot_of(uint64_t xh, uint64_t xl)
{
    uint32_t h = (uint32_t)xl ^ (uint32_t)(xl >> 29);
    if (NW == 2) h ^= (uint32_t)xh ^ (uint32_t)(xh >> 32) * 0x85EBCA77u;
    return (h * 0x9E3779B1u) >> (32 - __builtin_ctz(CAP));
}

// Forward and reverse-complement 2k-bit k-mer whose first base is at bit
// offset 2*sh of word wi of src (words wi .. wi+2, or wi+4 when NW == 2).
// NW == 1: (fl, rl), k <= 32; NW == 2: (fh:fl, rh:rl), 32 < k <= 64.
template <int NW>
__device__ __forceinline__ void kmer_at(const uint32_t* src, uint32_t wi, uint32_t sh, uint32_t k,
                                        uint64_t& fh, uint64_t& fl, uint64_t& rh, uint64_t& rl)
{
    const uint64_t g = 16ull * wi + sh;
    if (NW == 1) {
        fl = load_kmer64(src, g, k);
        rl = (~rev_pairs64(fl)) >> (64u - 2u * k);
        fh = rh = 0;
    } else {
        load_kmer128(src, g, k, fh, fl);
        rh = ~rev_pairs64(fl);
        rl = ~rev_pairs64(fh);
        const uint32_t r = 128u - 2u * k;
        if (r) { rl = (rl >> r) | (rh << (64u - r)); rh >>= r; }
    }
}

// Roll base b (0..3) into the forward k-mer and its complement into the
// reverse complement.  kmask: the valid bits of the forward word that holds
// the first base (fl when NW == 1, fh when NW == 2); rsh: the bit offset of
// the new reverse-complement base in that word.
// KH (one-word keys): 1 when k >= 17, 0 when k <= 16, -1 to decide at run time.
template <int NW, int KH = -1>
__device__ __forceinline__ void roll_base(uint64_t& fh, uint64_t& fl, uint64_t& rh, uint64_t& rl,
                                          uint32_t b, uint64_t kmask, uint32_t rsh)
{
    if (NW == 1) {
        // 32-bit halves: funnel shifts instead of 64-bit shifts and masks
        uint32_t flo = (uint32_t)fl, fhi = (uint32_t)(fl >> 32);
        uint32_t rlo = (uint32_t)rl, rhi = (uint32_t)(rl >> 32);
        const uint32_t cb = b ^ 3u;
        if (KH == 1 || (KH < 0 && rsh >= 32u)) {  // k >= 17: the low forward word needs no mask
            fhi = __builtin_amdgcn_alignbit(fhi, flo, 30u) & (uint32_t)(kmask >> 32);
            flo = (flo << 2) | b;
            rlo = __builtin_amdgcn_alignbit(rhi, rlo, 2u);
            rhi = (rhi >> 2) | (cb << (rsh - 32u));
        } else {           // k <= 16: both fit in the low word
            flo = ((flo << 2) | b) & (uint32_t)kmask;
            rlo = (rlo >> 2) | (cb << rsh);
        }
        fl = ((uint64_t)fhi << 32) | flo;
        rl = ((uint64_t)rhi << 32) | rlo;
    } else {
        fh = ((fh << 2) | (fl >> 62)) & kmask;
        fl = (fl << 2) | b;
        rl = (rl >> 2) | (rh << 62);
        rh = (rh >> 2) | ((uint64_t)(b ^ 3u) << rsh);
    }
}

// Canonical k-mer: the smaller of forward and reverse complement.
template <int NW>
__device__ __forceinline__ void canon_pick(uint64_t fh, uint64_t fl, uint64_t rh, uint64_t rl,
                                           uint64_t& xh, uint64_t& xl)
{
    if (NW == 1) {
        xh = 0;
        xl = fl < rl ? fl : rl;
    } else {
        const bool f = fh < rh || (fh == rh && fl <= rl);
        xh = f ? fh : rh;
        xl = f ? fl : rl;
    }
}


// Block-wide exclusive scan of one u32 per thread (blockDim multiple of 64).
__device__ __forceinline__ uint32_t block_excl_u32(uint32_t v, uint32_t* sh, uint32_t& total)
{
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t i = 0; i < nw; ++i) {
        if (i < wv) base += sh[i];
        tot += sh[i];
    }
    total = tot;
    __syncthreads();
    return base + x - v;
}

template <int NW>
__device__ __forceinline__ void emit_out(const CountArgs& a, uint64_t pos, uint64_t hi, uint64_t lo, uint32_t c)
{
    if (pos >= a.out_cap) return;
    if (NW == 1) {
        a.out_keys[pos] = lo;
    } else {
        a.out_keys[2 * pos] = lo;
        a.out_keys[2 * pos + 1] = hi;
    }
    a.out_cnt[pos] = c;
}

__device__ __forceinline__ uint32_t wave_incl_u32(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(v, d, 64);
        if (lane >= d) v += y;
    }
    return v;
}

// Occurrences of the key in LDS slot s.  One-word table: the count word holds
// the occurrences after the one that claimed the slot (a claim needs no add).
// Two-word table: the count word is the count (the claim writes 1).
template <int NW>
__device__ __forceinline__ uint32_t slot_count(const uint64_t* k0, const uint32_t* cnt, uint32_t s)
{
    if (NW == 1) return k0[s] != EMPTY64 ? cnt[s] + 1u : 0u;
    return cnt[s];
}

// Keys waiting on a wave's stack for their next probe (k <= 32): the count
// loop makes exactly one probe per k-mer and pushes the misses, so its lanes
// stay in step; the stack is then worked off 64 entries per pass, one probe
// per entry, the misses pushed back one step further along their sequence.
constexpr uint32_t DEFER_CAP = 128;  // entries per wave (>= 64 + 64)
constexpr uint32_t DEFER_PSH = 13;   // entry state: slot | probes << 13 (CAP <= 8192)
constexpr uint32_t DEFER_WSH = 20;   //   | weight << 20 (occurrences the entry stands for)

// A fat record whose bases repeat with a period P (tandem repeats,
// homopolymers) has only P distinct k-mer positions: k-mer j equals k-mer
// j + P.  Its k-mers are counted as positions 0..P-1 with weights: position j
// stands for the floor((nk - 1 - j) / P) + 1 k-mers j, j + P, ... of the record,
// i.e. nk / P, plus one for the first nk % P positions.  Periods tried: 12, 10
// and 8, which cover every period dividing one of them (1-6, 8, 10, 12); the
// smallest of those three that holds is used.  bases: the record's base words
// (first base in bits 31:30 of b[0]), L its length (<= 48).
__device__ __forceinline__ uint32_t record_period(const uint32_t* b, uint32_t L)
{
    const uint64_t hi = ((uint64_t)b[0] << 32) | b[1];  // bases 0..31
    const uint64_t lo = (uint64_t)b[2] << 32;            // bases 32..47 on top
    uint32_t best = 0;
#pragma unroll
    for (uint32_t P = 12; P >= 8; P -= 2) {
        if (P >= L) continue;
        const uint32_t sh = 2u * P;
        const uint64_t xh = hi ^ ((hi << sh) | (lo >> (64u - sh)));
        const uint64_t xl = lo ^ (lo << sh);
        const uint32_t n = 2u * (L - P);  // bits that must agree (from the top)
        const bool ok = n <= 64u ? (xh >> (64u - n)) == 0ull
                                 : xh == 0ull && (n == 64u || (xl >> (128u - n)) == 0ull);
        if (ok) best = P;
    }
    return best;
}

template <uint32_t CAP, bool BATCH>
__device__ __forceinline__ uint32_t defer_passes(uint64_t* keys, uint32_t* cnt, uint64_t* qk,
                                                 uint32_t* qs, uint32_t qn, bool all,
                                                 const CountArgs& a, uint32_t tag, uint32_t& spilled,
                                                 uint32_t lane)
{
    static_assert(CAP <= (1u << DEFER_PSH), "slot field");
    // qn is wave-uniform: the loop is a scalar branch
    qn = __builtin_amdgcn_readfirstlane(qn);
    while (qn >= 64u || (all && qn)) {
        const uint32_t take = min(qn, 64u), base = qn - take;
        const bool has = lane < take;
        uint64_t x = 0;
        uint32_t e = 0;
        if (has) {
            x = qk[base + lane];
            e = qs[base + lane];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t slot = e & ((1u << DEFER_PSH) - 1u), p = (e >> DEFER_PSH) & ((1u << (DEFER_WSH - DEFER_PSH)) - 1u);
        const uint32_t w = e >> DEFER_WSH;
        bool again = false;
        if (has) {
            const uint64_t old = atomicCAS((unsigned long long*)&keys[slot], (unsigned long long)EMPTY64,
                                           (unsigned long long)x);
            if (old == EMPTY64 || old == x) {
                // a claim is one occurrence, counted by the drain
                const uint32_t add = old == x ? w : w - 1u;
                if (add) atomicAdd(&cnt[slot], add);
            } else if (++p == a.probe_max) {
                if (BATCH) spill_insert_w<2>(spill_ref(a), SpKey{{(uint64_t)tag, x, 0ull}}, w);
                else spill_insert1(spill_ref(a), x, w);
                spilled += w;
            } else {
                again = true;
                slot = (slot + p) & (CAP - 1u);  // triangular steps
            }
        }
        const uint64_t am = __ballot(again);
        if (again) {
            const uint32_t at = base + lane_rank(am);
            qk[at] = x;
            qs[at] = slot | (p << DEFER_PSH) | (w << DEFER_WSH);
        }
        qn = __builtin_amdgcn_readfirstlane(base + (uint32_t)__popcll(am));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    return qn;
}

// packed words staged per wave: 64 records x (RW-1) words, + 1 read ahead
template <int NW>
constexpr uint32_t count_stage() { return NW == 1 ? 64u * 3u + 8u : 64u * 7u + 8u; }

// One workgroup per work item.  A wave takes 64 fat records at a time from a
// workgroup dispenser (one coalesced vector load per lane), compacts the
// records of the item's sub-partition into its LDS stage at a fixed stride,
// and then their k-mers form a flat index space (wave scan) cut into 64 equal
// pieces, one per lane.  A lane finds its first record once (binary search)
// and rolls through consecutive k-mers from LDS, re-seeding from the next
// record's first k-mer (computed once per record) when it crosses into it.
//
// Tables: NW == 1, one 64-bit word per key claimed by compare-and-swap, plus a
// count word for the occurrences after the claiming one; the first probe of a
// k-mer only reads the slot, and claims go through the deferred pass.  NW == 2,
// two key words per slot and the count word as the slot
// state (0 empty, LOCKED32 being written, otherwise the count).  A key probes
// slot h, h+1, h+3, h+6, ... (triangular steps, which visit every slot of a
// power-of-two table and break up the clusters of linear probing).  Either
// way a slot goes from empty to one key exactly once, so a key absent from
// its full probe window can never enter the table later: spilling it keeps
// every occurrence of a key in exactly one table.
//
// Each loop iteration makes ONE probe for each lane's current k-mer; a lane
// whose probe ends (counted or spilled) rolls to its next k-mer while the
// others keep probing, so a wave pays the probes of its lanes' sums, not the
// worst lane at every k-mer.
//
// BATCH: a persistent launch; workgroups take items from a device queue until
// the device-side item count is reached, spill keys carry the item's job, and
// every item records its output segment for the per-job gather.
template <int NW, uint32_t CAP, bool BATCH, int KH>
__global__ __launch_bounds__(count_block(NW)) void k_count(CountArgs a)
{
    constexpr uint32_t NWAVE = count_block(NW) / 64;
    __shared__ __align__(16) uint64_t s_k0[CAP];
    __shared__ uint64_t s_k1[NW == 2 ? CAP : 1];
    __shared__ uint32_t s_cnt[CAP];
    __shared__ uint32_t s_stage[NWAVE][count_stage<NW>()];
    __shared__ uint32_t s_incl[NWAVE][64];
    __shared__ uint32_t s_wgt[NW == 1 ? NWAVE : 1][64];  // k <= 32: weights of a record's positions
    __shared__ uint64_t s_seed[NWAVE][64][2 * NW];  // first k-mer of each record: fwd, rc
    __shared__ uint64_t s_qk[NWAVE][NW == 1 ? DEFER_CAP : 1];  // deferred keys (k <= 32)
    __shared__ uint32_t s_qs[NWAVE][NW == 1 ? DEFER_CAP : 1];  // their probe state
    __shared__ uint32_t s_red[NWAVE];
    __shared__ uint32_t s_next, s_tot_u, s_tot_s;
    // the next work item, fetched by one thread while the current one is counted
    __shared__ uint4 s_nit;
    __shared__ uint32_t s_nitem, s_ntag, s_nstop;
    __shared__ unsigned long long s_base;

    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t k = a.k;
    constexpr uint32_t RW = NW == 1 ? 4u : 8u, PW = RW - 1u;  // record words, base words
    const uint32_t* stage = s_stage[wv];
    // roll constants: valid bits of the word holding the first base, and the
    // bit offset of a new reverse-complement base
    const uint32_t topb = NW == 1 ? 2u * k : 2u * k - 64u;
    const uint64_t kmask = topb == 64u ? ~0ull : ((1ull << topb) - 1u);
    const uint32_t rsh = topb - 2u;
    const uint32_t n_items = BATCH ? *a.n_items_dev : a.n_items;
    // Work item fetch: the item (from the device queue when BATCH), its
    // bounds, job and the spill-overflow flag (after an overflow the job is
    // counted again, so the rest is skipped), into the s_n* slots.  One thread
    // fetches the next item while the workgroup counts the current one, so
    // the queue and item reads are off the critical path.
    auto fetch = [&](uint32_t ni) {
        if (BATCH) ni = atomicAdd(a.item_q, 1u);
        uint4 nit = make_uint4(0u, 0u, 0u, 0u);
        uint32_t ntag = 0;
        if (ni < n_items) {
            nit = a.items[ni];
            if (BATCH) ntag = a.item_job[ni];
        }
        s_nstop = __hip_atomic_load(a.sp_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_nitem = ni;
        s_nit = nit;
        s_ntag = ntag;
    };
    if (threadIdx.x == 0) fetch(blockIdx.x);
    __syncthreads();
    uint32_t item = s_nitem, tag = s_ntag, stop = s_nstop;
    uint4 it = s_nit;
    // the table starts empty; the drain empties every slot it reads
    for (uint32_t i = threadIdx.x; i < CAP; i += blockDim.x) {
        if (NW == 1) s_k0[i] = EMPTY64;
        s_cnt[i] = 0;
    }
    // An item's first record batch is wave wv's records [x + 64 wv, +64) of
    // the item, loaded before the item starts (for later items during the
    // previous item's drain); the dispenser hands out the rest from x + 64 NWAVE.
    uint32_t rb = it.x + 64u * wv;
    Rec<RW> nxt;
    if (item < n_items && rb + lane < (stop ? it.x : it.y)) nxt = load_rec<RW>(a.rec, rb + lane);
    while (item < n_items) {
        if (threadIdx.x == 0) {
            s_next = it.x + 64u * NWAVE;
            s_tot_u = s_tot_s = 0;
        }
        __syncthreads();
        if (BATCH && threadIdx.x == 0) fetch(0u);

        const uint32_t r1 = stop ? it.x : it.y, smask = it.z, sval = it.w;
        uint32_t spilled = 0;
        uint32_t qn = 0;  // deferred keys on this wave's stack (k <= 32)
        // records are loaded one batch ahead: the next 64 are in flight while
        // the current ones are counted
        for (;;) {
            if (rb >= r1) break;
            const uint32_t i = rb + lane;
            const Rec<RW> rr = nxt;
            {
                uint32_t rn = 0;
                if (lane == 0) rn = atomicAdd(&s_next, 64u);
                rb = __shfl(rn, 0, 64);
                if (rb + lane < r1) nxt = load_rec<RW>(a.rec, rb + lane);
            }
            // the item's records, compacted to the front of the wave's stage
            const uint32_t len = rr.w[0] & REC_FAT_LEN_MASK;
            const bool live = i < r1 && len >= k && ((rr.w[0] >> REC_HSEL_SHIFT) & smask) == sval;
            uint32_t nk = live ? len - k + 1u : 0u;
            // k <= 32: a periodic record (longer than a superkmer of one
            // minimizer occurrence can be, a.per_len) visits only its first P
            // positions, weighted: wb = nk / P each, one more for the first
            // nk % P (packed wb | rem << 16)
            uint32_t wgt = 1u;
            if (NW == 1 && live && len >= a.per_len) {
                const uint32_t P = record_period(rr.w + 1, len);
                if (P && P < nk) {
                    wgt = (nk / P) | ((nk % P) << 16);
                    nk = P;
                }
            }
            const uint64_t lm = __ballot(live);
            const uint32_t ci = lane_rank(lm), nlive = (uint32_t)__popcll(lm);
            const uint32_t incl = wave_incl_u32(nk, lane);
            const uint32_t total = __shfl(incl, 63, 64);
            if (live) {
#pragma unroll
                for (uint32_t q = 0; q < PW; ++q) s_stage[wv][ci * PW + q] = rr.w[1 + q];
                s_incl[wv][ci] = incl;
                if (NW == 1) s_wgt[NW == 1 ? wv : 0][ci] = wgt;
                uint64_t fh, fl, rh, rl;  // the record's first k-mer, from registers
                kmer_at<NW>(rr.w + 1, 0u, 0u, k, fh, fl, rh, rl);
                s_seed[wv][ci][0] = fl;
                s_seed[wv][ci][1] = rl;
                if (NW == 2) {
                    s_seed[wv][ci][NW == 2 ? 2 : 0] = fh;
                    s_seed[wv][ci][NW == 2 ? 3 : 1] = rh;
                }
            }
            if (lane >= nlive) s_incl[wv][lane] = total;  // keeps the search in range
            const bool weighted = NW == 1 && __ballot(wgt != 1u) != 0ull;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t q = (total + 63u) >> 6;
            const uint32_t j0 = lane * q, j1 = min(j0 + q, total);
            if constexpr (NW == 1) {
                // k <= 32: every lane advances every iteration; a probe that
                // misses moves the key to the wave's stack
                const uint32_t nmine = j1 > j0 ? j1 - j0 : 0u;
                uint32_t s = 0, end = 0, pc = 1u;
                uint32_t rpos = 0, wcur = 1u;  // position in the record, its weights (wb | rem << 16)
                uint64_t fh = 0, fl = 0, rh = 0, rl = 0;
                if (nmine) {
#pragma unroll
                    for (uint32_t step = 32; step; step >>= 1)
                        if (s_incl[wv][s + step - 1u] <= j0) s += step;
                    const uint32_t ex = s ? s_incl[wv][s - 1u] : 0u;
                    end = s_incl[wv][s];
                    wcur = s_wgt[NW == 1 ? wv : 0][s];
                    const uint32_t gb = j0 - ex;
                    rpos = gb;
                    kmer_at<1>(stage, s * PW + (gb >> 4), gb & 15u, k, fh, fl, rh, rl);
                    pc = 2u * (s * (16u * PW) + gb + k) + 1u;
                }
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): settle the set-up reads
                // A probe is a plain LDS read of the key's slot, used one
                // iteration after it is issued, so two probes per wave are in
                // flight.  The slot holds the key: count it; empty: push the key
                // with that slot, to be claimed by compare-and-swap in the
                // wave's deferred pass; another key: push the key with its next
                // slot.  A read is cheaper than a compare-and-swap, and on
                // repeated k-mers most probes find their key.
                // WT: the batch holds a periodic record, so a probe stands for a
                // weight pw of occurrences (the loop without weights is kept for
                // the common case: pw == 1, no weight bookkeeping)
                auto settle = [&](auto WT, bool pa, uint64_t px, uint32_t ps, uint64_t po, uint32_t pw) {
                    const bool hit = pa && po == px;
                    const bool miss = pa && !hit;
                    if (hit) atomicAdd(&s_cnt[ps], decltype(WT)::value ? pw : 1u);
                    const uint64_t mm = __ballot(miss);
                    if (mm) {
                        if (miss) {
                            // an empty slot is claimed by the deferred pass at the same slot
                            const bool emp = po == EMPTY64;
                            const uint32_t at = qn + lane_rank(mm);
                            s_qk[wv][at] = px;
                            s_qs[wv][at] = ((ps + (emp ? 0u : 1u)) & (CAP - 1u)) | ((emp ? 0u : 1u) << DEFER_PSH) |
                                           ((decltype(WT)::value ? pw : 1u) << DEFER_WSH);
                        }
                        qn = __builtin_amdgcn_readfirstlane(qn + (uint32_t)__popcll(mm));
                    }
                    if (__builtin_amdgcn_readfirstlane(qn) > DEFER_CAP - 64u)
                        qn = __builtin_amdgcn_readfirstlane(defer_passes<CAP, BATCH>(s_k0, s_cnt, s_qk[wv], s_qs[wv], qn, false, a, tag,
                                                      spilled, lane));
                };
                // k-mer i: issue its probe, then roll the lane on to k-mer i+1
                auto step = [&](auto WT, uint32_t i, uint64_t& x, uint32_t& slot, uint64_t& old, bool& act,
                                uint32_t& w) {
                    act = i < nmine;
                    const uint32_t wn = stage[pc >> 5];
                    const bool xing = act && j0 + i + 1u == end && i + 1u < nmine;
                    uint32_t nend, nw;
                    uint64_t nfl, nrl;  // read only when xing
                    if (xing) {
                        const uint32_t sn = min(s + 1u, 63u);
                        nend = s_incl[wv][sn];
                        nfl = s_seed[wv][sn][0];
                        nrl = s_seed[wv][sn][1];
                        if constexpr (decltype(WT)::value) nw = s_wgt[NW == 1 ? wv : 0][sn];
                    }
                    x = fl < rl ? fl : rl;
                    slot = slot_of<1, CAP>(0ull, x);
                    if constexpr (decltype(WT)::value) w = (wcur & 0xFFFFu) + (rpos < (wcur >> 16) ? 1u : 0u);
                    else w = 1u;
                    // every lane reads a slot; a lane without a k-mer ignores it
                    old = __hip_atomic_load(&s_k0[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const uint32_t b = __builtin_amdgcn_ubfe(wn, ~pc, 2u);
                    if (act && i + 1u < nmine) {
                        if (xing) {
                            ++s;
                            end = nend;
                            fl = nfl;
                            rl = nrl;
                            if constexpr (decltype(WT)::value) {
                                wcur = nw;
                                rpos = 0;
                            }
                            pc = 2u * (s * (16u * PW) + k) + 1u;
                        } else {
                            pc += 2u;
                            if constexpr (decltype(WT)::value) ++rpos;
                            roll_base<1, KH>(fh, fl, rh, rl, b, kmask, rsh);
                        }
                    }
                };
                // two k-mers per trip with their own result registers, so a probe
                // result crosses the loop edge without a copy (which would wait)
                auto run = [&](auto WT) {
                    uint64_t xB = 0, oB = 0;
                    uint32_t sB = 0, wB = 1;
                    bool aB = false;
                    for (uint32_t i = 0; i < q; i += 2u) {
                        uint64_t xA, oA;
                        uint32_t sA, wA;
                        bool aA;
                        step(WT, i, xA, sA, oA, aA, wA);
                        settle(WT, aB, xB, sB, oB, wB);
                        aB = false;
                        if (i + 1u < q) step(WT, i + 1u, xB, sB, oB, aB, wB);
                        settle(WT, aA, xA, sA, oA, wA);
                    }
                    settle(WT, aB, xB, sB, oB, wB);
                };
                if (weighted) run(std::true_type{});
                else run(std::false_type{});
            } else if (j0 < j1) {
                uint32_t s = 0;
#pragma unroll
                for (uint32_t step = 32; step; step >>= 1)
                    if (s_incl[wv][s + step - 1u] <= j0) s += step;
                const uint32_t ex = s ? s_incl[wv][s - 1u] : 0u;
                uint32_t end = s_incl[wv][s];
                const uint32_t gb = j0 - ex;  // k-mer offset inside record s
                uint64_t fh, fl, rh, rl;
                kmer_at<NW>(stage, s * PW + (gb >> 4), gb & 15u, k, fh, fl, rh, rl);
                // stage bit index, plus one, of the base rolled in next (odd, so
                // that its complement's low five bits give the base's shift)
                uint32_t pc = 2u * (s * (16u * PW) + gb + k) + 1u;
                uint64_t xh, xl;
                canon_pick<NW>(fh, fl, rh, rl, xh, xl);
                uint32_t slot = slot_of<NW, CAP>(xh, xl), p = 0;
                // settle the set-up reads here: a read still in flight at the loop
                // entry would otherwise be waited for inside every iteration
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
                for (uint32_t j = j0;;) {
                    // the LDS reads an advance may need are issued with the probe,
                    // so that one wait covers them all
                    const uint32_t wn = stage[pc >> 5];
                    const bool xing = j + 1u == end;  // advancing would cross
                    uint32_t nend;
                    uint64_t nfl, nrl, nfh, nrh;  // read only when xing
                    if (xing) {
                        const uint32_t sn = min(s + 1u, 63u);
                        nend = s_incl[wv][sn];
                        nfl = s_seed[wv][sn][0];
                        nrl = s_seed[wv][sn][1];
                        nfh = s_seed[wv][sn][2];
                        nrh = s_seed[wv][sn][3];
                    }
                    bool adv = false, miss = false;
                    // the slot state, then its key words, read without a wait in
                    // between: LDS serves a wave's accesses in order and a claim
                    // writes the key before it publishes the count, so a published
                    // count read first means the key read after it is complete
                    const uint32_t c = __hip_atomic_load(&s_cnt[slot], __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
                    __asm__ volatile("" ::: "memory");
                    const uint64_t sk0 = s_k0[slot], sk1 = s_k1[slot];
                    if (c == 0u) {
                        // claim; a lost race retries the same slot next time
                        if (atomicCAS(&s_cnt[slot], 0u, LOCKED32) == 0u) {
                            s_k0[slot] = xh;
                            s_k1[slot] = xl;
                            __hip_atomic_store(&s_cnt[slot], 1u, __ATOMIC_RELEASE,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                            adv = true;
                        }
                    } else if (c != LOCKED32) {
                        if (sk0 == xh && sk1 == xl) {
                            atomicAdd(&s_cnt[slot], 1u);
                            adv = true;
                        } else {
                            miss = true;
                        }
                    }
                    if (miss) {
                        if (++p == a.probe_max) {
                            if (BATCH) spill_insert_w<3>(spill_ref(a), SpKey{{xh, xl, (uint64_t)tag}});
                            else spill_insert2(spill_ref(a), xh, xl);
                            ++spilled;
                            adv = true;
                        } else {
                            slot = (slot + p) & (CAP - 1u);  // triangular steps
                        }
                    }
                    if (adv) {
                        if (++j == j1) break;
                        if (xing) {
                            ++s;
                            end = nend;
                            fl = nfl;
                            rl = nrl;
                            fh = nfh;
                            rh = nrh;
                            pc = 2u * (s * (16u * PW) + k) + 1u;
                        } else {
                            const uint32_t b = __builtin_amdgcn_ubfe(wn, ~pc, 2u);
                            pc += 2u;
                            roll_base<NW>(fh, fl, rh, rl, b, kmask, rsh);
                        }
                        canon_pick<NW>(fh, fl, rh, rl, xh, xl);
                        slot = slot_of<NW, CAP>(xh, xl);
                        p = 0;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if constexpr (NW == 1) {
            if (qn) defer_passes<CAP, BATCH>(s_k0, s_cnt, s_qk[wv], s_qs[wv], qn, true, a, tag, spilled, lane);
        }
        __syncthreads();

        // Drain.  Wave w owns slots [w, w+1) * CAP / NWAVE, taken 64 at a time:
        // the count pass keeps each slot's count and the ballot of the slots
        // to write, the wave totals give each wave its output base (one
        // reservation per workgroup), and the write pass compacts the written
        // slots with the ballots, so consecutive lanes write consecutive
        // entries (coalesced stores), and empties the wave's slots for the
        // next item (the next item's first barrier orders that).
        constexpr uint32_t DR = CAP / NWAVE / 64;  // drain rounds per wave
        static_assert(CAP % (NWAVE * 64) == 0, "drain rounds");
        uint32_t cv[DR];
        uint64_t em[DR];
        uint32_t ne = 0, nu = 0;
#pragma unroll
        for (uint32_t r = 0; r < DR; ++r) {
            const uint32_t s = (wv * DR + r) * 64u + lane;
            const uint32_t c = slot_count<NW>(s_k0, s_cnt, s);
            cv[r] = c;
            em[r] = __ballot(c && c >= a.ci && (uint64_t)c <= a.cx);
            ne += (uint32_t)__popcll(em[r]);
            nu += (uint32_t)__popcll(__ballot(c != 0u));
        }
        uint32_t ss = spilled;
#pragma unroll
        for (uint32_t d = 32; d; d >>= 1) ss += __shfl_xor(ss, d, 64);
        if (lane == 0) {
            s_red[wv] = ne;
            if (nu) atomicAdd(&s_tot_u, nu);
            if (ss) atomicAdd(&s_tot_s, ss);
        }
        __syncthreads();
        uint32_t wbase = 0, tot_out = 0;
#pragma unroll
        for (uint32_t w = 0; w < NWAVE; ++w) {
            const uint32_t n = s_red[w];
            wbase += w < wv ? n : 0u;
            tot_out += n;
        }
        const uint32_t tot_uniq = s_tot_u, tot_sp = s_tot_s;
        if (BATCH) {
            // the next item (fetched during this one, visible since the barrier
            // above): this wave's first batch, in flight during the writes
            const uint4 nit = s_nit;
            rb = nit.x + 64u * wv;
            if (s_nitem < n_items && rb + lane < (s_nstop ? nit.x : nit.y)) nxt = load_rec<RW>(a.rec, rb + lane);
        }
        if (threadIdx.x == 0) {
            s_base = tot_out ? atomicAdd(&a.counters[0], (unsigned long long)tot_out) : 0ull;
            if (tot_uniq) atomicAdd(&a.counters[1], (unsigned long long)tot_uniq);
            if (tot_sp) atomicAdd(&a.counters[2], (unsigned long long)tot_sp);
            if (a.item_uniq) a.item_uniq[item] = tot_uniq + tot_sp;
            if (a.item_sp) a.item_sp[item] = tot_sp;
            if (BATCH) {
                a.seg_pos[item] = s_base;
                a.seg_n[item] = tot_out;
                if (tot_uniq) atomicAdd(&a.job_uniq[tag], (unsigned long long)tot_uniq);
            }
        }
        __syncthreads();
        uint64_t pos = s_base + wbase;
#pragma unroll
        for (uint32_t r = 0; r < DR; ++r) {
            const uint32_t s = (wv * DR + r) * 64u + lane;
            if ((em[r] >> lane) & 1ull)
                emit_out<NW>(a, pos + lane_rank(em[r]), NW == 2 ? s_k0[s] : 0ull, NW == 2 ? s_k1[s] : s_k0[s],
                             cv[r]);
            pos += (uint32_t)__popcll(em[r]);
            if (NW == 1) s_k0[s] = EMPTY64;
            s_cnt[s] = 0;
        }
        if (!BATCH) break;
        item = s_nitem;
        tag = s_ntag;
        stop = s_nstop;
        it = s_nit;
    }
}

void launch_count(hipStream_t st, const CountArgs& a, uint32_t nw)
{
    if (a.n_items == 0) return;
    if (nw == 1 && a.k >= 17) k_count<1, LDS_CAP_NW1, false, 1><<<a.n_items, count_block(1), 0, st>>>(a);
    else if (nw == 1) k_count<1, LDS_CAP_NW1, false, 0><<<a.n_items, count_block(1), 0, st>>>(a);
    else k_count<2, LDS_CAP_NW2, false, -1><<<a.n_items, count_block(2), 0, st>>>(a);
}

// Resident workgroups of the persistent count launch: occupancy x CUs.
uint32_t count_persistent_blocks(uint32_t nw)
{
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (nw == 1)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, (const void*)k_count<1, LDS_CAP_NW1, true, 1>, count_block(1), 0);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, (const void*)k_count<2, LDS_CAP_NW2, true, -1>, count_block(2), 0);
    if (cus <= 0) cus = 256;
    if (per <= 0) per = 1;
    return (uint32_t)(cus * per);
}

void launch_count_batch(hipStream_t st, const CountArgs& a, uint32_t nw, uint32_t blocks)
{
    if (nw == 1 && a.k >= 17) k_count<1, LDS_CAP_NW1, true, 1><<<blocks, count_block(1), 0, st>>>(a);
    else if (nw == 1) k_count<1, LDS_CAP_NW1, true, 0><<<blocks, count_block(1), 0, st>>>(a);
    else k_count<2, LDS_CAP_NW2, true, -1><<<blocks, count_block(2), 0, st>>>(a);
}

template <int NW>
__global__ __launch_bounds__(256) void k_spill_drain(CountArgs a, uint64_t n_slots)
{
    __shared__ uint32_t s_red[4];
    __shared__ unsigned long long s_base;
    const uint64_t per_block = 256u * 16u;
    const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
    uint32_t mine = 0, uniq = 0;
    for (uint32_t j = 0; j < 16u; ++j) {
        const uint64_t s = b0 + j * 256u + threadIdx.x;
        if (s >= n_slots) break;
        const uint32_t c = a.sp_cnt[s];
        if (c) {
            ++uniq;
            if (c >= a.ci && (uint64_t)c <= a.cx) ++mine;
        }
    }
    uint32_t tot_out, tot_uniq;
    const uint32_t my_off = block_excl_u32(mine, s_red, tot_out);
    block_excl_u32(uniq, s_red, tot_uniq);
    if (threadIdx.x == 0) {
        s_base = tot_out ? atomicAdd(&a.counters[0], (unsigned long long)tot_out) : 0ull;
        if (tot_uniq) atomicAdd(&a.counters[1], (unsigned long long)tot_uniq);
    }
    __syncthreads();
    uint64_t pos = s_base + my_off;
    for (uint32_t j = 0; j < 16u; ++j) {
        const uint64_t s = b0 + j * 256u + threadIdx.x;
        if (s >= n_slots) break;
        const uint32_t c = a.sp_cnt[s];
        if (c && c >= a.ci && (uint64_t)c <= a.cx) {
            if (NW == 1) emit_out<1>(a, pos, 0, a.sp_keys[s], c);
            else emit_out<2>(a, pos, a.sp_keys[2 * s], a.sp_keys[2 * s + 1], c);
            ++pos;
        }
    }
}

void launch_spill_drain(hipStream_t st, const CountArgs& a, uint32_t nw, uint64_t n_slots)
{
    const uint64_t blocks = (n_slots + 4095) / 4096;
    if (!blocks) return;
    if (nw == 1) k_spill_drain<1><<<(unsigned)blocks, 256, 0, st>>>(a, n_slots);
    else k_spill_drain<2><<<(unsigned)blocks, 256, 0, st>>>(a, n_slots);
}

// ============================================================ batches ====
// Largest j in [0, n) with v[j] <= x (v non-decreasing, v[0] <= x).
__device__ __forceinline__ uint32_t last_le(const uint32_t* v, uint32_t n, uint64_t x, uint32_t mask)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((v[mid] & mask) <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Batch-relative chunk offsets (padded to the capacity with empty chunks), the
// job of every chunk and the job of every bucket.
__global__ void k_batch_setup(const BatchDesc* __restrict__ d, BatchJobs j, uint32_t cap_jobs,
                              uint64_t* __restrict__ boff, uint32_t* __restrict__ chunk_job,
                              uint64_t cap_ch, uint32_t* __restrict__ bucket_job, uint32_t nsp,
                              uint32_t* __restrict__ err)
{
    const uint64_t nch = min(d->n_chunks, cap_ch), N = d->n_bases, ob = d->off_base;
    const uint32_t nj = min(d->n_jobs, cap_jobs);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint64_t i = t0; i <= cap_ch; i += stride) {
        uint64_t v = N;
        if (i <= nch) {
            // offsets outside the batch's bases are clamped and reported
            const uint64_t o = d->off[i];
            const bool bad = o < ob || o - ob > N || (i == nch && o - ob != N);
            if (bad) atomicOr(err, 4u);
            v = o < ob ? 0 : min(o - ob, N);
        }
        boff[i] = v;
        if (i < cap_ch) chunk_job[i] = (i < nch && nj) ? last_le(j.job_chunk, nj, i, ~0u) : 0u;
    }
    for (uint64_t b = t0; b < nsp; b += stride) {
        uint32_t job = 0xFFFFFFFFu;
        if (nj) {
            const uint32_t q = last_le(j.job_route, nj, b, 0xFFFFu);
            const uint32_t r = j.job_route[q];
            if (b < (r & 0xFFFFu) + (1u << (r >> 16))) job = q;
        }
        bucket_job[b] = job;
    }
}

void launch_batch_setup(hipStream_t st, const BatchDesc* d, BatchJobs j, uint32_t cap_jobs,
                        uint64_t* boff, uint32_t* chunk_job, uint64_t cap_ch, uint32_t* bucket_job,
                        uint32_t nsp, uint32_t* err)
{
    uint64_t blocks = (std::max<uint64_t>(cap_ch + 1, nsp) + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_batch_setup<<<(unsigned)blocks, 256, 0, st>>>(d, j, cap_jobs, boff, chunk_job, cap_ch,
                                                     bucket_job, nsp, err);
}

// Work items of a batch, built on the device: every non-empty bucket is one
// item, split by the low routing bits into r = 2^i items when its estimated
// distinct k-mers (k-mers x scale, scale = rho / target) exceed one table.
// Buckets never merge, so an item never spans two jobs.  Three steps: split
// counts per bucket, their exclusive scan (bucket_item), the items.
__device__ __forceinline__ uint32_t bucket_splits(const uint64_t* part_km, const uint64_t* part_off,
                                                  uint32_t b, uint32_t rmax, float scale)
{
    if (part_off[b + 1] == part_off[b]) return 0u;
    const float est = (float)part_km[b] * scale;
    uint32_t r = 1;
    while (r < rmax && est > (float)r) r <<= 1;
    return r;
}

__global__ __launch_bounds__(256) void k_item_counts(const uint64_t* __restrict__ part_km,
                                                     const uint64_t* __restrict__ part_off,
                                                     uint32_t nsp, uint32_t xb,
                                                     const float* __restrict__ scale_p,
                                                     uint64_t* __restrict__ nitem)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nsp) nitem[b] = bucket_splits(part_km, part_off, b, 1u << xb, *scale_p);
}

// A bucket split into r <= 2^split_bits items: item q counts the records of
// split groups [q, q+1) * 2^split_bits / r, one contiguous range (the fine
// scatter orders a bucket's records by split group), with no filter.  More
// splits than that: every item reads the whole bucket and keeps the records
// whose low split bits equal q.
__global__ __launch_bounds__(256) void k_items(const uint64_t* __restrict__ part_km,
                                               const uint64_t* __restrict__ part_off, uint32_t nsp,
                                               uint32_t xb, const uint32_t* __restrict__ bucket_job,
                                               const float* __restrict__ scale_p,
                                               const uint64_t* __restrict__ bucket_item,
                                               uint4* __restrict__ items, uint32_t* __restrict__ item_job,
                                               uint32_t* __restrict__ n_items,
                                               const uint64_t* __restrict__ split_off, uint32_t split_bits)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) *n_items = (uint32_t)bucket_item[nsp];
    if (b >= nsp) return;
    const uint32_t r = bucket_splits(part_km, part_off, b, 1u << xb, *scale_p);
    const uint32_t lo = (uint32_t)part_off[b], hi = (uint32_t)part_off[b + 1];
    const uint32_t job = bucket_job[b];
    const uint64_t cur = bucket_item[b];
    // (the fine scatter orders only coarse buckets holding a split bucket, so
    // an unsplit bucket keeps its whole range)
    const bool ranges = split_off && r > 1u && r <= (1u << split_bits);
    const uint32_t per = ranges ? (1u << split_bits) / r : 0u;
    for (uint32_t q = 0; q < r; ++q) {
        if (ranges) {
            const uint64_t g = ((uint64_t)b << split_bits) + (uint64_t)q * per;
            items[cur + q] = make_uint4((uint32_t)split_off[g], (uint32_t)split_off[g + per], 0u, 0u);
        } else {
            items[cur + q] = make_uint4(lo, hi, r - 1u, q);
        }
        item_job[cur + q] = job;
    }
}

void launch_build_items(hipStream_t st, const uint64_t* part_km, const uint64_t* part_off,
                        uint32_t nsp, uint32_t xb, const uint32_t* bucket_job, const float* scale,
                        uint4* items, uint32_t* item_job, uint64_t* bucket_item, uint32_t* n_items,
                        uint64_t* nitem, uint64_t* scan_tmp, const uint64_t* split_off,
                        uint32_t split_bits)
{
    const unsigned blocks = (nsp + 255) / 256;
    k_item_counts<<<blocks, 256, 0, st>>>(part_km, part_off, nsp, xb, scale, nitem);
    launch_scan_u64(st, nitem, nsp, bucket_item, scan_tmp);
    k_items<<<blocks, 256, 0, st>>>(part_km, part_off, nsp, xb, bucket_job, scale, bucket_item, items,
                                    item_job, n_items, split_off, split_bits);
}

// Spill entries per job: distinct keys, and reported keys (ci/cx filter).
template <int SW>
__global__ __launch_bounds__(256) void k_spill_jobs(CountArgs a, uint64_t n_slots,
                                                    unsigned long long* __restrict__ job_sp)
{
    if (a.counters[2] == 0) return;  // no k-mer reached the spill table: it is empty
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n_slots; s += stride) {
        const uint32_t c = a.sp_cnt[s];
        if (!c) continue;
        const uint32_t tag = (uint32_t)a.sp_keys[SW * s + (SW == 2 ? 0 : 2)];
        atomicAdd(&a.job_uniq[tag], 1ull);
        if (c >= a.ci && (uint64_t)c <= a.cx) atomicAdd(&job_sp[tag], 1ull);
    }
}

void launch_spill_jobs(hipStream_t st, const CountArgs& a, uint32_t nw, uint64_t n_slots,
                       unsigned long long* job_sp)
{
    uint64_t blocks = (n_slots + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (!blocks) return;
    if (nw == 1) k_spill_jobs<2><<<(unsigned)blocks, 256, 0, st>>>(a, n_slots, job_sp);
    else k_spill_jobs<3><<<(unsigned)blocks, 256, 0, st>>>(a, n_slots, job_sp);
}

// Per job: reported keys from its items (segments) and from the spill table,
// its superkmers and its k-mers.  Jobs past the batch get 0.
__global__ void k_job_totals(const BatchDesc* __restrict__ d, BatchJobs j, uint32_t cap_jobs,
                             const uint64_t* __restrict__ bucket_item,
                             const uint64_t* __restrict__ seg_x,
                             const unsigned long long* __restrict__ job_sk,
                             const unsigned long long* __restrict__ job_sp,
                             const uint64_t* __restrict__ nkm_x,
                             uint64_t* __restrict__ job_tot, uint64_t* __restrict__ job_item_tot,
                             uint64_t* __restrict__ job_nsk, uint64_t* __restrict__ job_nkm)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cap_jobs) return;
    if (q >= d->n_jobs) {
        job_tot[q] = 0;
        job_item_tot[q] = 0;
        job_nsk[q] = 0;
        job_nkm[q] = 0;
        return;
    }
    const uint32_t r = j.job_route[q];
    const uint32_t b = r & 0xFFFFu, be = b + (1u << (r >> 16));
    const uint64_t it = seg_x[bucket_item[be]] - seg_x[bucket_item[b]];
    job_item_tot[q] = it;
    job_tot[q] = it + job_sp[q];
    job_nsk[q] = job_sk[q];
    job_nkm[q] = nkm_x[j.job_chunk[q + 1]] - nkm_x[j.job_chunk[q]];
}

void launch_job_totals(hipStream_t st, const BatchDesc* d, BatchJobs j, uint32_t cap_jobs,
                       const uint64_t* bucket_item, const uint64_t* seg_x,
                       const unsigned long long* job_sk, const unsigned long long* job_sp,
                       const uint64_t* nkm_x, uint64_t* job_tot, uint64_t* job_item_tot,
                       uint64_t* job_nsk, uint64_t* job_nkm)
{
    k_job_totals<<<(cap_jobs + 255) / 256, 256, 0, st>>>(d, j, cap_jobs, bucket_item, seg_x, job_sk,
                                                         job_sp, nkm_x, job_tot, job_item_tot,
                                                         job_nsk, job_nkm);
}

// Item segments into job order: job j's items first (item order), its spill
// entries after them.
template <int NW>
__global__ __launch_bounds__(256) void k_gather_items(CountArgs a, BatchJobs j, const BatchDesc* __restrict__ d,
                                                      const uint64_t* __restrict__ bucket_item,
                                                      const uint64_t* __restrict__ seg_x,
                                                      const uint64_t* __restrict__ job_off,
                                                      uint64_t* __restrict__ dst_keys,
                                                      uint32_t* __restrict__ dst_cnt)
{
    // one job: its items' reservations already tile out_keys from 0, which
    // is then the result (the spill entries follow them)
    if (d->n_jobs == 1) return;
    const uint32_t n = *a.n_items_dev;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t job = a.item_job[i];
        const uint32_t b = j.job_route[job] & 0xFFFFu;
        const uint64_t dst = job_off[job] + seg_x[i] - seg_x[bucket_item[b]];
        const uint64_t src = a.seg_pos[i], cnt = a.seg_n[i];
        for (uint64_t e = threadIdx.x; e < cnt; e += blockDim.x) {
            if (NW == 1) {
                dst_keys[dst + e] = a.out_keys[src + e];
            } else {
                dst_keys[2 * (dst + e)] = a.out_keys[2 * (src + e)];
                dst_keys[2 * (dst + e) + 1] = a.out_keys[2 * (src + e) + 1];
            }
            dst_cnt[dst + e] = a.out_cnt[src + e];
        }
    }
}

void launch_gather_items(hipStream_t st, const CountArgs& a, uint32_t nw, BatchJobs j, const BatchDesc* d,
                         const uint64_t* bucket_item, const uint64_t* seg_x,
                         const uint64_t* job_off, uint64_t* dst_keys, uint32_t* dst_cnt,
                         uint32_t blocks)
{
    if (nw == 1)
        k_gather_items<1><<<blocks, 256, 0, st>>>(a, j, d, bucket_item, seg_x, job_off, dst_keys, dst_cnt);
    else
        k_gather_items<2><<<blocks, 256, 0, st>>>(a, j, d, bucket_item, seg_x, job_off, dst_keys, dst_cnt);
}

__global__ __launch_bounds__(256) void k_pack_back(const BatchDesc* __restrict__ d, uint32_t cap_jobs,
                                                   const uint64_t* __restrict__ job_off,
                                                   const uint64_t* __restrict__ job_nsk,
                                                   const uint64_t* __restrict__ job_nkm,
                                                   const unsigned long long* __restrict__ job_uniq,
                                                   const uint32_t* __restrict__ flags, uint32_t n_flags,
                                                   const unsigned long long* __restrict__ counters,
                                                   uint8_t* __restrict__ back)
{
    const uint32_t nj = min(d->n_jobs, cap_jobs);
    uint64_t* o = reinterpret_cast<uint64_t*>(back);
    uint64_t* nsk = o + cap_jobs + 1;
    uint64_t* nkm = nsk + cap_jobs;
    uint64_t* unq = nkm + cap_jobs;
    uint32_t* fl = reinterpret_cast<uint32_t*>(unq + cap_jobs);
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(fl + n_flags);
    for (uint32_t i = threadIdx.x; i <= nj; i += blockDim.x) {
        o[i] = job_off[i];
        if (i < nj) {
            nsk[i] = job_nsk[i];
            nkm[i] = job_nkm[i];
            unq[i] = job_uniq[i];
        }
    }
    if (threadIdx.x < n_flags) fl[threadIdx.x] = flags[threadIdx.x];
    if (threadIdx.x < 8) ctr[threadIdx.x] = counters[threadIdx.x];
}

void launch_pack_back(hipStream_t st, const BatchDesc* d, uint32_t cap_jobs, const uint64_t* job_off,
                      const uint64_t* job_nsk, const uint64_t* job_nkm,
                      const unsigned long long* job_uniq, const uint32_t* flags, uint32_t n_flags,
                      const unsigned long long* counters, uint8_t* back)
{
    k_pack_back<<<1, 256, 0, st>>>(d, cap_jobs, job_off, job_nsk, job_nkm, job_uniq, flags, n_flags,
                                   counters, back);
}

template <int SW>
__global__ __launch_bounds__(256) void k_spill_place(CountArgs a, const BatchDesc* __restrict__ d,
                                                     uint64_t n_slots,
                                                     const uint64_t* __restrict__ job_off,
                                                     const uint64_t* __restrict__ job_item_tot,
                                                     unsigned long long* __restrict__ job_cur,
                                                     uint64_t* __restrict__ dst_keys,
                                                     uint32_t* __restrict__ dst_cnt)
{
    if (a.counters[2] == 0) return;  // no k-mer reached the spill table: it is empty
    if (d->n_jobs == 1) {  // one job: the result is out_keys (no gather)
        dst_keys = a.out_keys;
        dst_cnt = a.out_cnt;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n_slots; s += stride) {
        const uint32_t c = a.sp_cnt[s];
        if (!c || c < a.ci || (uint64_t)c > a.cx) continue;
        const uint32_t tag = (uint32_t)a.sp_keys[SW * s + (SW == 2 ? 0 : 2)];
        const uint64_t pos = job_off[tag] + job_item_tot[tag] + atomicAdd(&job_cur[tag], 1ull);
        if (SW == 2) {
            dst_keys[pos] = a.sp_keys[SW * s + 1];
        } else {
            dst_keys[2 * pos] = a.sp_keys[SW * s + 1];
            dst_keys[2 * pos + 1] = a.sp_keys[SW * s];
        }
        dst_cnt[pos] = c;
    }
}

void launch_spill_place(hipStream_t st, const CountArgs& a, const BatchDesc* d, uint32_t nw, uint64_t n_slots,
                        const uint64_t* job_off, const uint64_t* job_item_tot,
                        unsigned long long* job_cur, uint64_t* dst_keys, uint32_t* dst_cnt)
{
    uint64_t blocks = (n_slots + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (!blocks) return;
    if (nw == 1)
        k_spill_place<2><<<(unsigned)blocks, 256, 0, st>>>(a, d, n_slots, job_off, job_item_tot, job_cur,
                                                           dst_keys, dst_cnt);
    else
        k_spill_place<3><<<(unsigned)blocks, 256, 0, st>>>(a, d, n_slots, job_off, job_item_tot, job_cur,
                                                           dst_keys, dst_cnt);
}

}  // namespace tg
