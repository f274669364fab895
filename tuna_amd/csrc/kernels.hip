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

namespace tg {

// ============================================================== pack ====
__device__ __forceinline__ uint32_t pack4(uint32_t x)
{
    const uint32_t t = ((x >> 2) ^ (x >> 1)) & 0x03030303u;
    return ((t & 3u) << 6) | (((t >> 8) & 3u) << 4) | (((t >> 16) & 3u) << 2) | (t >> 24);
}

__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ b, uint64_t n,
                                              uint32_t* __restrict__ out, uint64_t n_words,
                                              int aligned)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_words; i += stride) {
        const uint64_t b0 = i * 16u;
        uint32_t w = 0;
        if (aligned && b0 + 16u <= n) {
            const uint4 v = *reinterpret_cast<const uint4*>(b + b0);
            w = (pack4(v.x) << 24) | (pack4(v.y) << 16) | (pack4(v.z) << 8) | pack4(v.w);
        } else {
            for (uint32_t j = 0; j < 16u; ++j)
                if (b0 + j < n) w |= kache_code(b[b0 + j]) << (30u - 2u * j);
        }
        out[i] = w;
    }
}

void launch_pack(hipStream_t st, const uint8_t* bases, uint64_t n, uint32_t* out, uint64_t n_words)
{
    const int aligned = ((uintptr_t)bases & 15u) == 0;
    uint64_t blocks = (n_words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    k_pack<<<(unsigned)blocks, 256, 0, st>>>(bases, n, out, n_words, aligned);
}

// ============================================================ strips ====
__global__ void k_chunk_counts(const uint64_t* __restrict__ off, uint64_t nch, uint32_t k,
                               uint32_t S, uint64_t* __restrict__ nstr, uint64_t* __restrict__ nkm)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nch; c += stride) {
        const uint64_t L = off[c + 1] - off[c];
        const uint64_t K = L >= k ? L - k + 1 : 0;
        nkm[c] = K;
        nstr[c] = (K + S - 1) / S;
    }
}

void launch_chunk_counts(hipStream_t st, const uint64_t* off, uint64_t n_chunks, uint32_t k,
                         uint32_t S, uint64_t* nstr, uint64_t* nkm)
{
    uint64_t blocks = (n_chunks + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    k_chunk_counts<<<(unsigned)blocks, 256, 0, st>>>(off, n_chunks, k, S, nstr, nkm);
}

__global__ void k_fill_strips(uint64_t nch, const uint64_t* __restrict__ nstr_excl,
                              const uint64_t* __restrict__ nkm_excl,
                              const uint64_t* __restrict__ nkm, uint32_t S,
                              uint32_t* __restrict__ strip_chunk, uint32_t* __restrict__ strip_tb,
                              uint64_t* __restrict__ strip_slot)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nch; c += stride) {
        const uint64_t K = nkm[c];
        const uint64_t base = nstr_excl[c], kb = nkm_excl[c];
        uint64_t i = 0;
        for (uint64_t t = 0; t < K; t += S, ++i) {
            strip_chunk[base + i] = (uint32_t)c;
            strip_tb[base + i] = (uint32_t)t;
            strip_slot[base + i] = kb + t;
        }
    }
}

void launch_fill_strips(hipStream_t st, uint64_t n_chunks, const uint64_t* nstr_excl,
                        const uint64_t* nkm_excl, const uint64_t* nkm, uint32_t S,
                        uint32_t* strip_chunk, uint32_t* strip_tb, uint64_t* strip_slot)
{
    uint64_t blocks = (n_chunks + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    k_fill_strips<<<(unsigned)blocks, 256, 0, st>>>(n_chunks, nstr_excl, nkm_excl, nkm, S,
                                                     strip_chunk, strip_tb, strip_slot);
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
#pragma unroll
    for (uint32_t i = 0; i < SCAN_I; ++i)
        if (base + i < n) s += in[base + i];
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
#pragma unroll
    for (uint32_t i = 0; i < SCAN_I; ++i) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    uint64_t tot;
    uint64_t run = block_excl_scan_u64(s, sh, tot) + (boff ? boff[blockIdx.x] : 0);
#pragma unroll
    for (uint32_t i = 0; i < SCAN_I; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
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

// ======================================================== superkmers ====
// Lane state for one strip: rolling ntHash over the packed stream plus the
// superkmer cut state.  The cut rule follows src/partition_hash.hpp:78-93: a
// superkmer starts at t = 0, where the window minimum hash changes, and where
// the current superkmer already holds capk = 256-k k-mers (the 255-base cap of
// the one-byte record header, :61-63,81).
struct SkLane {
    // packed-stream reader + rolling hash (include/nt_hash.hpp:114-118)
    const uint32_t* packed;
    uint64_t widx;
    uint32_t wrd, rem;
    uint64_t f, rv, lmer;
    uint32_t osh;
    // cut state
    uint32_t ts, tb, te, K;
    uint32_t t0, P0;
    uint64_t H0, prevH;
    bool have, done, flagged;
    uint32_t nrec;
    uint2* out;
    uint64_t cb;
};

__device__ __forceinline__ void sk_roll(SkLane& L, const uint64_t* tf, const uint64_t* tr)
{
    if (L.rem == 0) {
        L.wrd = L.packed[++L.widx];
        L.rem = 16;
    }
    const uint32_t kb = L.wrd >> 30;
    L.wrd <<= 2;
    --L.rem;
    const uint32_t in = kb ^ (kb >> 1);  // kache -> nt code
    const uint32_t o = (uint32_t)(L.lmer >> L.osh) & 3u;
    L.lmer = (L.lmer << 2) | in;
    const uint32_t ix = (o << 2) | in;
    L.f = rotl1(L.f) ^ tf[ix];
    L.rv = rotr1(L.rv) ^ tr[ix];
}

__device__ __forceinline__ uint32_t route_hsel(uint64_t H, uint32_t log2n, uint32_t log2s, uint32_t sb)
{
    const uint64_t n1 = (1ull << log2n) - 1u, s1 = (1ull << log2s) - 1u;
    const uint32_t pid = (uint32_t)(((H & n1) << log2s) | ((H >> log2n) & s1));
    const uint32_t split = (uint32_t)(H >> (log2n + log2s)) & ((1u << sb) - 1u);
    return (pid << sb) | split;
}

__device__ __forceinline__ void sk_emit(SkLane& L, const SkArgs& a, uint32_t tend)
{
    const uint32_t len = tend - L.t0 + a.k - 1u;
    const uint32_t mp = L.P0 - L.t0;
    const uint32_t hsel = route_hsel(L.H0, a.log2n, a.log2s, a.sb);
    L.out[L.nrec++] = make_uint2((uint32_t)(L.cb + L.t0),
                                 len | (mp << REC_MINPOS_SHIFT) | (hsel << REC_HSEL_SHIFT));
}

// Decision for k-mer t with window minimum (Hc, Pc).
__device__ __forceinline__ void sk_kmer(SkLane& L, const SkArgs& a, uint32_t t, uint64_t Hc, uint32_t Pc)
{
    if (t == L.ts) {
        if (L.ts == 0) {
            L.t0 = 0; L.H0 = Hc; L.P0 = Pc; L.have = true;
        }
    } else {
        const bool cut = (Hc != L.prevH) || (L.have && (t - L.t0 >= a.capk));
        if (cut) {
            if (L.have && L.t0 >= L.tb) sk_emit(L, a, t);
            if (t >= L.te) {
                L.done = true;
            } else {
                L.t0 = t; L.H0 = Hc; L.P0 = Pc; L.have = true;
            }
        } else if (t == L.tb && !L.have) {
            L.flagged = true;
            L.done = true;
        }
    }
    L.prevH = Hc;
    if (!L.done && t == L.K - 1u) {
        if (L.have && L.t0 >= L.tb) sk_emit(L, a, L.K);
        L.done = true;
    }
}

__device__ __forceinline__ void sk_init(SkLane& L, const SkArgs& a, uint32_t s, const uint64_t* s_init)
{
    const uint32_t c = a.strip_chunk[s];
    L.tb = a.strip_tb[s];
    L.cb = a.off[c];
    const uint32_t len = (uint32_t)(a.off[c + 1] - L.cb);
    L.K = len - a.k + 1u;
    L.te = min(L.tb + a.S, L.K);
    L.ts = 0;
    if (a.lookback != LOOKBACK_FULL && L.tb > a.lookback) L.ts = ((L.tb - a.lookback) / a.W) * a.W;
    L.out = a.rec + a.strip_slot[s];
    L.nrec = 0;
    L.packed = a.packed;
    const uint64_t gq = L.cb + L.ts;
    L.widx = gq >> 4;
    L.wrd = a.packed[L.widx] << (2u * (uint32_t)(gq & 15u));
    L.rem = 16u - (uint32_t)(gq & 15u);
    L.f = s_init[0];
    L.rv = s_init[1];
    L.lmer = 0;
    L.osh = 2u * (a.m - 1u);
    L.t0 = 0; L.P0 = 0; L.H0 = 0; L.prevH = 0;
    L.have = false; L.done = false; L.flagged = false;
}

__device__ __forceinline__ void sk_tables(const SkArgs& a, uint64_t* tf, uint64_t* tr, uint64_t* s_init)
{
    const uint32_t m = a.m;
    if (threadIdx.x < 16) {
        const uint32_t o = threadIdx.x >> 2, i = threadIdx.x & 3u;
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

template <int WMAX>
__global__ __launch_bounds__(256) void k_superkmers(SkArgs a)
{
    __shared__ uint64_t tf[16], tr[16], s_init[2];
    sk_tables(a, tf, tr, s_init);
    __syncthreads();
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    const uint32_t s = a.list ? a.list[idx] : idx;
    SkLane L;
    sk_init(L, a, s, s_init);
    const uint32_t W = a.W;

    for (uint32_t i = 1; i < a.m; ++i) sk_roll(L, tf, tr);
    uint64_t sv[WMAX];
    uint8_t sp[WMAX];
#pragma unroll
    for (int r = 0; r < WMAX; ++r) {
        if ((uint32_t)r < W) {
            sk_roll(L, tf, tr);
            sv[r] = L.f ^ L.rv;
        }
    }
    sk_backward<WMAX>(sv, sp, W);
    uint32_t bs = L.ts;
    uint64_t pv = 0;
    uint32_t pp = 0;
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
                sk_kmer(L, a, t, Hc, Pc);
                if (!L.done) {
                    sk_roll(L, tf, tr);
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
    a.nsk[s] = L.flagged ? 0u : L.nrec;
    if (L.flagged) {
        const uint32_t slot = atomicAdd(a.flag_count, 1u);
        a.flag_list[slot] = s;
    }
}

void launch_superkmers(hipStream_t st, const SkArgs& a)
{
    if (a.n == 0) return;
    const unsigned blocks = (a.n + 255u) / 256u;
    if (a.W <= 8) k_superkmers<8><<<blocks, 256, 0, st>>>(a);
    else if (a.W <= 16) k_superkmers<16><<<blocks, 256, 0, st>>>(a);
    else if (a.W <= 32) k_superkmers<32><<<blocks, 256, 0, st>>>(a);
    else k_superkmers<64><<<blocks, 256, 0, st>>>(a);
}

// =========================================================== routing ====
constexpr uint32_t ROUTE_T = 1024;

__global__ __launch_bounds__(ROUTE_T) void k_route_hist(RouteArgs a)
{
    extern __shared__ uint32_t sh[];  // [0, nsp): superkmers, [nsp, 2nsp): k-mers
    const uint32_t nsp = a.nsp;
    for (uint32_t i = threadIdx.x; i < 2u * nsp; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const uint32_t s0 = blockIdx.x * a.spt;
    const uint32_t s1 = min(s0 + a.spt, a.n_strips);
    const uint32_t shift = REC_HSEL_SHIFT + a.sb;
    for (uint32_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const uint32_t n = a.nsk[s];
        const uint2* r = a.rec + a.strip_slot[s];
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t y = r[i].y;
            const uint32_t pid = y >> shift;
            atomicAdd(&sh[pid], 1u);
            atomicAdd(&sh[nsp + pid], (y & REC_LEN_MASK) - a.k + 1u);
        }
    }
    __syncthreads();
    uint32_t* hs = a.hsk + (uint64_t)blockIdx.x * nsp;
    uint32_t* hk = a.hkm + (uint64_t)blockIdx.x * nsp;
    for (uint32_t i = threadIdx.x; i < nsp; i += blockDim.x) {
        hs[i] = sh[i];
        hk[i] = sh[nsp + i];
    }
}

void launch_route_hist(hipStream_t st, const RouteArgs& a, uint32_t ntiles)
{
    if (ntiles == 0) return;
    (void)hipFuncSetAttribute((const void*)k_route_hist, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(2u * a.nsp * sizeof(uint32_t)));
    k_route_hist<<<ntiles, ROUTE_T, 2u * a.nsp * sizeof(uint32_t), st>>>(a);
}

__global__ __launch_bounds__(256) void k_route_colscan(RouteArgs a, uint32_t ntiles,
                                                     uint64_t* __restrict__ part_sk,
                                                     uint64_t* __restrict__ part_km)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nsp) return;
    uint64_t run = 0, km = 0;
    for (uint32_t t = 0; t < ntiles; ++t) {
        const uint64_t i = (uint64_t)t * a.nsp + p;
        const uint32_t v = a.hsk[i];
        a.hsk[i] = (uint32_t)run;  // tile-exclusive offset inside partition p
        run += v;
        km += a.hkm[i];
    }
    part_sk[p] = run;
    part_km[p] = km;
}

void launch_route_colscan(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint64_t* part_sk,
                          uint64_t* part_km)
{
    k_route_colscan<<<(a.nsp + 255u) / 256u, 256, 0, st>>>(a, ntiles, part_sk, part_km);
}

__global__ __launch_bounds__(ROUTE_T) void k_route_scatter(RouteArgs a)
{
    extern __shared__ uint32_t cur[];  // nsp cursors
    const uint32_t nsp = a.nsp;
    const uint32_t* hs = a.hsk + (uint64_t)blockIdx.x * nsp;
    for (uint32_t i = threadIdx.x; i < nsp; i += blockDim.x) cur[i] = (uint32_t)a.part_off[i] + hs[i];
    __syncthreads();
    const uint32_t s0 = blockIdx.x * a.spt;
    const uint32_t s1 = min(s0 + a.spt, a.n_strips);
    const uint32_t shift = REC_HSEL_SHIFT + a.sb;
    for (uint32_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const uint32_t n = a.nsk[s];
        const uint2* r = a.rec + a.strip_slot[s];
        for (uint32_t i = 0; i < n; ++i) {
            const uint2 v = r[i];
            const uint32_t pos = atomicAdd(&cur[v.y >> shift], 1u);
            a.dst[pos] = v;
        }
    }
}

void launch_route_scatter(hipStream_t st, const RouteArgs& a, uint32_t ntiles)
{
    if (ntiles == 0) return;
    (void)hipFuncSetAttribute((const void*)k_route_scatter, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(a.nsp * sizeof(uint32_t)));
    k_route_scatter<<<ntiles, ROUTE_T, a.nsp * sizeof(uint32_t), st>>>(a);
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

// ========================================================== counting ====
constexpr uint32_t LOCKED32 = 0xFFFFFFFFu;

// Global spill table, one 64-bit word per key: claim by CAS on the key.
__device__ bool spill_insert1(const CountArgs& a, uint64_t x)
{
    uint64_t slot = mix64(x ^ 0x9e3779b97f4a7c15ull) & a.sp_mask;
    for (uint32_t p = 0; p <= a.sp_mask && p < (1u << 16); ++p) {
        unsigned long long* kp = (unsigned long long*)&a.sp_keys[slot];
        const uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == x) {
            atomicAdd(&a.sp_cnt[slot], 1u);
            return true;
        }
        if (cur == EMPTY64) {
            const uint64_t old = atomicCAS(kp, (unsigned long long)EMPTY64, (unsigned long long)x);
            if (old == EMPTY64 || old == x) {
                atomicAdd(&a.sp_cnt[slot], 1u);
                return true;
            }
        }
        slot = (slot + 1) & a.sp_mask;
    }
    atomicOr(a.sp_err, 1u);
    return false;
}

// Global spill table, two words per key: the count word is the slot state
// (0 empty, LOCKED32 being written, otherwise published count).
__device__ bool spill_insert2(const CountArgs& a, uint64_t hi, uint64_t lo)
{
    uint64_t slot = mix64(lo ^ mix64(hi ^ 0x9e3779b97f4a7c15ull)) & a.sp_mask;
    for (uint32_t p = 0; p <= a.sp_mask && p < (1u << 16);) {
        uint32_t* cp = &a.sp_cnt[slot];
        const uint32_t c = __hip_atomic_load(cp, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (c == 0u) {
            if (atomicCAS(cp, 0u, LOCKED32) == 0u) {
                __hip_atomic_store(&a.sp_keys[2 * slot], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&a.sp_keys[2 * slot + 1], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(cp, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                return true;
            }
            continue;
        }
        if (c == LOCKED32) continue;
        const uint64_t kh = __hip_atomic_load(&a.sp_keys[2 * slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t kl = __hip_atomic_load(&a.sp_keys[2 * slot + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (kh == hi && kl == lo) {
            atomicAdd(cp, 1u);
            return true;
        }
        slot = (slot + 1) & a.sp_mask;
        ++p;
    }
    atomicOr(a.sp_err, 1u);
    return false;
}

// LDS table, one word per key.  Slots move EMPTY -> key exactly once, so a key
// absent from a full probe window can never enter the table later: spilling
// it keeps every occurrence of a key in exactly one table.
__device__ __forceinline__ bool lds_insert1(uint64_t* keys, uint32_t* cnt, uint32_t mask,
                                            uint64_t x, uint32_t probe_max)
{
    uint32_t slot = (uint32_t)mix64(x) & mask;
    for (uint32_t p = 0; p < probe_max; ++p) {
        const uint64_t cur = __hip_atomic_load((unsigned long long*)&keys[slot], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == x) {
            atomicAdd(&cnt[slot], 1u);
            return true;
        }
        if (cur == EMPTY64) {
            const uint64_t old = atomicCAS((unsigned long long*)&keys[slot],
                                           (unsigned long long)EMPTY64, (unsigned long long)x);
            if (old == EMPTY64 || old == x) {
                atomicAdd(&cnt[slot], 1u);
                return true;
            }
        }
        slot = (slot + 1) & mask;
    }
    return false;
}

// LDS table, two words per key, count word as slot state.
__device__ __forceinline__ bool lds_insert2(uint64_t* khi, uint64_t* klo, uint32_t* cnt,
                                            uint32_t mask, uint64_t hi, uint64_t lo,
                                            uint32_t probe_max)
{
    uint32_t slot = (uint32_t)mix64(lo ^ mix64(hi)) & mask;
    for (uint32_t p = 0; p < probe_max;) {
        const uint32_t c = __hip_atomic_load(&cnt[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (c == 0u) {
            if (atomicCAS(&cnt[slot], 0u, LOCKED32) == 0u) {
                khi[slot] = hi;
                klo[slot] = lo;
                __hip_atomic_store(&cnt[slot], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                return true;
            }
            continue;
        }
        if (c == LOCKED32) continue;
        if (khi[slot] == hi && klo[slot] == lo) {
            atomicAdd(&cnt[slot], 1u);
            return true;
        }
        slot = (slot + 1) & mask;
        ++p;
    }
    return false;
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

// One workgroup per work item.  Waves flatten 64 records at a time into a
// k-mer index space (wave scan of k-mer counts) and every lane takes k-mers
// lane, lane+64, ... so superkmers of any length load-balance.
template <int NW, uint32_t CAP>
__global__ __launch_bounds__(COUNT_BLOCK) void k_count(CountArgs a)
{
    constexpr uint32_t NWAVE = COUNT_BLOCK / 64;
    __shared__ uint64_t s_k0[CAP];
    __shared__ uint64_t s_k1[NW == 2 ? CAP : 1];
    __shared__ uint32_t s_cnt[CAP];
    __shared__ uint32_t s_incl[NWAVE][64];
    __shared__ uint32_t s_start[NWAVE][64];
    __shared__ uint32_t s_red[NWAVE];
    __shared__ unsigned long long s_base;

    const uint4 it = a.items[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < CAP; i += blockDim.x) {
        if (NW == 1) s_k0[i] = EMPTY64;
        s_cnt[i] = 0;
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t k = a.k;
    const uint32_t r0 = it.x, r1 = it.y, smask = it.z, sval = it.w;
    const uint32_t split_mask_bits = 0;  // split bits sit at REC_HSEL_SHIFT
    (void)split_mask_bits;
    uint32_t spilled = 0;
    for (uint32_t rb = r0 + wv * 64u; rb < r1; rb += NWAVE * 64u) {
        const uint32_t i = rb + lane;
        uint32_t nk = 0, st = 0;
        if (i < r1) {
            const uint2 rr = a.rec[i];
            if (((rr.y >> REC_HSEL_SHIFT) & smask) == sval) {
                nk = (rr.y & REC_LEN_MASK) - k + 1u;
                st = rr.x;
            }
        }
        uint32_t incl = nk;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        s_incl[wv][lane] = incl;
        s_start[wv][lane] = st;
        const uint32_t total = __shfl(incl, 63, 64);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t j = lane; j < total; j += 64u) {
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1)
                if (s_incl[wv][lo + step - 1u] <= j) lo += step;
            const uint32_t ex = lo ? s_incl[wv][lo - 1u] : 0u;
            const uint64_t g = (uint64_t)s_start[wv][lo] + (j - ex);
            if (NW == 1) {
                const uint64_t x = canon64(load_kmer64(a.packed, g, k), k);
                if (!lds_insert1(s_k0, s_cnt, CAP - 1u, x, a.probe_max)) {
                    spill_insert1(a, x);
                    ++spilled;
                }
            } else {
                uint64_t hi, lo2;
                load_kmer128(a.packed, g, k, hi, lo2);
                canon128(hi, lo2, k);
                if (!lds_insert2(s_k0, s_k1, s_cnt, CAP - 1u, hi, lo2, a.probe_max)) {
                    spill_insert2(a, hi, lo2);
                    ++spilled;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // drain: count, reserve once per workgroup, write
    uint32_t mine = 0, uniq = 0;
    for (uint32_t s = threadIdx.x; s < CAP; s += blockDim.x) {
        const uint32_t c = s_cnt[s];
        if (c) {
            ++uniq;
            if (c >= a.ci && (uint64_t)c <= a.cx) ++mine;
        }
    }
    uint32_t tot_out, tot_uniq, tot_sp;
    const uint32_t my_off = block_excl_u32(mine, s_red, tot_out);
    block_excl_u32(uniq, s_red, tot_uniq);
    block_excl_u32(spilled, s_red, tot_sp);
    if (threadIdx.x == 0) {
        s_base = tot_out ? atomicAdd(&a.counters[0], (unsigned long long)tot_out) : 0ull;
        if (tot_uniq) atomicAdd(&a.counters[1], (unsigned long long)tot_uniq);
        if (tot_sp) atomicAdd(&a.counters[2], (unsigned long long)tot_sp);
    }
    __syncthreads();
    uint64_t pos = s_base + my_off;
    for (uint32_t s = threadIdx.x; s < CAP; s += blockDim.x) {
        const uint32_t c = s_cnt[s];
        if (c && c >= a.ci && (uint64_t)c <= a.cx) {
            emit_out<NW>(a, pos, NW == 2 ? s_k0[s] : 0ull, NW == 2 ? s_k1[s] : s_k0[s], c);
            ++pos;
        }
    }
}

void launch_count(hipStream_t st, const CountArgs& a, uint32_t nw)
{
    if (a.n_items == 0) return;
    if (nw == 1) k_count<1, LDS_CAP_NW1><<<a.n_items, COUNT_BLOCK, 0, st>>>(a);
    else k_count<2, LDS_CAP_NW2><<<a.n_items, COUNT_BLOCK, 0, st>>>(a);
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

}  // namespace tg
