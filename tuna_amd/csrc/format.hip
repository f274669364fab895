// format.hip — output formats of a result, built on the device (row A10):
// TSV lines "<k-mer>\t<count>\n" (src/count.hpp:180-216) and the records of a
// KFF raw section (src/kff_output.hpp:20-57).  The host only copies bytes.
//
// TSV: a line's length depends on its count's digits, so a tile of lines
// first reports its byte total, a scan places the tiles, and each workgroup
// then builds its lines in LDS and writes them out as aligned 32-bit words
// (bytes only at the two ends a neighbouring tile shares).  KFF records have a
// fixed size, ceil(k/4) key bytes then a 4-byte big-endian count.
#include "engine_impl.h"
#include "kff.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

namespace tg {

constexpr uint32_t FMT_BLOCK = 256;
constexpr uint32_t TSV_TILE = 512;                 // lines per workgroup (k <= 64)
constexpr uint32_t TSV_LINE_MAX = 64u + 1u + 10u + 1u;  // k <= 64, tab, count, newline
constexpr uint32_t TSV_TILE_W = 256;               // lines per workgroup (64 < k <= 256)
constexpr uint32_t TSV_LINE_MAX_W = 256u + 1u + 10u + 1u;
constexpr uint32_t KFF_REC_MAX = 64u + 4u;         // k <= 256

__device__ __forceinline__ uint32_t dec_digits(uint32_t c)
{
    uint32_t d = 1;
#pragma unroll
    for (uint32_t p = 10u, i = 0; i < 9; ++i, p *= 10u) d += c >= p ? 1u : 0u;
    return d;
}

// base j (0 = first) of a key: kache code at bits 2(k-1-j) of (hi:lo)
__device__ __forceinline__ uint32_t key_base(uint64_t lo, uint64_t hi, uint32_t k, uint32_t j)
{
    const uint32_t bit = 2u * (k - 1u - j);
    const uint64_t w = bit >= 64u ? hi : lo;
    return (uint32_t)(w >> (bit & 63u)) & 3u;
}

__device__ __forceinline__ void load_key(const uint64_t* keys, uint32_t nw, uint64_t e, uint64_t& lo,
                                         uint64_t& hi)
{
    lo = keys[e * nw];
    hi = nw == 2 ? keys[e * nw + 1] : 0ull;
}

// The k letters of key e (nw words, word 0 the lowest) into dst: one or two
// words from registers; wider keys one word at a time as the bases pass
// through it (32 bases per word, first base in the top word).
__device__ __forceinline__ void put_label(uint8_t* dst, const uint64_t* keys, uint32_t nw, uint64_t e,
                                          uint32_t k)
{
    constexpr uint32_t ACGT = 0x54474341u;  // "ACGT", byte b = letter of code b
    if (nw <= 2) {
        uint64_t lo, hi;
        load_key(keys, nw, e, lo, hi);
        for (uint32_t j = 0; j < k; ++j) dst[j] = (uint8_t)(ACGT >> (8u * key_base(lo, hi, k, j)));
        return;
    }
    uint32_t cw = 0xFFFFFFFFu;
    uint64_t w = 0;
    for (uint32_t j = 0; j < k; ++j) {
        const uint32_t bit = 2u * (k - 1u - j), wi = bit >> 6;
        if (wi != cw) {
            w = keys[e * nw + wi];
            cw = wi;
        }
        dst[j] = (uint8_t)(ACGT >> (8u * ((uint32_t)(w >> (bit & 63u)) & 3u)));
    }
}

// base j of key e, any width
__device__ __forceinline__ uint32_t key_base_w(const uint64_t* keys, uint32_t nw, uint64_t e, uint32_t k,
                                               uint32_t j)
{
    const uint32_t bit = 2u * (k - 1u - j);
    return (uint32_t)(keys[e * nw + (bit >> 6)] >> (bit & 63u)) & 3u;
}

template <uint32_t TILE>
__global__ __launch_bounds__(FMT_BLOCK) void k_tsv_len(const uint32_t* __restrict__ cnt, uint64_t n,
                                                       uint32_t k, uint64_t* __restrict__ tile_bytes)
{
    __shared__ uint32_t s_red[FMT_BLOCK / 64];
    const uint64_t t0 = (uint64_t)blockIdx.x * TILE;
    uint32_t sum = 0;
    for (uint32_t i = threadIdx.x; i < TILE; i += FMT_BLOCK)
        if (t0 + i < n) sum += k + 2u + dec_digits(cnt[t0 + i]);
#pragma unroll
    for (uint32_t d = 32; d; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if ((threadIdx.x & 63u) == 0) s_red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < FMT_BLOCK / 64; ++w) t += s_red[w];
        tile_bytes[blockIdx.x] = t;
    }
}

// Bytes [0, len) of the LDS image go to out[g0 .. g0+len); the image starts
// at byte a = g0 & 3 of an aligned word, so word i of the image is the
// aligned output word at (g0 - a) + 4i.  The first and last words may be
// shared with the neighbouring tiles: those are written byte by byte.
__device__ __forceinline__ void store_image(const uint8_t* img, uint32_t a, uint32_t len, uint8_t* out,
                                            uint64_t g0)
{
    if (len == 0) return;
    const uint64_t w0 = g0 - a;                 // aligned
    const uint32_t end = a + len;               // image bytes used: [a, end)
    const uint32_t nwords = (end + 3u) >> 2;
    for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) {
        const uint32_t b0 = 4u * i, b1 = b0 + 4u;
        if (b0 >= a && b1 <= end) {
            *reinterpret_cast<uint32_t*>(out + w0 + b0) = *reinterpret_cast<const uint32_t*>(img + b0);
        } else {
            for (uint32_t b = max(b0, a); b < min(b1, end); ++b) out[w0 + b] = img[b];
        }
    }
}

template <uint32_t TILE, uint32_t LINE>
__global__ __launch_bounds__(FMT_BLOCK) void k_tsv_write(const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ cnt, uint64_t n,
                                                         uint32_t k, uint32_t nw,
                                                         const uint64_t* __restrict__ tile_off,
                                                         uint8_t* __restrict__ out)
{
    constexpr uint32_t PER = TILE / FMT_BLOCK;
    __shared__ __align__(16) uint8_t s_img[TILE * LINE + 8];
    __shared__ uint32_t s_red[FMT_BLOCK / 64];
    const uint64_t t0 = (uint64_t)blockIdx.x * TILE;
    const uint64_t g0 = tile_off[blockIdx.x];
    const uint32_t a = (uint32_t)(g0 & 3u);
    // lines of this thread: t0 + threadIdx.x*PER + q (consecutive, so one
    // scan value per thread)
    uint32_t len[PER], mine = 0;
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint64_t e = t0 + threadIdx.x * PER + q;
        len[q] = e < n ? k + 2u + dec_digits(cnt[e]) : 0u;
        mine += len[q];
    }
    // block exclusive scan of mine
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t x = mine;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_red[wv] = x;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < FMT_BLOCK / 64; ++w) {
        if (w < wv) base += s_red[w];
        total += s_red[w];
    }
    uint32_t o = a + base + x - mine;
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint64_t e = t0 + threadIdx.x * PER + q;
        if (e >= n) break;
        put_label(s_img + o, keys, nw, e, k);
        o += k;
        s_img[o++] = '\t';
        uint32_t c = cnt[e];
        const uint32_t d = len[q] - k - 2u;
        for (uint32_t i = d; i; --i) {
            s_img[o + i - 1u] = (uint8_t)('0' + c % 10u);
            c /= 10u;
        }
        o += d;
        s_img[o++] = '\n';
    }
    __syncthreads();
    store_image(s_img, a, total, out, g0);
}

// KFF raw-section record of key e: ceil(k/4) bytes, base j at bits
// 6-2(j%4) of byte j/4 (first base in the top bits), zero padded, then the
// count as 4 big-endian bytes (src/kff_output.hpp:20-57).
__global__ __launch_bounds__(FMT_BLOCK) void k_kff_write(const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ cnt, uint64_t n,
                                                         uint32_t k, uint32_t nw, uint8_t* __restrict__ out)
{
    __shared__ __align__(16) uint8_t s_img[FMT_BLOCK * KFF_REC_MAX + 8];
    const uint32_t kb = (k + 3u) / 4u, rb = kb + 4u;
    const uint64_t r0 = (uint64_t)blockIdx.x * FMT_BLOCK;
    const uint64_t e = r0 + threadIdx.x;
    if (e < n) {
        uint64_t lo = 0, hi = 0;
        if (nw <= 2) load_key(keys, nw, e, lo, hi);
        uint8_t* r = s_img + threadIdx.x * rb;
        for (uint32_t b = 0; b < kb; ++b) {
            uint32_t v = 0;
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                const uint32_t j = 4u * b + t;
                if (j < k)
                    v |= (nw <= 2 ? key_base(lo, hi, k, j) : key_base_w(keys, nw, e, k, j)) << (6u - 2u * t);
            }
            r[b] = (uint8_t)v;
        }
        const uint32_t c = cnt[e];
        r[kb] = (uint8_t)(c >> 24);
        r[kb + 1] = (uint8_t)(c >> 16);
        r[kb + 2] = (uint8_t)(c >> 8);
        r[kb + 3] = (uint8_t)c;
    }
    __syncthreads();
    const uint64_t recs = n - r0 < FMT_BLOCK ? n - r0 : FMT_BLOCK;
    // the block's region starts at r0*rb, a multiple of 4 (FMT_BLOCK is)
    store_image(s_img, 0u, (uint32_t)(recs * rb), out, r0 * rb);
}

}  // namespace tg

using namespace tgi;

namespace {

// Entries formatted per device pass when a whole result is written.
constexpr uint64_t WRITE_CHUNK = 1ull << 21;

int check_fmt(tg_ctx* c, uint32_t k, uint64_t first, uint64_t n)
{
    if (!c) return fail(TG_EINVAL, "null context");
    if (k < 1 || k > 256) return fail(TG_EINVAL, "k must be in [1, 256] for formatting");
    if ((k + 31u) / 32u != c->nw && c->n_out) return fail(TG_EINVAL, "k does not match the result's key width");
    if (first > c->n_out || n > c->n_out - first) return fail(TG_EINVAL, "entries out of range");
    return TG_OK;
}

// TSV of entries [first, first+n) into the device buffer "fmt_out"; the byte
// total goes to *total (host, after a stream synchronisation).
int tsv_device(tg_ctx* c, uint32_t k, uint64_t first, uint64_t n, uint8_t** d_out, uint64_t* total)
{
    const uint64_t tile = k <= 64 ? tg::TSV_TILE : tg::TSV_TILE_W;
    const uint64_t tiles = (n + tile - 1) / tile;
    uint64_t *tb, *toff, *tmp;
    uint8_t* out;
    int rc;
    if ((rc = ensure_t(c, "fmt_tiles", tiles + 1, &tb)) || (rc = ensure_t(c, "fmt_toff", tiles + 1, &toff)) ||
        (rc = ensure_t(c, "fmt_scan", tg::scan_tmp_words(tiles) + 2, &tmp)) ||
        (rc = ensure_t(c, "fmt_out", n * (k + 12ull) + 16, &out)))
        return rc;
    const uint64_t* keys = c->d_keys + first * c->nw;
    const uint32_t* cnt = c->d_cnt + first;
    if (tiles) {
        if (k <= 64) tg::k_tsv_len<tg::TSV_TILE><<<(unsigned)tiles, tg::FMT_BLOCK, 0, c->stream>>>(cnt, n, k, tb);
        else tg::k_tsv_len<tg::TSV_TILE_W><<<(unsigned)tiles, tg::FMT_BLOCK, 0, c->stream>>>(cnt, n, k, tb);
        tg::launch_scan_u64(c->stream, tb, tiles, toff, tmp);
        if (k <= 64)
            tg::k_tsv_write<tg::TSV_TILE, tg::TSV_LINE_MAX>
                <<<(unsigned)tiles, tg::FMT_BLOCK, 0, c->stream>>>(keys, cnt, n, k, c->nw, toff, out);
        else
            tg::k_tsv_write<tg::TSV_TILE_W, tg::TSV_LINE_MAX_W>
                <<<(unsigned)tiles, tg::FMT_BLOCK, 0, c->stream>>>(keys, cnt, n, k, c->nw, toff, out);
        HIPOK(hipGetLastError());
        HIPOK(hipMemcpyAsync(total, toff + tiles, 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        *total = 0;
    }
    HIPOK(hipStreamSynchronize(c->stream));
    *d_out = out;
    return TG_OK;
}

int kff_device(tg_ctx* c, uint32_t k, uint64_t first, uint64_t n, uint8_t** d_out)
{
    const uint64_t rb = (k + 3) / 4 + 4;
    uint8_t* out;
    int rc;
    if ((rc = ensure_t(c, "fmt_out", n * rb + 16, &out))) return rc;
    const uint64_t blocks = (n + tg::FMT_BLOCK - 1) / tg::FMT_BLOCK;
    if (blocks)
        tg::k_kff_write<<<(unsigned)blocks, tg::FMT_BLOCK, 0, c->stream>>>(c->d_keys + first * c->nw,
                                                                          c->d_cnt + first, n, k, c->nw, out);
    HIPOK(hipGetLastError());
    *d_out = out;
    return TG_OK;
}

// ---- the TSV entry point that allocates nothing -------------------------
// tg_result_tsv_noalloc formats up to NOALLOC_TSV_MAX lines per call with all
// of its device scratch (tile byte totals, their scan, the formatted bytes) in
// one file-scope device array.  Every GPU holds its own copy of the array; its
// address on a device is taken once (hipGetSymbolAddress, at the first context
// created on that device) and only that address is handed to kernels and
// copies.  Calls on one GPU share the array: calls that overlap on one GPU race.
constexpr uint64_t NOALLOC_TSV_MAX = 1ull << 18;                   // lines per call
constexpr uint64_t NOALLOC_TILES = NOALLOC_TSV_MAX / tg::TSV_TILE;  // 512
constexpr uint64_t NOALLOC_TILE_BYTES = 8ull * (NOALLOC_TILES + 1);
constexpr uint64_t NOALLOC_SCAN_BYTES = 8ull * 1024;                // >= scan_tmp_words(512) + 2 words
constexpr uint64_t NOALLOC_OUT_OFF = 2 * NOALLOC_TILE_BYTES + NOALLOC_SCAN_BYTES;
constexpr uint64_t NOALLOC_OUT_BYTES = NOALLOC_TSV_MAX * tg::TSV_LINE_MAX + 256;
constexpr uint64_t SCRATCH_BYTES = (NOALLOC_OUT_OFF + NOALLOC_OUT_BYTES + 255) / 256 * 256;

__device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[MAX_DEVICES] = {};
hipError_t g_scratch_err[MAX_DEVICES] = {};

}  // namespace

namespace tgi {

// Called with `device` current, once per context creation; resolves the
// device's copy of the scratch array once per device.
void noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_scratch_err[device] = e;
        return;
    }
    g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
}

}  // namespace tgi

namespace {

}  // namespace

extern "C" {

int64_t tg_result_tsv(tg_ctx* c, uint32_t k, uint64_t first, uint64_t n, char* out, uint64_t cap)
{
    int rc;
    if ((rc = check_fmt(c, k, first, n))) return rc;
    HIPOK(hipSetDevice(c->device));
    uint8_t* d_out;
    uint64_t total = 0;
    if ((rc = tsv_device(c, k, first, n, &d_out, &total))) return rc;
    if (!out) return (int64_t)total;
    if (total > cap) return fail(TG_ECAP, "TSV needs %llu bytes, buffer holds %llu", (unsigned long long)total,
                                 (unsigned long long)cap);
    if (total) HIPOK(hipMemcpyAsync(out, d_out, total, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return (int64_t)total;
}

int64_t tg_result_kff(tg_ctx* c, uint32_t k, uint64_t first, uint64_t n, uint8_t* out, uint64_t cap)
{
    int rc;
    if ((rc = check_fmt(c, k, first, n))) return rc;
    const uint64_t total = n * ((k + 3) / 4 + 4ull);
    if (!out) return (int64_t)total;
    if (total > cap) return fail(TG_ECAP, "KFF records need %llu bytes, buffer holds %llu",
                                 (unsigned long long)total, (unsigned long long)cap);
    HIPOK(hipSetDevice(c->device));
    uint8_t* d_out;
    if ((rc = kff_device(c, k, first, n, &d_out))) return rc;
    if (total) HIPOK(hipMemcpyAsync(out, d_out, total, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return (int64_t)total;
}

int tg_result_write(tg_ctx* c, uint32_t k, const char* path, int kff)
{
    int rc;
    if ((rc = check_fmt(c, k, 0, 0))) return rc;
    if (!path) return fail(TG_EINVAL, "null path");
    HIPOK(hipSetDevice(c->device));
    FILE* f = fopen(path, "wb");
    if (!f) return fail(TG_EINPUT, "cannot open output file: %s", path);
    const uint64_t n = c->n_out;
    const uint64_t line_max = kff ? (k + 3) / 4 + 4ull : k + 12ull;
    const uint64_t chunk_bytes = WRITE_CHUNK * line_max;
    // two pinned host buffers: the file write of one chunk overlaps the
    // device formatting and copy of the next
    void* h[2] = {nullptr, nullptr};
    uint64_t h_len[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    rc = TG_OK;
    auto io_fail = [&](const char* what) {
        rc = fail(TG_EINPUT, "%s output file: %s", what, path);
    };
    do {
        if (hipHostMalloc(&h[0], chunk_bytes) != hipSuccess || hipHostMalloc(&h[1], chunk_bytes) != hipSuccess ||
            hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            rc = fail(TG_ENOMEM, "pinned staging for the output failed");
            break;
        }
        if (kff) {
            const std::string hdr = kff_header(k, n);
            if (fwrite(hdr.data(), 1, hdr.size(), f) != hdr.size()) {
                io_fail("failed while writing");
                break;
            }
        }
        int pending = -1;  // buffer whose copy is in flight
        for (uint64_t first = 0, i = 0; first < n && rc == TG_OK; first += WRITE_CHUNK, ++i) {
            const uint64_t m = std::min(WRITE_CHUNK, n - first);
            const int b = (int)(i & 1u);
            uint8_t* d_out = nullptr;
            uint64_t bytes = m * line_max;
            // formatting waits for the previous copy (stream order), which
            // frees buffer b^1 for its file write below
            if (kff) rc = kff_device(c, k, first, m, &d_out);
            else rc = tsv_device(c, k, first, m, &d_out, &bytes);
            if (rc) break;
            if (hipMemcpyAsync(h[b], d_out, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipEventRecord(ev[b], c->stream) != hipSuccess) {
                rc = fail(TG_EDEVICE, "result copy failed: %s", hipGetErrorString(hipGetLastError()));
                break;
            }
            h_len[b] = bytes;
            if (pending >= 0) {
                if (hipEventSynchronize(ev[pending]) != hipSuccess) {
                    rc = fail(TG_EDEVICE, "result copy failed");
                    break;
                }
                if (fwrite(h[pending], 1, h_len[pending], f) != h_len[pending]) {
                    io_fail("failed while writing");
                    break;
                }
            }
            pending = b;
        }
        if (rc == TG_OK && pending >= 0) {
            if (hipEventSynchronize(ev[pending]) != hipSuccess) rc = fail(TG_EDEVICE, "result copy failed");
            else if (fwrite(h[pending], 1, h_len[pending], f) != h_len[pending]) io_fail("failed while writing");
        }
        if (rc == TG_OK && kff) {
            const std::string ft = kff_footer();
            if (fwrite(ft.data(), 1, ft.size(), f) != ft.size()) io_fail("failed while writing");
        }
    } while (false);
    (void)hipStreamSynchronize(c->stream);
    for (int b = 0; b < 2; ++b) {
        if (h[b]) (void)hipHostFree(h[b]);
        if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
    if (fclose(f) != 0 && rc == TG_OK) io_fail("failed while writing");
    return rc;
}

int64_t tg_result_tsv_noalloc(tg_ctx* c, uint32_t k, uint64_t first, uint64_t n, char* out, uint64_t cap)
{
    int rc;
    if ((rc = check_fmt(c, k, first, n))) return rc;
    if (k > 64)
        return fail(TG_EINVAL, "tg_result_tsv_noalloc formats k <= 64 (its fixed scratch holds lines of at "
                               "most 76 bytes); use tg_result_tsv for k = %u", k);
    if (n > NOALLOC_TSV_MAX)
        return fail(TG_EINVAL, "tg_result_tsv_noalloc formats at most %llu lines per call (%llu asked)",
                    (unsigned long long)NOALLOC_TSV_MAX, (unsigned long long)n);
    if (c->device < 0 || c->device >= MAX_DEVICES) return fail(TG_EINVAL, "device %d out of range", c->device);
    uint8_t* base;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        base = g_scratch_ptr[c->device];
        if (!base)
            return fail(TG_EDEVICE, "scratch array of device %d unavailable: %s", c->device,
                        hipGetErrorString(g_scratch_err[c->device]));
    }
    HIPOK(hipSetDevice(c->device));
    const uint64_t tiles = (n + tg::TSV_TILE - 1) / tg::TSV_TILE;
    uint64_t* tb = reinterpret_cast<uint64_t*>(base);
    uint64_t* toff = reinterpret_cast<uint64_t*>(base + NOALLOC_TILE_BYTES);
    uint64_t* tmp = reinterpret_cast<uint64_t*>(base + 2 * NOALLOC_TILE_BYTES);
    uint8_t* d_out = base + NOALLOC_OUT_OFF;
    uint64_t total = 0;
    if (tiles) {
        tg::k_tsv_len<tg::TSV_TILE><<<(unsigned)tiles, tg::FMT_BLOCK, 0, c->stream>>>(c->d_cnt + first, n, k, tb);
        tg::launch_scan_u64(c->stream, tb, tiles, toff, tmp);
        tg::k_tsv_write<tg::TSV_TILE, tg::TSV_LINE_MAX>
            <<<(unsigned)tiles, tg::FMT_BLOCK, 0, c->stream>>>(c->d_keys + first * c->nw,
                                                                           c->d_cnt + first, n, k, c->nw,
                                                                           toff, d_out);
        HIPOK(hipGetLastError());
        HIPOK(hipMemcpyAsync(&total, toff + tiles, 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPOK(hipStreamSynchronize(c->stream));
    if (!out) return (int64_t)total;
    if (total > cap) return fail(TG_ECAP, "TSV needs %llu bytes, buffer holds %llu", (unsigned long long)total,
                                 (unsigned long long)cap);
    if (total) HIPOK(hipMemcpyAsync(out, d_out, total, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return (int64_t)total;
}

}  // extern "C"
