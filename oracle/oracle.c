/*
 * oracle.c — CPU restatement of tuna's k-mer counting path (reference
 * vicLeva/tuna v1.4.0).  TEST INFRASTRUCTURE ONLY: the product never links or
 * calls this file; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it as the checker.
 *
 * Pinning (what fixes each function independently of its own formula):
 *   orc_parse          — hand fixtures in tests/golden (N-splits, lowercase,
 *                        multi-line FASTA, '>' mid-line, '\r', FASTQ N,
 *                        64-byte-boundary EOF) whose expected chunks come from
 *                        the reference's own SeqSource/helicase parser
 *                        (oracle/_ref build), plus the written rules.
 *   orc_nthash_*       — rolling vs closed form, strand symmetry h(s)==h(rc s),
 *                        and the reference's own Roller (oracle/_ref).
 *   orc_window_*       — streaming two-stack vs brute-force normative rule on
 *                        adversarial inputs (homopolymers, repeats, equal
 *                        hashes); reference MinimizerWindow via oracle/_ref.
 *   orc_superkmers     — reference extract_superkmers_from_actg output
 *                        (oracle/_ref), totals identities (sum of k-mers).
 *   orc_pack_record    — reference SuperkmerWriter bytes (oracle/_ref).
 *   orc_canonical/count— textbook sort-and-count; invariants canon(x) ==
 *                        canon(rc x); reference kache table via oracle/_ref.
 *   orc_tsv_line       — label round trip; reference Kmer::get_label (_ref).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- A1 ---- */
/* ACGT test, case-insensitive: src/seq_source.hpp:33-39 (split_non_actg),
 * include/helicase_cpp/simd.hpp:305-307,330. */
static inline int is_acgt(uint8_t c)
{
    const uint8_t u = c & 0xDFu;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

/* FASTA rules: include/helicase_cpp/lexer.hpp:46-71 (header = '>' through the
 * next '\n'), parser_fasta.hpp:233-316 (content before the first header is
 * skipped, '\n' inside sequence joins lines, any other non-ACGT byte or a
 * header ends the chunk). */
static int64_t parse_fasta(const uint8_t* d, uint64_t len, uint8_t* ob,
                           uint64_t* off, uint64_t cap)
{
    uint64_t i = 0, nb = 0, nc = 0;
    off[0] = 0;
    while (i < len && d[i] != '>') ++i;
    while (i < len) {
        /* header: through the next line feed (inclusive) */
        while (i < len && d[i] != '\n') ++i;
        ++i;
        while (i < len && d[i] != '>') {
            if (is_acgt(d[i])) {
                while (i < len) {
                    const uint8_t c = d[i];
                    if (is_acgt(c)) { ob[nb++] = c; ++i; }
                    else if (c == '\n') ++i;
                    else break;
                }
                if (nc + 1 >= cap) return -1;
                off[++nc] = nb;
            } else {
                ++i;
            }
        }
    }
    return (int64_t)nc;
}

static inline uint64_t find_lf(const uint8_t* d, uint64_t len, uint64_t from)
{
    const void* p = from < len ? memchr(d + from, '\n', len - from) : NULL;
    return p ? (uint64_t)((const uint8_t*)p - d) : len;
}

/* FASTQ rules: include/helicase_cpp/parser_fastq.hpp:112-285.  Strict 4-line
 * records; only line 2 is DNA, split at non-ACGT.  Two block-boundary
 * behaviours of the 64-byte lexer are reproduced:
 *   - a DNA run that reaches end-of-file exactly at a multiple of 64 bytes is
 *     not emitted (:181, advance_block() returns before the chunk event);
 *   - on a header line the first byte is stepped over by increment_pos() (:115);
 *     when that byte is a '\n' at offset 63 of a block the line feed is lost
 *     with the block, so the header runs to the following '\n'. */
static int64_t parse_fastq(const uint8_t* d, uint64_t len, uint8_t* ob,
                           uint64_t* off, uint64_t cap)
{
    uint64_t pos = 0, nb = 0, nc = 0;
    unsigned line = 0;
    off[0] = 0;
    for (;;) {
        switch (line & 3u) {
        case 0: {
            if (pos + 1 >= len) return (int64_t)nc;
            const uint64_t from = (pos % 64 == 63) ? pos + 1 : pos;
            const uint64_t lf = find_lf(d, len, from);
            if (lf >= len) return (int64_t)nc;
            pos = lf + 1; ++line;
            break;
        }
        case 1: {
            uint64_t s = pos;
            while (s < len && !is_acgt(d[s]) && d[s] != '\n') ++s;
            if (s >= len) return (int64_t)nc;
            if (d[s] == '\n') { pos = s + 1; ++line; break; }
            uint64_t r = s;
            while (r < len && is_acgt(d[r])) ++r;
            if (r == len && len % 64 == 0) return (int64_t)nc;
            memcpy(ob + nb, d + s, r - s);
            nb += r - s;
            if (nc + 1 >= cap) return -1;
            off[++nc] = nb;
            pos = r;
            if (r < len && d[r] == '\n') { pos = r + 1; ++line; }
            break;
        }
        default: {
            const uint64_t lf = find_lf(d, len, pos);
            if (lf >= len) return (int64_t)nc;
            pos = lf + 1; ++line;
            break;
        }
        }
    }
}

/* Format by first byte: src/seq_source.hpp:169-185. Empty input throws
 * "Empty file" in the reference (include/helicase_cpp/input.hpp:107). */
int64_t orc_parse(const uint8_t* data, uint64_t len, uint8_t* out_bases,
                  uint64_t* out_off, uint64_t off_cap)
{
    if (len == 0) return -2;
    if (off_cap < 1) return -1;
    return data[0] == '@' ? parse_fastq(data, len, out_bases, out_off, off_cap)
                          : parse_fasta(data, len, out_bases, out_off, off_cap);
}

/* ---------------------------------------------------------------- A2 ---- */
/* kache: A0 C1 G2 T3 (src/partition_hash.hpp:69-70); nt: A0 C1 T2 G3
 * (include/nt_hash.hpp:53-55). */
uint32_t orc_kache_code(uint8_t c) { return ((c >> 2) ^ (c >> 1)) & 3u; }
uint32_t orc_nt_code(uint8_t c) { return (c >> 1) & 3u; }

/* ---------------------------------------------------------------- A3 ---- */
/* Published ntHash seeds for A, C, T, G (Mohamadi et al. 2016), indexed by the
 * nt code; include/nt_hash.hpp:26-40.  The reverse strand uses the seed of the
 * complement, nt code b^2. */
static const uint64_t NT_SEED[4] = {
    0x3c8bfbb395c60474ULL, 0x3193c18562a02b4cULL,
    0x295549f54be24456ULL, 0x20323ed082572324ULL,
};

static inline uint64_t rol(uint64_t x, unsigned s)
{
    s &= 63u;
    return s ? (x << s) | (x >> (64u - s)) : x;
}
static inline uint64_t ror(uint64_t x, unsigned s)
{
    s &= 63u;
    return s ? (x >> s) | (x << (64u - s)) : x;
}

/* fwd = XOR_i rol(F[s_i], m-1-i); rev = XOR_i rol(F[s_i ^ 2], i);
 * canonical = fwd ^ rev (include/nt_hash.hpp:71-72,104-111,123). */
uint64_t orc_nthash_direct(const uint8_t* a, int m)
{
    uint64_t f = 0, r = 0;
    for (int i = 0; i < m; ++i) {
        const uint32_t b = orc_nt_code(a[i]);
        f ^= rol(NT_SEED[b], (unsigned)(m - 1 - i));
        r ^= rol(NT_SEED[b ^ 2u], (unsigned)i);
    }
    return f ^ r;
}

/* Rolling update, include/nt_hash.hpp:114-118. */
void orc_nthash_roll(const uint8_t* a, uint64_t len, int m, uint64_t* out)
{
    if (len < (uint64_t)m) return;
    uint64_t f = 0, r = 0;
    for (int i = 0; i < m; ++i) {
        const uint32_t b = orc_nt_code(a[i]);
        f ^= rol(NT_SEED[b], (unsigned)(m - 1 - i));
        r ^= rol(NT_SEED[b ^ 2u], (unsigned)i);
    }
    out[0] = f ^ r;
    for (uint64_t j = 1; j + m <= len; ++j) {
        const uint32_t o = orc_nt_code(a[j - 1]), n = orc_nt_code(a[j + m - 1]);
        f = rol(f, 1) ^ rol(NT_SEED[o], (unsigned)m) ^ NT_SEED[n];
        r = ror(r, 1) ^ ror(NT_SEED[o ^ 2u], 1) ^ rol(NT_SEED[n ^ 2u], (unsigned)(m - 1));
        out[j] = f ^ r;
    }
}

/* ---------------------------------------------------------------- A4 ---- */
/* Streaming restatement of include/minimizer_window.hpp:55-199.  The window
 * holds W hashes in a ring written right-to-left; at every wrap the prefix
 * minima of the whole ring are rebuilt (reset_windows :66-78, strict '<' keeps
 * the lower ring index = the later position), new hashes feed a running
 * minimum (advance_impl :86-91, strict '<' keeps the first written).  The
 * query (:176-178, :191-199) takes the older-part minimum only when strictly
 * smaller. */
void orc_window_stream(const uint64_t* h, uint64_t M, int W, uint64_t* H, uint32_t* pos)
{
    if (M < (uint64_t)W) return;
    const int w = W - 1;
    uint64_t* ring = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W);
    uint64_t* pre = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W);
    int* pre_idx = (int*)malloc(sizeof(int) * (size_t)W);
    uint64_t suf = 0, h0pos;
    int suf_idx = 0, pivot;
#define REBUILD()                                                     \
    do {                                                              \
        pre[0] = ring[0]; pre_idx[0] = 0;                             \
        for (int i_ = 1; i_ <= w; ++i_) {                             \
            const int lt_ = ring[i_] < pre[i_ - 1];                   \
            pre[i_] = lt_ ? ring[i_] : pre[i_ - 1];                   \
            pre_idx[i_] = lt_ ? i_ : pre_idx[i_ - 1];                 \
        }                                                             \
        suf = UINT64_MAX; suf_idx = 0; pivot = w;                     \
    } while (0)
    /* reset(): first W m-mers, position p stored at ring[w - p] */
    for (int p = 0; p <= w; ++p) ring[w - p] = h[p];
    h0pos = (uint64_t)w;
    REBUILD();
    for (uint64_t t = 0;; ++t) {
        const int use_pre = pre[pivot] < suf;
        H[t] = use_pre ? pre[pivot] : suf;
        pos[t] = use_pre ? (uint32_t)(h0pos - (uint64_t)pre_idx[pivot])
                         : (uint32_t)(h0pos + (uint64_t)w + 1u - (uint64_t)suf_idx);
        if (t + (uint64_t)W >= M) break;
        const uint64_t x = h[t + (uint64_t)W];
        ring[pivot] = x;
        if (x < suf) { suf = x; suf_idx = pivot; }
        if (pivot > 0) --pivot;
        else { h0pos += (uint64_t)w + 1u; REBUILD(); }
    }
#undef REBUILD
    free(ring); free(pre); free(pre_idx);
}

/* The same rule stated per window: blocks of W m-mers anchored at the chunk
 * start; window of k-mer t = A (tail of block floor(t/W), from t) + B (head of
 * the next block).  Minimum from A only if strictly smaller, then its
 * rightmost argmin; otherwise the leftmost argmin of B. */
void orc_window_brute(const uint64_t* h, uint64_t M, int W, uint64_t* H, uint32_t* pos)
{
    if (M < (uint64_t)W) return;
    for (uint64_t t = 0; t + (uint64_t)W <= M; ++t) {
        const uint64_t bend = (t / (uint64_t)W + 1u) * (uint64_t)W; /* first index of B */
        uint64_t ma = UINT64_MAX, pa = 0, mb = UINT64_MAX, pb = 0;
        int has_b = 0;
        for (uint64_t j = t; j < bend; ++j)
            if (h[j] <= ma) { ma = h[j]; pa = j; }          /* rightmost */
        for (uint64_t j = bend; j < t + (uint64_t)W; ++j) {
            if (!has_b || h[j] < mb) { mb = h[j]; pb = j; } /* leftmost */
            has_b = 1;
        }
        if (!has_b || ma < mb) { H[t] = ma; pos[t] = (uint32_t)pa; }
        else { H[t] = mb; pos[t] = (uint32_t)pb; }
    }
}

/* ---------------------------------------------------------------- A5 ---- */
/* src/partition_hash.hpp:48-101: a new superkmer starts at t = 0, wherever the
 * window hash changes, and when the current one already spans HDR_MAX bases
 * (:81, HDR_MAX = 255 for the 1-byte header, 65535 for the 2-byte one). */
int64_t orc_superkmers(const uint8_t* bases, const uint64_t* off, uint64_t n_chunks,
                       int k, int m, uint64_t* sk_start, uint32_t* sk_len,
                       uint32_t* sk_minpos, uint64_t* sk_hash, uint64_t cap,
                       uint64_t* kmers_out)
{
    const int W = k - m + 1;
    const uint64_t hdr_max = (2u * (unsigned)k - (unsigned)m <= 255u) ? 255u : 65535u;
    uint64_t n = 0, kmers = 0;
    uint64_t maxlen = 0;
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (off[c + 1] - off[c] > maxlen) maxlen = off[c + 1] - off[c];
    uint64_t* h = (uint64_t*)malloc(sizeof(uint64_t) * (maxlen + 1));
    uint64_t* H = (uint64_t*)malloc(sizeof(uint64_t) * (maxlen + 1));
    uint32_t* P = (uint32_t*)malloc(sizeof(uint32_t) * (maxlen + 1));
    for (uint64_t c = 0; c < n_chunks; ++c) {
        const uint64_t L = off[c + 1] - off[c];
        if (L < (uint64_t)k) continue;
        const uint8_t* s = bases + off[c];
        orc_nthash_roll(s, L, m, h);
        orc_window_stream(h, L - (uint64_t)m + 1u, W, H, P);
        const uint64_t K = L - (uint64_t)k + 1u;
        uint64_t t0 = 0;
        for (uint64_t t = 1; t <= K; ++t) {
            /* base index of the k-mer's last base is t + k - 1 ("pos") */
            const int cut = (t == K) || H[t] != H[t0] ||
                            (t + (uint64_t)k - 1u - t0 >= hdr_max);
            if (!cut) continue;
            if (n >= cap) { free(h); free(H); free(P); return -1; }
            sk_start[n] = off[c] + t0;
            sk_len[n] = (uint32_t)(t - t0 + (uint64_t)k - 1u);
            sk_minpos[n] = (uint32_t)(P[t0] - t0);
            sk_hash[n] = H[t0];
            kmers += t - t0;
            ++n;
            t0 = t;
        }
    }
    free(h); free(H); free(P);
    if (kmers_out) *kmers_out = kmers;
    return (int64_t)n;
}

/* ---------------------------------------------------------------- A6 ---- */
/* src/superkmer_io.hpp:47-48 (header type), :154-172 (append_kache). */
uint64_t orc_pack_record(const uint8_t* a, uint32_t len, uint32_t min_pos,
                         int k, int m, uint8_t* out)
{
    const int wide = (2u * (unsigned)k - (unsigned)m) > 255u;
    uint64_t o = 0;
    if (wide) {
        out[o++] = (uint8_t)(len & 0xFF); out[o++] = (uint8_t)(len >> 8);
        out[o++] = (uint8_t)(min_pos & 0xFF); out[o++] = (uint8_t)(min_pos >> 8);
    } else {
        out[o++] = (uint8_t)len; out[o++] = (uint8_t)min_pos;
    }
    const uint64_t nbytes = (len + 3u) / 4u;
    memset(out + o, 0, nbytes);
    for (uint32_t i = 0; i < len; ++i)
        out[o + i / 4] |= (uint8_t)(orc_kache_code(a[i]) << (6u - 2u * (i % 4u)));
    return o + nbytes;
}

/* ---------------------------------------------------------------- A8 ---- */
/* include/kache-hash/Kmer.hpp:44-48 (layout), :410-439 (reverse complement,
 * complement 3-b per DNA_Utility.hpp:47-50), :442-456 + :642-646 (canonical =
 * smaller of the two compared from the top word). */
typedef unsigned __int128 u128;

static inline u128 kmer_fwd(const uint8_t* a, int k)
{
    u128 x = 0;
    for (int i = 0; i < k; ++i) x = (x << 2) | orc_kache_code(a[i]);
    return x;
}
static inline u128 kmer_rc(const uint8_t* a, int k)
{
    u128 x = 0;
    for (int i = k - 1; i >= 0; --i) x = (x << 2) | (3u - orc_kache_code(a[i]));
    return x;
}

/* k > 64: ceil(k/32) words, word 0 the lowest (the same layout, Kmer.hpp:44-48) */
#define ORC_MAXW 8
static void kw_push(uint64_t* w, int nw, uint64_t b)
{
    for (int i = nw - 1; i > 0; --i) w[i] = (w[i] << 2) | (w[i - 1] >> 62);
    w[0] = (w[0] << 2) | b;
}
static void kw_fwd(const uint8_t* a, int k, int nw, uint64_t* w)
{
    memset(w, 0, sizeof(uint64_t) * (size_t)nw);
    for (int i = 0; i < k; ++i) kw_push(w, nw, orc_kache_code(a[i]));
}
static void kw_rc(const uint8_t* a, int k, int nw, uint64_t* w)
{
    memset(w, 0, sizeof(uint64_t) * (size_t)nw);
    for (int i = k - 1; i >= 0; --i) kw_push(w, nw, 3u - orc_kache_code(a[i]));
}
static int kw_cmp(const uint64_t* a, const uint64_t* b, int nw)
{
    for (int i = nw - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

void orc_canonical(const uint8_t* a, int k, uint64_t* words)
{
    if (k > 64) {
        const int nw = (k + 31) / 32;
        uint64_t f[ORC_MAXW], r[ORC_MAXW];
        kw_fwd(a, k, nw, f);
        kw_rc(a, k, nw, r);
        memcpy(words, kw_cmp(f, r, nw) <= 0 ? f : r, sizeof(uint64_t) * (size_t)nw);
        return;
    }
    const u128 f = kmer_fwd(a, k), r = kmer_rc(a, k);
    const u128 c = f < r ? f : r;
    words[0] = (uint64_t)c;
    if (k > 32) words[1] = (uint64_t)(c >> 64);
}

/* ---------------------------------------------------------------- A9 ---- */
/* Textbook count: every canonical k-mer of every chunk of length >= k
 * (src/partition_hash.hpp:65), sorted, runs counted.  The reference table's
 * observable semantics (Streaming_Kmer_Hash_Table.hpp:1095-1136, value 1 on
 * insert, +1 after) are exactly the run lengths. */
static void radix_u64(uint64_t* a, uint64_t* tmp, uint64_t n, int bits)
{
    size_t cnt[1 << 16];
    for (int sh = 0; sh < bits; sh += 16) {
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; ++i) cnt[(a[i] >> sh) & 0xFFFF]++;
        size_t s = 0;
        for (int d = 0; d < (1 << 16); ++d) { const size_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint64_t i = 0; i < n; ++i) tmp[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
        memcpy(a, tmp, n * sizeof(uint64_t));
    }
}

static int cmp_u128(const void* x, const void* y)
{
    const u128 a = *(const u128*)x, b = *(const u128*)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}

static int g_kw_nw; /* key words of the multi-word sort (k > 64) */
static int cmp_kw(const void* x, const void* y)
{
    return kw_cmp((const uint64_t*)x, (const uint64_t*)y, g_kw_nw);
}

int64_t orc_count(const uint8_t* bases, const uint64_t* off, uint64_t n_chunks,
                  int k, uint32_t ci, uint64_t cx, uint64_t* keys,
                  uint32_t* counts, uint64_t cap, uint64_t* total_kmers,
                  uint64_t* n_unique)
{
    uint64_t K = 0;
    for (uint64_t c = 0; c < n_chunks; ++c) {
        const uint64_t L = off[c + 1] - off[c];
        if (L >= (uint64_t)k) K += L - (uint64_t)k + 1u;
    }
    if (total_kmers) *total_kmers = K;
    int64_t nout = 0;
    uint64_t nu = 0;
    if (k > 64) {
        /* multi-word keys: every canonical k-mer written out, sorted, runs counted */
        const int nw = (k + 31) / 32;
        uint64_t* a = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nw * (K + 1));
        uint64_t i = 0;
        for (uint64_t c = 0; c < n_chunks; ++c) {
            const uint64_t L = off[c + 1] - off[c];
            if (L < (uint64_t)k) continue;
            for (uint64_t j = 0; j + (uint64_t)k <= L; ++j) orc_canonical(bases + off[c] + j, k, a + nw * i++);
        }
        g_kw_nw = nw;
        qsort(a, K, sizeof(uint64_t) * (size_t)nw, cmp_kw);
        for (uint64_t x = 0; x < K;) {
            uint64_t y = x + 1;
            while (y < K && kw_cmp(a + nw * y, a + nw * x, nw) == 0) ++y;
            const uint64_t cnt = y - x;
            ++nu;
            if (cnt >= ci && cnt <= cx) {
                if ((uint64_t)nout >= cap) { free(a); return -1; }
                memcpy(keys + (uint64_t)nw * (uint64_t)nout, a + nw * x, sizeof(uint64_t) * (size_t)nw);
                counts[nout] = (uint32_t)cnt;
                ++nout;
            }
            x = y;
        }
        free(a);
    } else if (k <= 32) {
        uint64_t* a = (uint64_t*)malloc(sizeof(uint64_t) * (K + 1));
        uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * (K + 1));
        const uint64_t mask = k == 32 ? ~0ULL : ((1ULL << (2 * k)) - 1u);
        uint64_t i = 0;
        for (uint64_t c = 0; c < n_chunks; ++c) {
            const uint64_t L = off[c + 1] - off[c];
            if (L < (uint64_t)k) continue;
            const uint8_t* s = bases + off[c];
            uint64_t f = 0, r = 0;
            for (int j = 0; j < k - 1; ++j) {
                const uint64_t b = orc_kache_code(s[j]);
                f = ((f << 2) | b) & mask;
                r = (r >> 2) | ((3u - b) << (2 * (k - 1)));
            }
            for (uint64_t j = (uint64_t)k - 1u; j < L; ++j) {
                const uint64_t b = orc_kache_code(s[j]);
                f = ((f << 2) | b) & mask;
                r = (r >> 2) | ((3u - b) << (2 * (k - 1)));
                a[i++] = f < r ? f : r;
            }
        }
        radix_u64(a, t, K, 2 * k);
        for (uint64_t x = 0; x < K;) {
            uint64_t y = x + 1;
            while (y < K && a[y] == a[x]) ++y;
            const uint64_t cnt = y - x;
            ++nu;
            if (cnt >= ci && cnt <= cx) {
                if ((uint64_t)nout >= cap) { free(a); free(t); return -1; }
                keys[nout] = a[x];
                counts[nout] = (uint32_t)cnt;
                ++nout;
            }
            x = y;
        }
        free(a); free(t);
    } else {
        u128* a = (u128*)malloc(sizeof(u128) * (K + 1));
        uint64_t i = 0;
        for (uint64_t c = 0; c < n_chunks; ++c) {
            const uint64_t L = off[c + 1] - off[c];
            if (L < (uint64_t)k) continue;
            for (uint64_t j = 0; j + (uint64_t)k <= L; ++j) {
                const u128 f = kmer_fwd(bases + off[c] + j, k);
                const u128 r = kmer_rc(bases + off[c] + j, k);
                a[i++] = f < r ? f : r;
            }
        }
        qsort(a, K, sizeof(u128), cmp_u128);
        for (uint64_t x = 0; x < K;) {
            uint64_t y = x + 1;
            while (y < K && a[y] == a[x]) ++y;
            const uint64_t cnt = y - x;
            ++nu;
            if (cnt >= ci && cnt <= cx) {
                if ((uint64_t)nout >= cap) { free(a); return -1; }
                keys[2 * nout] = (uint64_t)a[x];
                keys[2 * nout + 1] = (uint64_t)(a[x] >> 64);
                counts[nout] = (uint32_t)cnt;
                ++nout;
            }
            x = y;
        }
        free(a);
    }
    if (n_unique) *n_unique = nu;
    return nout;
}

/* --------------------------------------------------------------- A10 ---- */
/* src/count.hpp:180-216 with Kmer::get_label (Kmer.hpp:659-682): uppercase
 * ACGT label, first base first, '\t', decimal count, '\n'. */
uint64_t orc_tsv_line(const uint64_t* words, int k, uint32_t count, char* out)
{
    static const char B[4] = {'A', 'C', 'G', 'T'};
    for (int i = 0; i < k; ++i) {
        const int bit = 2 * (k - 1 - i);
        const uint64_t w = words[bit / 64];
        out[i] = B[(w >> (bit % 64)) & 3u];
    }
    uint64_t o = (uint64_t)k;
    out[o++] = '\t';
    char tmp[16];
    int n = 0;
    uint32_t v = count;
    do { tmp[n++] = (char)('0' + v % 10u); v /= 10u; } while (v);
    while (n) out[o++] = tmp[--n];
    out[o++] = '\n';
    return o;
}

/* --------------------------------------------------------------- A12 ---- */
/* src/setup.hpp:53-59. */
uint32_t orc_round_pow2(uint32_t n)
{
    if (n <= 1) return 1;
    if ((n & (n - 1)) == 0) return n;
    while (n & (n - 1)) n &= n - 1;
    return n << 1;
}
