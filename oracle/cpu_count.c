/*
 * cpu_count.c — multithreaded CPU restatement of tuna's two-phase counting,
 * timed as the CPU baseline of the benchmark when the reference itself is not
 * on the machine.  TEST INFRASTRUCTURE ONLY: the product never links or calls
 * this file; bench.py's cpu_baseline leg and tests/ use it.
 *
 * The phases follow the reference pipeline (src/pipeline.hpp:111-178):
 *   Phase 1 (src/partition_hash.hpp:48-101, 543-565): threads take contiguous
 *     ranges of chunks, cut superkmers with the oracle's rolling canonical
 *     ntHash and window minimum (orc_nthash_roll, orc_window_stream, the
 *     superkmer rule of orc_superkmers) and route each to partition
 *     hash & (n-1).  A record here is (first base, length), pointing into the
 *     resident bases instead of a packed copy.
 *   Phase 2 (src/count.hpp:71-175, 875-970): threads take partitions from a
 *     shared counter; each counts its partition's k-mers into an
 *     open-addressed table (one per partition, sized from its k-mer count, as
 *     src/count.hpp:905-916 sizes tables), then applies ci <= count <= cx.
 * Pinned by tests/test_oracle.py against orc_count (the sorted multiset of
 * every canonical k-mer), which is independent of partitions and threads.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

typedef struct {
    uint64_t* v;  /* start << 8 | len */
    uint64_t n, cap;
} RecList;

typedef struct {
    const uint8_t* bases;
    const uint64_t* off;
    uint64_t c0, c1;
    int k, m;
    uint32_t n_parts;
    RecList* parts; /* n_parts lists of this thread */
    uint64_t kmers, superkmers;
    int fail;
} P1;

static int push(RecList* l, uint64_t x)
{
    if (l->n == l->cap) {
        const uint64_t c = l->cap ? 2 * l->cap : 64;
        uint64_t* v = (uint64_t*)realloc(l->v, c * sizeof(uint64_t));
        if (!v) return -1;
        l->v = v;
        l->cap = c;
    }
    l->v[l->n++] = x;
    return 0;
}

static void* phase1(void* arg)
{
    P1* a = (P1*)arg;
    const int k = a->k, m = a->m, W = k - m + 1;
    const uint64_t hdr_max = (2u * (unsigned)k - (unsigned)m <= 255u) ? 255u : 65535u;
    uint64_t cap = 0;
    uint64_t *h = NULL, *H = NULL;
    uint32_t* P = NULL;
    for (uint64_t c = a->c0; c < a->c1; ++c) {
        const uint64_t L = a->off[c + 1] - a->off[c];
        if (L < (uint64_t)k) continue;
        if (L + 1 > cap) {
            cap = L + 1;
            free(h); free(H); free(P);
            h = (uint64_t*)malloc(cap * sizeof(uint64_t));
            H = (uint64_t*)malloc(cap * sizeof(uint64_t));
            P = (uint32_t*)malloc(cap * sizeof(uint32_t));
            if (!h || !H || !P) { a->fail = 1; break; }
        }
        const uint8_t* s = a->bases + a->off[c];
        orc_nthash_roll(s, L, m, h);
        orc_window_stream(h, L - (uint64_t)m + 1u, W, H, P);
        const uint64_t K = L - (uint64_t)k + 1u;
        uint64_t t0 = 0;
        for (uint64_t t = 1; t <= K; ++t) {
            if (!(t == K || H[t] != H[t0] || t + (uint64_t)k - 1u - t0 >= hdr_max)) continue;
            const uint64_t len = t - t0 + (uint64_t)k - 1u;
            if (push(&a->parts[H[t0] & (a->n_parts - 1u)], ((a->off[c] + t0) << 8) | len)) {
                a->fail = 1;
                break;
            }
            a->kmers += t - t0;
            a->superkmers += 1;
            t0 = t;
        }
    }
    free(h); free(H); free(P);
    return NULL;
}

typedef struct {
    const uint8_t* bases;
    int k, n_threads;
    uint32_t n_parts, ci;
    uint64_t cx;
    P1* p1;
    atomic_uint_fast64_t* next_part;
    atomic_uint_fast64_t* out_cursor;
    uint64_t* keys;
    uint32_t* counts;
    uint64_t cap;
    uint64_t unique, out;
    int fail;
} P2;

static inline uint64_t mix(uint64_t x)
{
    x ^= x >> 31;
    x *= 0x7fb5d329728ea185ull;
    x ^= x >> 27;
    x *= 0x81dadef4bc2dd44dull;
    x ^= x >> 33;
    return x;
}

static void* phase2(void* arg)
{
    P2* a = (P2*)arg;
    const int k = a->k;
    const int wide = k > 32;
    const uint64_t mask = k >= 32 ? ~0ull : ((1ull << (2 * k)) - 1u);
    const u128 mask2 = k >= 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1u);
    uint64_t tcap = 0;
    uint64_t* tk = NULL;   /* k <= 32: key words, empty = ~0 */
    u128* tk2 = NULL;      /* k > 32: keys, empty = all ones */
    uint32_t* tc = NULL;
    for (;;) {
        const uint64_t p = atomic_fetch_add(a->next_part, 1);
        if (p >= a->n_parts) break;
        uint64_t km = 0;
        for (int t = 0; t < a->n_threads; ++t) {
            const RecList* l = &a->p1[t].parts[p];
            for (uint64_t i = 0; i < l->n; ++i) km += (l->v[i] & 0xFF) - (uint64_t)k + 1u;
        }
        uint64_t want = 1024;
        while (want < 2 * km) want <<= 1;
        if (want > tcap) {
            free(tk); free(tk2); free(tc);
            tk = NULL; tk2 = NULL;
            tcap = want;
            if (wide) tk2 = (u128*)malloc(tcap * sizeof(u128));
            else tk = (uint64_t*)malloc(tcap * sizeof(uint64_t));
            tc = (uint32_t*)malloc(tcap * sizeof(uint32_t));
            if ((wide ? !tk2 : !tk) || !tc) { a->fail = 1; break; }
        }
        const uint64_t tm = want - 1;
        if (wide) memset(tk2, 0xFF, want * sizeof(u128));
        else memset(tk, 0xFF, want * sizeof(uint64_t));
        memset(tc, 0, want * sizeof(uint32_t));
        for (int t = 0; t < a->n_threads; ++t) {
            const RecList* l = &a->p1[t].parts[p];
            for (uint64_t i = 0; i < l->n; ++i) {
                const uint8_t* s = a->bases + (l->v[i] >> 8);
                const uint32_t len = (uint32_t)(l->v[i] & 0xFF);
                if (!wide) {
                    uint64_t f = 0, r = 0;
                    for (uint32_t j = 0; j < len; ++j) {
                        const uint64_t b = orc_kache_code(s[j]);
                        f = ((f << 2) | b) & mask;
                        r = (r >> 2) | ((3u - b) << (2 * (k - 1)));
                        if (j + 1 < (uint32_t)k) continue;
                        const uint64_t x = f < r ? f : r;
                        uint64_t q = mix(x) & tm;
                        while (tk[q] != x && tk[q] != ~0ull) q = (q + 1) & tm;
                        tk[q] = x;
                        tc[q] += 1;
                    }
                } else {
                    u128 f = 0, r = 0;
                    for (uint32_t j = 0; j < len; ++j) {
                        const u128 b = orc_kache_code(s[j]);
                        f = ((f << 2) | b) & mask2;
                        r = (r >> 2) | ((u128)(3u - (uint32_t)b) << (2 * (k - 1)));
                        if (j + 1 < (uint32_t)k) continue;
                        const u128 x = f < r ? f : r;
                        uint64_t q = mix((uint64_t)x ^ mix((uint64_t)(x >> 64))) & tm;
                        while (tk2[q] != x && tk2[q] != ~(u128)0) q = (q + 1) & tm;
                        tk2[q] = x;
                        tc[q] += 1;
                    }
                }
            }
        }
        uint64_t u = 0, o = 0;
        for (uint64_t q = 0; q < want; ++q) {
            if (!tc[q]) continue;
            ++u;
            if (tc[q] >= a->ci && (uint64_t)tc[q] <= a->cx) ++o;
        }
        a->unique += u;
        a->out += o;
        if (a->keys && o) {
            uint64_t at = atomic_fetch_add(a->out_cursor, o);
            if (at + o > a->cap) { a->fail = 1; continue; }
            for (uint64_t q = 0; q < want; ++q) {
                if (!tc[q] || tc[q] < a->ci || (uint64_t)tc[q] > a->cx) continue;
                if (wide) {
                    a->keys[2 * at] = (uint64_t)tk2[q];
                    a->keys[2 * at + 1] = (uint64_t)(tk2[q] >> 64);
                } else {
                    a->keys[at] = tk[q];
                }
                a->counts[at] = tc[q];
                ++at;
            }
        }
    }
    free(tk); free(tk2); free(tc);
    return NULL;
}

int64_t orc_count_parallel(const uint8_t* bases, const uint64_t* off, uint64_t n_chunks, int k, int m,
                           uint32_t n_parts, uint32_t ci, uint64_t cx, int threads, uint64_t* keys,
                           uint32_t* counts, uint64_t cap, uint64_t* stats, double* phase_s)
{
    if (k < 2 || k > 64 || m < 1 || m >= k || threads < 1 || n_parts == 0 || (n_parts & (n_parts - 1)))
        return -2;
    struct timespec t0, t1, t2;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    P1* p1 = (P1*)calloc((size_t)threads, sizeof(P1));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    const uint64_t total = off[n_chunks] - off[0];
    uint64_t c = 0;
    for (int t = 0; t < threads; ++t) {
        /* contiguous chunk ranges of about equal bases */
        const uint64_t until = off[0] + total / (uint64_t)threads * (uint64_t)(t + 1);
        const uint64_t c0 = c;
        while (c < n_chunks && (t == threads - 1 || off[c + 1] <= until)) ++c;
        p1[t] = (P1){bases, off, c0, c, k, m, n_parts, (RecList*)calloc(n_parts, sizeof(RecList)), 0, 0, 0};
        pthread_create(&th[t], NULL, phase1, &p1[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    atomic_uint_fast64_t next = 0, cursor = 0;
    P2* p2 = (P2*)calloc((size_t)threads, sizeof(P2));
    for (int t = 0; t < threads; ++t) {
        p2[t] = (P2){bases, k, threads, n_parts, ci, cx, p1, &next, &cursor, keys, counts, cap, 0, 0, 0};
        pthread_create(&th[t], NULL, phase2, &p2[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t2);
    uint64_t kmers = 0, sk = 0, uniq = 0, out = 0;
    int fail = 0;
    for (int t = 0; t < threads; ++t) {
        kmers += p1[t].kmers;
        sk += p1[t].superkmers;
        uniq += p2[t].unique;
        out += p2[t].out;
        fail |= p1[t].fail | p2[t].fail;
        for (uint32_t p = 0; p < n_parts; ++p) free(p1[t].parts[p].v);
        free(p1[t].parts);
    }
    free(p1); free(p2); free(th);
    if (stats) {
        stats[0] = kmers;
        stats[1] = sk;
        stats[2] = uniq;
    }
    if (phase_s) {
        phase_s[0] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        phase_s[1] = (double)(t2.tv_sec - t1.tv_sec) + 1e-9 * (double)(t2.tv_nsec - t1.tv_nsec);
    }
    return fail ? -1 : (int64_t)out;
}
