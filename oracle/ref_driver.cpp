// ref_driver.cpp — a driver of our own around the reference's Phase-1 code and
// kache-hash table, compiled in place from /root/reference by oracle/Makefile
// into oracle/_ref/ (never copied into this repository).  TEST INFRASTRUCTURE:
// it produces golden vectors and the reference-side CPU timing.
//
// Modes
//   parse FILE                      ACGT chunks (reference SeqSource/helicase)
//   records K M N FILE OUT          per-partition record bytes of
//                                   extract_superkmers_from_actg
//   count K M N T CI CX OUT FILE..  phase 1 with partition_kmers_mem (the
//                                   reference in-memory harness, T threads),
//                                   phase 2 with the reference table through
//                                   the count_partition loop, TSV to OUT
//   hash M STRING                   canonical ntHash of every m-mer (Roller)
//   window K M STRING               MinimizerWindow hash()/min_lmer_pos()
//
// count.hpp is not included: it pulls kff_io.hpp, which is not available
// offline.  The phase-2 loop below restates count_partition (src/count.hpp:
// 71-175) over the reference Kmer_Window / Streaming_Kmer_Hash_Table.
#include "partition_hash.hpp"
#include "kache-hash/Streaming_Kmer_Hash_Table.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

static double now_s()
{
    using C = std::chrono::steady_clock;
    return std::chrono::duration<double>(C::now().time_since_epoch()).count();
}

template <uint16_t k, uint16_t m>
static int run_records(uint32_t n, const char* file, const char* out)
{
    std::vector<std::string> parts(n);
    std::vector<SuperkmerWriter<k, m>> writers(n, SuperkmerWriter<k, m>(4096));
    MinimizerWindow<k, m> win;
    std::vector<uint8_t> kbuf;
    uint64_t kmers = 0, sks = 0;
    const uint64_t mask = n - 1;
    SeqSource src;
    src.process(file, [&](const char* p, size_t len) {
        extract_superkmers_from_actg<k, m>(
            p, len, [mask](uint64_t h) -> size_t { return h & mask; }, win, writers, kmers, sks,
            [&](std::vector<SuperkmerWriter<k, m>>& w, size_t pid) {
                parts[pid].append(w[pid].data(), w[pid].size());
                w[pid].clear();
            },
            kbuf);
    });
    std::ofstream o(out, std::ios::binary);
    o.write((const char*)&n, 4);
    for (auto& s : parts) {
        const uint64_t sz = s.size();
        o.write((const char*)&sz, 8);
        o.write(s.data(), (std::streamsize)sz);
    }
    fprintf(stderr, "kmers: %llu\nsuperkmers: %llu\n", (unsigned long long)kmers, (unsigned long long)sks);
    return 0;
}

template <uint16_t k, uint16_t m>
static int run_count(uint32_t n, uint32_t t, uint32_t ci, uint64_t cx, const char* out,
                     const std::vector<std::string>& files)
{
    Config cfg;
    cfg.input_files = files;
    cfg.k = k;
    cfg.l = m;
    cfg.num_partitions = n;
    cfg.num_threads = t;
    cfg.ci = ci;
    cfg.cx = cx;
    cfg.hide_progress = true;
    std::vector<std::string> bufs(n);
    const double t0 = now_s();
    const PartitionStats st = partition_kmers_mem<k, m>(cfg, bufs);
    const double t1 = now_s();

    using table_t = kache_hash::Streaming_Kmer_Hash_Table<k, false, uint32_t, m>;
    std::atomic<uint32_t> next{0};
    std::vector<std::vector<std::pair<std::string, uint32_t>>> res(t);
    std::atomic<uint64_t> inserted{0};
    auto worker = [&](uint32_t tid) {
        typename table_t::Token token;
        auto inc = [](uint32_t v) { return v + 1; };
        std::string label;
        for (;;) {
            const uint32_t p = next.fetch_add(1);
            if (p >= n) break;
            const size_t per = std::clamp<size_t>(st.kmers / n * 2, 1u << 15, 1u << 22);
            table_t table(per, 1);
            MemoryReader<k, m> rd(bufs[p]);
            uint64_t ins = 0;
            kache_hash::Kmer_Window<k, m> win;
            while (rd.next()) {
                const uint8_t* pk = rd.packed_data();
                const size_t len = rd.size();
                if (len < k) continue;
                win.init_packed_with_min(pk, rd.min_pos());
                table.upsert(win, inc, uint32_t(1), token);
                ++ins;
                const uint8_t* bp = pk + (k >> 2);
                int sh = (int)(6u - 2u * (k & 3u));
                for (size_t i = k; i < len; ++i) {
                    const auto b = static_cast<kache_hash::DNA::Base>((*bp >> sh) & 3u);
                    sh -= 2;
                    if (sh < 0) { sh = 6; ++bp; }
                    win.advance(b);
                    table.upsert(win, inc, uint32_t(1), token);
                    ++ins;
                }
            }
            inserted += ins;
            table.for_each([&](const auto& e) {
                const uint64_t c = e.second;
                if (c < ci || c > cx) return;
                e.first.get_label(label);
                res[tid].emplace_back(label, (uint32_t)c);
            });
        }
    };
    std::vector<std::thread> th;
    for (uint32_t i = 0; i < t; ++i) th.emplace_back(worker, i);
    for (auto& x : th) x.join();
    const double t2 = now_s();
    uint64_t written = 0;
    if (out && std::string(out) != "-") {
        std::ofstream o(out);
        for (auto& v : res)
            for (auto& kv : v) {
                o << kv.first << '\t' << kv.second << '\n';
                ++written;
            }
    } else {
        for (auto& v : res) written += v.size();
    }
    fprintf(stderr, "phase1: %.6f\nphase2: %.6f\nsuperkmers: %llu\nkmers: %llu\ninserted: %llu\nunique_kmers: %llu\nthreads: %u\n",
            t1 - t0, t2 - t1, (unsigned long long)st.superkmers, (unsigned long long)st.kmers,
            (unsigned long long)inserted.load(), (unsigned long long)written, t);
    return 0;
}

template <uint16_t m>
static int run_hash(const std::string& s)
{
    if (s.size() < m) return 0;
    nt_hash::Roller<m> r;
    r.init(s.c_str());
    printf("%llu\n", (unsigned long long)r.canonical());
    for (size_t i = m; i < s.size(); ++i) {
        r.roll(nt_hash::to_2bit(s[i - m]), nt_hash::to_2bit(s[i]));
        printf("%llu\n", (unsigned long long)r.canonical());
    }
    return 0;
}

template <uint16_t k, uint16_t m>
static int run_window(const std::string& s)
{
    if (s.size() < k) return 0;
    MinimizerWindow<k, m> w;
    w.reset(s.c_str());
    printf("%llu %llu\n", (unsigned long long)w.hash(), (unsigned long long)w.min_lmer_pos());
    for (size_t i = k; i < s.size(); ++i) {
        w.advance(s[i]);
        printf("%llu %llu\n", (unsigned long long)w.hash(), (unsigned long long)w.min_lmer_pos());
    }
    return 0;
}

// (k, m) pairs compiled in: the five workloads plus small test shapes.
#define KM_LIST(X) X(31, 21) X(31, 23) X(21, 15) X(63, 21) X(11, 5) X(15, 7) X(5, 3) \
    X(9, 8) X(33, 15) X(25, 11) X(17, 2) X(64, 31) X(32, 16) X(3, 2) X(45, 9)

int main(int argc, char** argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: ref_tuna parse|records|count|hash|window ...\n");
        return 2;
    }
    const std::string mode = argv[1];
    try {
        if (mode == "parse" && argc >= 3) {
            SeqSource src;
            src.process(argv[2], [&](const char* p, size_t len) {
                fwrite(p, 1, len, stdout);
                fputc('\n', stdout);
            });
            return 0;
        }
        if (mode == "records" && argc >= 7) {
            const int k = atoi(argv[2]), m = atoi(argv[3]);
            const uint32_t n = (uint32_t)atoi(argv[4]);
#define X(K, M) if (k == K && m == M) return run_records<K, M>(n, argv[5], argv[6]);
            KM_LIST(X)
#undef X
        }
        if (mode == "count" && argc >= 10) {
            const int k = atoi(argv[2]), m = atoi(argv[3]);
            const uint32_t n = (uint32_t)atoi(argv[4]), t = (uint32_t)atoi(argv[5]);
            const uint32_t ci = (uint32_t)strtoul(argv[6], nullptr, 10);
            const uint64_t cx = strtoull(argv[7], nullptr, 10);
            std::vector<std::string> files(argv + 9, argv + argc);
#define X(K, M) if (k == K && m == M) return run_count<K, M>(n, t, ci, cx, argv[8], files);
            KM_LIST(X)
#undef X
        }
        if (mode == "hash" && argc >= 4) {
            const int m = atoi(argv[2]);
#define X(K, M) if (m == M) return run_hash<M>(argv[3]);
            KM_LIST(X)
#undef X
        }
        if (mode == "window" && argc >= 5) {
            const int k = atoi(argv[2]), m = atoi(argv[3]);
#define X(K, M) if (k == K && m == M) return run_window<K, M>(argv[4]);
            KM_LIST(X)
#undef X
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "ref_tuna: error: %s\n", e.what());
        return 1;
    }
    fprintf(stderr, "ref_tuna: unsupported mode or (k, m)\n");
    return 2;
}
