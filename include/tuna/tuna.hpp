#pragma once
// tuna — public C++ API of the MI355X (gfx950) engine.
//
// Same templates, option names, callback forms and exceptions as the
// reference library (include/tuna/tuna.hpp:46-195 of vicLeva/tuna v1.4.0):
//
//   auto kmers = tuna::count_to<31>({"genome.fa"});              // unordered_map
//   tuna::count<31>({"reads.fq"}, [](std::string_view s, uint32_t c) { ... });
//   auto raw = tuna::count_to_raw<31>({"reads.fq"});             // (Kmer<31>, count)
//
// Counting runs on the GPU through the C ABI of include/tuna_gpu.h (link
// libtuna_gpu.so).  Callbacks are invoked from the calling thread only, a
// legal subset of the reference's "may be called concurrently" contract
// (include/tuna/tuna.hpp:82-86).  No temporary directory is used: partition
// buffers live in device memory.

#include "../tuna_gpu.h"

#include <cstdint>
#include <cstdlib>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

namespace tuna {

// ─── K-mer type ───────────────────────────────────────────────────────────────
// Binary layout of the reference Kmer<k> (include/kache-hash/Kmer.hpp:44-48):
// ceil(k/32) 64-bit words, 2 bits per base (A0 C1 G2 T3), word 0 holds the
// last 32 bases with the last base in bits 1:0.  Values are canonical.
template <uint16_t k>
class Kmer {
    static_assert(k >= 1 && k <= 256, "tuna: k must be in [1, 256]");

public:
    static constexpr uint16_t NUM_INTS = (k + 31) / 32;

    Kmer() = default;
    explicit Kmer(const uint64_t* words)
    {
        for (uint16_t i = 0; i < NUM_INTS; ++i) kmer_data[i] = words[i];
    }

    // Base at index i counted from the last base (0 = last), kache code.
    uint8_t base_at(uint16_t i) const { return (kmer_data[i / 32] >> (2 * (i % 32))) & 3u; }

    template <typename T_container_>
    void get_label(T_container_& label) const
    {
        static constexpr char B[4] = {'A', 'C', 'G', 'T'};
        label.resize(k);
        for (uint16_t i = 0; i < k; ++i) label[k - 1 - i] = B[base_at(i)];
    }
    std::string string_label() const
    {
        std::string s;
        get_label(s);
        return s;
    }
    explicit operator std::string() const { return string_label(); }

    // 2-bit MSB-first bytes, first base first (Kmer.hpp:684-693)
    void write_packed_2bit_msb(uint8_t* dst) const
    {
        constexpr size_t nb = (k + 3) / 4;
        for (size_t i = 0; i < nb; ++i) dst[i] = 0;
        for (uint16_t i = 0; i < k; ++i)
            dst[i >> 2] |= static_cast<uint8_t>(base_at(k - 1 - i) << (6u - 2u * (i & 3u)));
    }

    bool operator<(const Kmer& rhs) const
    {
        for (int i = NUM_INTS - 1; i >= 0; --i)
            if (kmer_data[i] != rhs.kmer_data[i]) return kmer_data[i] < rhs.kmer_data[i];
        return false;
    }
    bool operator==(const Kmer& rhs) const
    {
        for (uint16_t i = 0; i < NUM_INTS; ++i)
            if (kmer_data[i] != rhs.kmer_data[i]) return false;
        return true;
    }
    const uint64_t* data() const { return kmer_data; }

private:
    uint64_t kmer_data[NUM_INTS] = {};
};

// ─── Options (include/tuna/tuna.hpp:59-66) ───────────────────────────────────
struct Options {
    uint32_t threads = 1;     // host threads (input parsing)
    uint32_t min_count = 1;   // only yield k-mers with count >= min_count
    uint32_t max_count = std::numeric_limits<uint32_t>::max();  // ...and <= max_count
    uint32_t partitions = 0;  // 0 = auto-tune (recommended)
    uint32_t ram_gb = 0;      // accepted for compatibility; device memory is managed
    std::string work_dir;     // accepted for compatibility; no temporary files are made
};

namespace detail {

inline void check(int rc)
{
    if (rc != TG_OK) throw std::runtime_error(tg_last_error());
}

// One device context per process, created on first use.
inline tg_ctx* context()
{
    static std::mutex mu;
    static std::unique_ptr<tg_ctx, void (*)(tg_ctx*)> ctx(nullptr, tg_ctx_destroy);
    std::lock_guard<std::mutex> g(mu);
    if (!ctx) {
        tg_ctx* c = nullptr;
        check(tg_ctx_create(0, &c));
        ctx.reset(c);
    }
    return ctx.get();
}

struct Chunks {
    std::vector<uint8_t> bases;
    std::vector<uint64_t> off{0};
};

inline void read_inputs(const std::vector<std::string>& files, Chunks& out)
{
    for (const auto& f : files) {
        uint8_t* data = nullptr;
        uint64_t len = 0;
        check(tg_read_file(f.c_str(), &data, &len));
        std::unique_ptr<uint8_t, void (*)(void*)> hold(data, tg_free);
        std::vector<uint8_t> b(len);
        std::vector<uint64_t> o(len / 2 + 2);
        const int64_t n = tg_parse_buffer(data, len, b.data(), o.data(), o.size());
        if (n < 0) throw std::runtime_error("cannot parse input: " + f);
        const uint64_t base = out.bases.size();
        out.bases.insert(out.bases.end(), b.begin(), b.begin() + (std::ptrdiff_t)o[n]);
        for (int64_t i = 1; i <= n; ++i) out.off.push_back(base + o[i]);
    }
}

template <uint16_t k, uint16_t m>
inline void run(const std::vector<std::string>& files, const Options& opts,
                std::vector<uint64_t>& keys, std::vector<uint32_t>& counts)
{
    static_assert(k >= 2 && k <= 64, "tuna (gfx950 engine): k must be in [2, 64]");
    static_assert(m >= 1 && m < k && m <= 32, "tuna: m must be in [1, min(k-1, 32)]");
    if (files.empty()) throw std::invalid_argument("tuna::count: no input files");
    Chunks ch;
    read_inputs(files, ch);
    uint32_t n = opts.partitions;
    if (n == 0) {
        std::vector<const char*> fp;
        for (const auto& f : files) fp.push_back(f.c_str());
        n = tg_auto_partitions(fp.data(), (uint32_t)fp.size(), k, opts.threads ? opts.threads : 1);
    }
    tg_params p{k, m, n, opts.min_count, opts.max_count};
    tg_ctx* c = context();
    check(tg_count_host(c, &p, ch.bases.data(), ch.bases.size(), ch.off.data(), ch.off.size() - 1,
                        nullptr));
    const uint64_t nout = tg_result_size(c);
    keys.resize(nout * ((k + 31) / 32));
    counts.resize(nout);
    check(tg_result_copy(c, keys.data(), counts.data(), nout));
}

}  // namespace detail

// ─── count() ──────────────────────────────────────────────────────────────────
// cb(kmer, count) for each canonical k-mer with min_count <= count <= max_count.
//   ASCII  (default): cb(std::string_view kmer, uint32_t count)
//   Binary           : cb(const tuna::Kmer<k>& kmer, uint32_t count)
// Throws std::invalid_argument if files is empty, std::runtime_error on I/O or
// device errors.
template <uint16_t k, uint16_t m = 21, typename Callback>
void count(const std::vector<std::string>& files, Callback&& cb, Options opts = {})
{
    std::vector<uint64_t> keys;
    std::vector<uint32_t> counts;
    detail::run<k, m>(files, opts, keys, counts);
    constexpr size_t nw = (k + 31) / 32;
    std::string label;
    for (size_t i = 0; i < counts.size(); ++i) {
        const Kmer<k> km(keys.data() + i * nw);
        if constexpr (std::is_invocable_v<Callback, const Kmer<k>&, uint32_t>) {
            cb(km, counts[i]);
        } else {
            km.get_label(label);
            cb(std::string_view(label), counts[i]);
        }
    }
}

// ─── count_to() ───────────────────────────────────────────────────────────────
template <uint16_t k, uint16_t m = 21, typename Container = std::unordered_map<std::string, uint32_t>>
Container count_to(const std::vector<std::string>& files, Options opts = {})
{
    Container result;
    count<k, m>(files, [&](std::string_view kmer, uint32_t cnt) { result.emplace(std::string(kmer), cnt); },
                opts);
    return result;
}

// ─── count_to_raw() ───────────────────────────────────────────────────────────
template <uint16_t k, uint16_t m = 21>
std::vector<std::pair<Kmer<k>, uint32_t>> count_to_raw(const std::vector<std::string>& files, Options opts = {})
{
    std::vector<std::pair<Kmer<k>, uint32_t>> result;
    count<k, m>(files, [&](const Kmer<k>& kmer, uint32_t cnt) { result.emplace_back(kmer, cnt); }, opts);
    return result;
}

}  // namespace tuna
