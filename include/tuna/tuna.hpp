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
#include <ostream>
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

// DNA bases (include/kache-hash/DNA.hpp).
// 2-bit codes A0 C1 G2 T3 (the complement of b is 3 - b); N and E are the
// reference's markers for an ambiguous and an empty base.  Extended_Base is
// the de Bruijn edge encoding (E0 A1 C2 G3 T4 N7).
namespace DNA {
enum Base : uint8_t { A = 0, C = 1, G = 2, T = 3, N = 4, E = 5 };
enum class Extended_Base : uint8_t { E = 0, A = 1, C = 2, G = 3, T = 4, OP_non_branch = 5, OP_branching = 6, N = 7 };
}  // namespace DNA

// K-mers.
// The reference exposes kache_hash::Kmer<k> as tuna::Kmer<k>
// (include/tuna/tuna.hpp:54-55; include/kache-hash/Kmer.hpp:29-221).  This
// class has its binary layout and its public members:
//   * ceil(k/32) 64-bit words, 2 bits per base (A0 C1 G2 T3); a k-mer
//     n_{k-1} ... n_1 n_0 (n_{k-1} first in text order) keeps n_i in bits
//     2(i%32)+1 : 2(i%32) of word i/32, and the bits above n_{k-1} are zero;
//   * comparisons order the words from the highest down, which is the
//     lexicographic A < C < G < T order of the labels;
//   * the canonical form is the smaller of a k-mer and its reverse complement.
// to_u64() hashes the words with this library's own 64-bit mix (the reference
// uses XXH3 there); the value is only a hash-table key, so no count, label or
// canonical form depends on it.
template <uint16_t k>
class Kmer {
    static_assert(k >= 1 && k <= 256, "tuna: k must be in [1, 256]");
    friend class Kmer<k - 1>;

public:
    static constexpr uint16_t NUM_INTS = (k + 31) / 32;

private:
    static constexpr uint16_t TOP_SHIFT = 2 * ((k - 1) % 32);  // bit of n_{k-1} in the top word
    static constexpr uint64_t TOP_KEEP = (k % 32) ? (~0ull >> (64 - 2 * (k % 32))) : ~0ull;

    // n_i -> n_{i+1} for every base (n_{k-1} leaves, n_0 becomes A)
    void shift_up()
    {
        for (int w = NUM_INTS - 1; w > 0; --w) kmer_data[w] = (kmer_data[w] << 2) | (kmer_data[w - 1] >> 62);
        kmer_data[0] <<= 2;
        kmer_data[NUM_INTS - 1] &= TOP_KEEP;
    }
    // n_i -> n_{i-1} for every base (n_0 leaves, n_{k-1} becomes A)
    void shift_down()
    {
        for (uint16_t w = 0; w + 1 < NUM_INTS; ++w) kmer_data[w] = (kmer_data[w] >> 2) | (kmer_data[w + 1] << 62);
        kmer_data[NUM_INTS - 1] >>= 2;
    }
    static uint8_t code_of(char c) { return static_cast<uint8_t>(((c >> 2) ^ (c >> 1)) & 3); }
    static DNA::Base edge_base(DNA::Extended_Base e)
    {
        const uint8_t v = static_cast<uint8_t>(e);
        return (v >= 1 && v <= 4) ? DNA::Base(v - 1) : (v == 0 ? DNA::E : DNA::N);
    }

public:
    // "AA...A"
    constexpr Kmer() = default;
    // The k characters label[idx .. idx + k) (A/C/G/T, either case).
    Kmer(const char* label, std::size_t idx) : Kmer(label + idx) {}
    Kmer(const char* label)
    {
        for (uint16_t j = 0; j < k; ++j) {
            const uint16_t i = k - 1 - j;  // label[j] is n_i
            kmer_data[i / 32] |= static_cast<uint64_t>(code_of(label[j])) << (2 * (i % 32));
        }
    }
    Kmer(const std::string& label) : Kmer(label.c_str()) {}
    Kmer(const std::string& label, std::size_t idx) : Kmer(label.c_str() + idx) {}
    // The k-mer whose encoding is int_val (k <= 32).
    Kmer(uint64_t int_val)
    {
        static_assert(k <= 32, "tuna: Kmer<k>(uint64_t) needs k <= 32");
        kmer_data[0] = int_val;
    }
    // The k-mer of NUM_INTS words in this layout (as the C ABI returns keys).
    static Kmer from_words(const uint64_t* words)
    {
        Kmer x;
        for (uint16_t i = 0; i < NUM_INTS; ++i) x.kmer_data[i] = words[i];
        return x;
    }

    constexpr static uint16_t num_words() { return NUM_INTS; }

    // A 64-bit hash of the k-mer (see the class comment).
    uint64_t to_u64(uint64_t seed = 0) const
    {
        uint64_t h = seed ^ (0x9E3779B97F4A7C15ull * (k + 1));
        for (uint16_t i = 0; i < NUM_INTS; ++i) {
            h ^= kmer_data[i];
            h = (h ^ (h >> 32)) * 0xD6E8FEB86659FD93ull;
            h = (h ^ (h >> 32)) * 0xD6E8FEB86659FD93ull;
            h ^= h >> 32;
        }
        return h;
    }

    // The prefix (first k bases) and suffix (last k bases) of a (k+1)-mer.
    void from_prefix(const Kmer<k + 1>& e)
    {
        Kmer<k + 1> t = e;
        t.shift_down();
        for (uint16_t i = 0; i < NUM_INTS; ++i) kmer_data[i] = t.kmer_data[i];
    }
    void from_suffix(const Kmer<k + 1>& e)
    {
        for (uint16_t i = 0; i < NUM_INTS; ++i) kmer_data[i] = e.kmer_data[i];
        kmer_data[NUM_INTS - 1] &= TOP_KEEP;
    }

    Kmer reverse_complement() const
    {
        Kmer r;
        for (uint16_t i = 0; i < k; ++i) {
            const uint16_t j = k - 1 - i;  // n_i complemented becomes base j of the result
            r.kmer_data[j / 32] |= static_cast<uint64_t>(3u - base_at(i)) << (2 * (j % 32));
        }
        return r;
    }
    void as_reverse_complement(const Kmer& other) { *this = other.reverse_complement(); }

    bool operator<(const Kmer& rhs) const
    {
        for (int i = NUM_INTS - 1; i >= 0; --i)
            if (kmer_data[i] != rhs.kmer_data[i]) return kmer_data[i] < rhs.kmer_data[i];
        return false;
    }
    bool operator>(const Kmer& rhs) const { return rhs < *this; }
    bool operator==(const Kmer& rhs) const
    {
        for (uint16_t i = 0; i < NUM_INTS; ++i)
            if (kmer_data[i] != rhs.kmer_data[i]) return false;
        return true;
    }
    bool operator!=(const Kmer& rhs) const { return !(*this == rhs); }

    // n_{k-1} (the first base of the label), n_0 (the last), n_idx.
    DNA::Base front() const { return DNA::Base((kmer_data[NUM_INTS - 1] >> TOP_SHIFT) & 3u); }
    DNA::Base back() const { return DNA::Base(kmer_data[0] & 3u); }
    DNA::Base base_at(uint16_t idx) const { return DNA::Base((kmer_data[idx / 32] >> (2 * (idx % 32))) & 3u); }

    const uint64_t* data() const { return kmer_data; }
    uint64_t* data() { return kmer_data; }

    bool in_forward(const Kmer& kmer_hat) const { return *this == kmer_hat; }

    // Drop the first base, append `base`; rev_compl (this k-mer's reverse
    // complement) follows.
    void roll_to_next_kmer(DNA::Base base, Kmer& rev_compl)
    {
        roll_forward(base);
        rev_compl.shift_down();
        rev_compl.kmer_data[NUM_INTS - 1] |= static_cast<uint64_t>(3u - base) << TOP_SHIFT;
    }
    void roll_to_next_kmer(char next_base, Kmer& rev_compl) { roll_to_next_kmer(DNA::Base(code_of(next_base)), rev_compl); }
    void roll_to_next_kmer(DNA::Extended_Base edge, Kmer& rev_compl) { roll_to_next_kmer(edge_base(edge), rev_compl); }
    // Drop the last base, prepend `base`; rev_compl follows.
    void roll_to_prev_kmer(DNA::Base base, Kmer& rev_compl)
    {
        shift_down();
        kmer_data[NUM_INTS - 1] |= static_cast<uint64_t>(base) << TOP_SHIFT;
        rev_compl.shift_up();
        rev_compl.kmer_data[0] |= 3u - base;
    }
    void roll_forward(DNA::Base base)
    {
        shift_up();
        kmer_data[0] |= base;
    }
    void roll_forward(DNA::Extended_Base edge) { roll_forward(edge_base(edge)); }
    void roll_backward(DNA::Extended_Base edge)
    {
        shift_down();
        kmer_data[NUM_INTS - 1] |= static_cast<uint64_t>(edge_base(edge)) << TOP_SHIFT;
    }

    Kmer canonical(const Kmer& rev_compl) const { return rev_compl < *this ? rev_compl : *this; }
    Kmer canonical() const { return canonical(reverse_complement()); }
    static const Kmer* canonical(const Kmer& kmer, const Kmer& rev_compl)
    {
        return kmer < rev_compl ? &kmer : &rev_compl;
    }

    // The l-mer starting at label index idx, as an integer (k <= 32).
    template <uint16_t l>
    uint64_t lmer_at(std::size_t idx) const
    {
        static_assert(k <= 32 && l >= 1 && l <= k, "tuna: lmer_at needs l <= k <= 32");
        const uint64_t mask = l == 32 ? ~0ull : ((1ull << (2 * l)) - 1);
        return (kmer_data[0] >> (2 * (k - idx - l))) & mask;
    }

    // The label, first base first, in upper case.
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
    operator std::string() const { return string_label(); }
    friend std::ostream& operator<<(std::ostream& out, const Kmer& x) { return out << x.string_label(); }

    // 2-bit MSB-first bytes, first base first (the KFF sequence payload).
    void write_packed_2bit_msb(uint8_t* dst) const
    {
        for (size_t i = 0; i < (k + 3) / 4u; ++i) dst[i] = 0;
        for (uint16_t j = 0; j < k; ++j)
            dst[j >> 2] |= static_cast<uint8_t>(base_at(k - 1 - j) << (6u - 2u * (j & 3u)));
    }

    // The encoding as an integer (k <= 32).
    uint64_t as_int() const
    {
        static_assert(k <= 32, "tuna: as_int needs k <= 32");
        return kmer_data[0];
    }

private:
    uint64_t kmer_data[NUM_INTS] = {};
};

// Hash functor for unordered containers keyed by Kmer<k>
// (include/kache-hash/Kmer.hpp:225-230).
template <uint16_t k>
class Kmer_Hasher {
public:
    uint64_t operator()(const Kmer<k>& key, uint64_t seed = 0) const { return key.to_u64(seed); }
};

// Options of a call (include/tuna/tuna.hpp:59-66).
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

// The process's device context.  A call holds its lock from the count until
// its result has been copied out, so concurrent tuna::count calls from
// several threads run one after the other and never see each other's
// results (the reference gives every call its own work directory,
// include/tuna/tuna.hpp:121-133).
struct Device {
    std::mutex lock;
    std::unique_ptr<tg_ctx, void (*)(tg_ctx*)> ctx{nullptr, tg_ctx_destroy};
};

inline Device& device()
{
    static Device d;
    return d;
}

template <uint16_t k, uint16_t m>
inline void run(const std::vector<std::string>& files, const Options& opts,
                std::vector<uint64_t>& keys, std::vector<uint32_t>& counts)
{
    static_assert(k >= 2 && k <= 256, "tuna: k must be in [2, 256]");
    static_assert(m >= 1 && m < k && m <= 32, "tuna: m must be in [1, min(k-1, 32)]");
    if (files.empty()) throw std::invalid_argument("tuna::count: no input files");
    uint32_t n = opts.partitions;
    if (n == 0) {
        std::vector<const char*> fp;
        for (const auto& f : files) fp.push_back(f.c_str());
        n = tg_auto_partitions(fp.data(), (uint32_t)fp.size(), k, opts.threads ? opts.threads : 1);
    }
    tg_params p{k, m, n, opts.min_count, opts.max_count};
    Device& d = device();
    std::lock_guard<std::mutex> hold(d.lock);
    if (!d.ctx) {
        tg_ctx* made = nullptr;
        check(tg_ctx_create(0, &made));
        d.ctx.reset(made);
    }
    tg_ctx* c = d.ctx.get();
    std::vector<const char*> paths;
    for (const auto& f : files) paths.push_back(f.c_str());
    // the files are read straight to the device, the upload overlapping the read
    check(tg_count_files(c, &p, paths.data(), (uint32_t)paths.size(), opts.threads ? opts.threads : 1, nullptr));
    const uint64_t nout = tg_result_size(c);
    keys.resize(nout * ((k + 31) / 32));
    counts.resize(nout);
    check(tg_result_copy(c, keys.data(), counts.data(), nout));
}

}  // namespace detail

// count: calls cb once per canonical k-mer whose count c satisfies
// min_count <= c <= max_count.  The callback's signature picks the k-mer form:
// a callable taking (const Kmer<k>&, uint32_t) receives the packed k-mer, any
// other is called with the upper-case label as a std::string_view.  The view
// and the reference are valid during the call only.  An empty file list throws
// std::invalid_argument; input and device failures throw std::runtime_error.
template <uint16_t k, uint16_t m = 21, typename Callback>
void count(const std::vector<std::string>& files, Callback&& cb, Options opts = {})
{
    std::vector<uint64_t> keys;
    std::vector<uint32_t> counts;
    detail::run<k, m>(files, opts, keys, counts);
    constexpr size_t nw = (k + 31) / 32;
    std::string label;
    for (size_t i = 0; i < counts.size(); ++i) {
        const Kmer<k> km = Kmer<k>::from_words(keys.data() + i * nw);
        if constexpr (std::is_invocable_v<Callback, const Kmer<k>&, uint32_t>) {
            cb(km, counts[i]);
        } else {
            km.get_label(label);
            cb(std::string_view(label), counts[i]);
        }
    }
}

// count_to: the labels and counts collected into a container (an
// unordered_map<string, uint32_t> unless another is named).
template <uint16_t k, uint16_t m = 21, typename Container = std::unordered_map<std::string, uint32_t>>
Container count_to(const std::vector<std::string>& files, Options opts = {})
{
    Container result;
    count<k, m>(files, [&](std::string_view kmer, uint32_t cnt) { result.emplace(std::string(kmer), cnt); },
                opts);
    return result;
}

// count_to_raw: the packed k-mers and counts as a vector of pairs.
template <uint16_t k, uint16_t m = 21>
std::vector<std::pair<Kmer<k>, uint32_t>> count_to_raw(const std::vector<std::string>& files, Options opts = {})
{
    std::vector<std::pair<Kmer<k>, uint32_t>> result;
    count<k, m>(files, [&](const Kmer<k>& kmer, uint32_t cnt) { result.emplace_back(kmer, cnt); }, opts);
    return result;
}

}  // namespace tuna
