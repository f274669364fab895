// tuna_cli.cpp — the `tuna` command line over the gfx950 engine.
//
// Flags, messages, exit codes and structured stderr lines follow the
// reference CLI (src/CLI.cpp:10-237, src/main.cpp:23-98,
// src/pipeline.hpp:64-263).  Phase 1 (parse + superkmers + routing) and
// Phase 2 (counting) run on the GPU; output is TSV or KFF.
#include "../../include/tuna_gpu.h"

#include <cctype>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#ifndef TUNA_VERSION
#define TUNA_VERSION "1.4.0"
#endif

namespace {

struct Config {  // src/Config.hpp:10-26
    std::vector<std::string> input_files;
    std::string output_file;
    std::string work_dir;
    uint16_t k = 31;
    uint16_t l = 21;
    uint32_t num_partitions = 0;
    uint32_t num_threads = 1;
    uint32_t ci = 1;
    uint64_t cx = std::numeric_limits<uint64_t>::max();
    bool hide_progress = false;
    bool keep_tmp = false;
    bool partition_only = false;
    bool debug_stats = false;
    bool output_kff = false;
    uint64_t ram_budget_bytes = 0;
};

// ---- command line ----------------------------------------------------------
// The option names, their value ranges, every error message and the exit
// codes are the reference CLI's interface (src/CLI.cpp:68-237): scripts rely
// on them.  Here they are data: one row per option, walked by a single loop
// that classifies each token and applies the row; the help text is printed
// from the same rows.

enum class Takes { flag, number, path, help, version };

struct OptionRow {
    const char* name;
    const char* alt;      // second spelling, or nullptr
    Takes takes;
    uint64_t limit;       // largest accepted number (Takes::number)
    void (*apply)(Config&, uint64_t number, const char* text);
    const char* help;     // help lines for this option
};

constexpr uint64_t kMax16 = 0xFFFFull, kMax32 = 0xFFFFFFFFull, kMax64 = ~0ull;

const OptionRow kRows[] = {
    {"-k", nullptr, Takes::number, kMax16, [](Config& c, uint64_t v, const char*) { c.k = (uint16_t)v; },
     "  -k  <int>   k-mer length               [default: 31]\n"
     "              this engine accepts any value in [2,64]\n"},
    {"-m", nullptr, Takes::number, kMax16, [](Config& c, uint64_t v, const char*) { c.l = (uint16_t)v; },
     "  -m  <int>   minimizer length            [default: 21]\n"
     "              any value in [1, min(k-1, 32)]\n"},
    {"-n", nullptr, Takes::number, kMax32,
     [](Config& c, uint64_t v, const char*) { c.num_partitions = (uint32_t)v; },
     "  -n  <int>   number of partitions        [default: auto, ~2 MB input/partition]\n"
     "              rounded up to a power of two; the counts do not depend on it\n"},
    {"-t", nullptr, Takes::number, kMax32,
     [](Config& c, uint64_t v, const char*) { c.num_threads = (uint32_t)v; },
     "  -t  <int>   host threads (input parsing) [default: 1]\n"},
    {"-ci", nullptr, Takes::number, kMax32, [](Config& c, uint64_t v, const char*) { c.ci = (uint32_t)v; },
     "  -ci <int>   minimum k-mer count         [default: 1]\n"},
    {"-cx", nullptr, Takes::number, kMax64, [](Config& c, uint64_t v, const char*) { c.cx = v; },
     "  -cx <int>   maximum k-mer count         [default: max]\n"},
    {"-ram", nullptr, Takes::number, kMax64 >> 30,
     [](Config& c, uint64_t v, const char*) { c.ram_budget_bytes = v << 30; },
     "  -ram <int>  RAM budget in GB            [accepted; device memory is managed]\n"},
    {"-w", nullptr, Takes::path, 0, [](Config& c, uint64_t, const char* t) { c.work_dir = t; },
     "  -w  <dir>   directory for the -kt partition files and -dbg CSVs\n"
     "              [default: tuna_tmp/ next to output file]\n\n"},
    {"-hp", nullptr, Takes::flag, 0, [](Config& c, uint64_t, const char*) { c.hide_progress = true; },
     "  -hp         hide progress messages\n"},
    {"-kt", nullptr, Takes::flag, 0, [](Config& c, uint64_t, const char*) { c.keep_tmp = true; },
     "  -kt         keep partition files (hash_<p>.superkmers in the work dir)\n"},
    {"-tp", nullptr, Takes::flag, 0, [](Config& c, uint64_t, const char*) { c.partition_only = true; },
     "  -tp         stop after partitioning (phase 1 only, for benchmarking)\n"},
    {"-dbg", nullptr, Takes::flag, 0, [](Config& c, uint64_t, const char*) { c.debug_stats = true; },
     "  -dbg        debug stats: device timings, per-partition table report,\n"
     "              minimizer coverage (CSV files in the work dir)\n"},
    {"-kff", nullptr, Takes::flag, 0, [](Config& c, uint64_t, const char*) { c.output_kff = true; },
     "  -kff        KFF binary output (also chosen by a .kff output name)\n"},
    {"-h", "--help", Takes::help, 0, nullptr, "  -h/--help   show this help\n"},
    {"-v", "--version", Takes::version, 0, nullptr, "  -v/--version print version\n"},
};

void print_usage(const char* prog)
{
    std::string text = "tuna — streaming k-mer counter (MI355X engine)\n\nUsage:\n";
    for (const char* form : {" [options] <input1.fa [input2.fa ...]> <output_file>\n",
                             " [options] @<input_list_file>          <output_file>\n"})
        text += std::string("  ") + prog + form;
    text += "\nOptions:\n";
    for (const OptionRow& r : kRows) text += r.help;
    text += "\nOutput: plain text TSV (<kmer>\\t<count>) by default; use -kff for KFF binary.\n";
    std::cerr << text;
}

void complain(const std::string& what) { std::cerr << "tuna: error: " << what << "\n"; }

// Decimal text to an unsigned 64-bit value, with the acceptance rules of
// std::stoull in base 10: optional leading white space and sign (a minus sign
// negates modulo 2^64), at least one digit, nothing after the digits, no
// overflow.
bool decimal_value(const char* text, uint64_t& out)
{
    const char* p = text;
    while (*p && std::isspace((unsigned char)*p)) ++p;
    bool negate = false;
    if (*p == '+' || *p == '-') negate = *p++ == '-';
    if (!std::isdigit((unsigned char)*p)) return false;
    uint64_t v = 0;
    for (; std::isdigit((unsigned char)*p); ++p) {
        const uint64_t d = (uint64_t)(*p - '0');
        if (v > (kMax64 - d) / 10) return false;
        v = v * 10 + d;
    }
    if (*p) return false;
    out = negate ? 0 - v : v;
    return true;
}

const OptionRow* find_row(const std::string& tok)
{
    for (const OptionRow& r : kRows)
        if (tok == r.name || (r.alt && tok == r.alt)) return &r;
    return nullptr;
}

// Inputs: the positionals before the output name, or the lines of an @list
// file (a leading "~/" in its path is the home directory).
bool resolve_inputs(std::vector<std::string> words, Config& cfg)
{
    cfg.output_file = words.back();
    words.pop_back();
    const bool listed = words.size() == 1 && !words[0].empty() && words[0][0] == '@';
    if (!listed) {
        cfg.input_files = std::move(words);
        return true;
    }
    std::string path = words[0].substr(1);
    const char* home = std::getenv("HOME");
    if (home && path.rfind("~/", 0) == 0) path.replace(0, 1, home);
    std::ifstream list(path);
    if (!list) {
        complain("cannot open input list: " + path);
        return false;
    }
    for (std::string line; std::getline(list, line);)
        if (!line.empty()) cfg.input_files.push_back(line);
    return true;
}

// false: print the usage and exit 1 (after any message already printed).
bool parse_args(int argc, char* argv[], Config& cfg)
{
    if (argc < 2) return false;
    std::vector<std::string> words;
    for (int at = 1; at < argc; ++at) {
        const std::string tok = argv[at];
        const OptionRow* row = find_row(tok);
        if (!row) {
            if (!tok.empty() && tok[0] == '-') {
                complain("unknown option: " + tok);
                return false;
            }
            words.push_back(tok);
            continue;
        }
        const std::string name = row->name;
        switch (row->takes) {
        case Takes::help:
            return false;
        case Takes::version:
            std::cerr << "tuna " << TUNA_VERSION << "\n";
            std::exit(0);
        case Takes::flag:
            row->apply(cfg, 0, nullptr);
            break;
        case Takes::path:
        case Takes::number: {
            if (at + 1 >= argc) {
                complain(name + " requires an argument");
                return false;
            }
            const char* text = argv[++at];
            uint64_t v = 0;
            if (row->takes == Takes::number) {
                if (!decimal_value(text, v)) {
                    complain("invalid integer for " + name + ": " + text);
                    return false;
                }
                if (v > row->limit) {
                    complain("value out of range for " + name + ": " + text);
                    return false;
                }
            }
            row->apply(cfg, v, text);
            break;
        }
        }
    }
    if (words.size() < 2) {
        complain("need at least one input file and an output file");
        return false;
    }
    if (!resolve_inputs(std::move(words), cfg)) return false;
    const std::string& out = cfg.output_file;
    if (out.size() >= 4 && out.substr(out.size() - 4) == ".kff") cfg.output_kff = true;
    // the first rule that fails is reported
    const std::string k = std::to_string(cfg.k), m = std::to_string(cfg.l);
    const std::pair<bool, std::string> rules[] = {
        {cfg.input_files.empty(), "no input files"},
        {cfg.k < 2 || cfg.k > 256, "k-mer length (-k " + k + ") must be in [2, 256]"},
        {cfg.l < 1 || cfg.l > 32, "minimizer length (-m " + m + ") must be in [1, 32]"},
        {cfg.l >= cfg.k,
         "minimizer length (-m " + m + ") must be strictly less than k-mer length (-k " + k + ")"},
        {cfg.num_threads == 0, "number of threads (-t) must be > 0"},
    };
    for (const auto& r : rules)
        if (r.first) {
            complain(r.second);
            return false;
        }
    return true;
}

uint32_t round_pow2(uint32_t n)  // src/setup.hpp:53-59
{
    if (n <= 1) return 1;
    if ((n & (n - 1)) == 0) return n;
    while (n & (n - 1)) n &= n - 1;
    return n << 1;
}

double since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::string fmt_s(double s)
{
    char b[32];
    snprintf(b, sizeof(b), "%.3fs", s);
    return b;
}

struct HostChunks {
    uint8_t* bases = nullptr;  // pinned (tg_host_alloc), so the upload runs at full rate
    uint64_t n_bases = 0;
    std::vector<uint64_t> off{0};
    uint64_t seqs = 0;
    HostChunks() = default;
    HostChunks(const HostChunks&) = delete;
    HostChunks& operator=(const HostChunks&) = delete;
    ~HostChunks()
    {
        if (bases) tg_host_free(bases);
    }
};

// Every input file is parsed (on -t threads) straight into one pinned buffer.
void read_inputs(const Config& cfg, HostChunks& h)
{
    // plain files are mapped and parsed in place; gzip input is inflated
    struct Input {
        const uint8_t* d = nullptr;
        uint64_t len = 0;
        bool mapped = false;
        void release()
        {
            if (mapped) tg_unmap_file(d, len);
            else tg_free((void*)d);
            d = nullptr;
        }
        const uint8_t* get() const { return d; }
    };
    std::vector<Input> data;
    std::vector<uint64_t> lens;
    uint64_t total = 0;
    struct Release {
        std::vector<Input>& v;
        ~Release()
        {
            for (auto& x : v)
                if (x.d) x.release();
        }
    } release{data};
    for (const auto& f : cfg.input_files) {
        Input in;
        const bool gz = f.size() > 3 && f.compare(f.size() - 3, 3, ".gz") == 0;
        if (gz) {
            uint8_t* d = nullptr;
            if (tg_read_file(f.c_str(), &d, &in.len) != TG_OK) throw std::runtime_error(tg_last_error());
            in.d = d;
        } else {
            if (tg_map_file(f.c_str(), &in.d, &in.len) != TG_OK) throw std::runtime_error(tg_last_error());
            in.mapped = true;
        }
        data.push_back(in);
        lens.push_back(in.len);
        total += in.len;
    }
    if (tg_host_alloc(std::max<uint64_t>(total, 1), (void**)&h.bases) != TG_OK)
        throw std::runtime_error(tg_last_error());
    const uint32_t threads = std::max(1u, cfg.num_threads);
    for (size_t i = 0; i < data.size(); ++i) {
        const uint64_t len = lens[i];
        // chunk offsets: a first guess, then the bound (every chunk needs a
        // base and a separator)
        for (uint64_t cap = len / 64 + 1024;;) {
            std::unique_ptr<uint64_t[]> o(new uint64_t[cap]);
            const int64_t n = tg_parse_buffer_mt(data[i].get(), len, h.bases + h.n_bases, o.get(), cap, threads);
            if (n == TG_ECAP && cap < len / 2 + 2) {
                cap = len / 2 + 2;
                continue;
            }
            if (n < 0) throw std::runtime_error("cannot parse input: " + cfg.input_files[i]);
            h.off.reserve(h.off.size() + (size_t)n);
            for (int64_t j = 1; j <= n; ++j) h.off.push_back(h.n_bases + o[j]);
            h.n_bases += o[n];
            h.seqs += (uint64_t)n;
            break;
        }
        data[i].release();
    }
}

void check(int rc)
{
    if (rc != TG_OK) throw std::runtime_error(tg_last_error());
}

int write_partition_files(const Config& cfg, tg_ctx* ctx, const tg_params& p, const HostChunks& h)
{
    // -kt: the reference keeps hash_<p>.superkmers (src/superkmer_io.hpp:30-33)
    const uint64_t K = h.n_bases;
    std::vector<uint64_t> part_off(cfg.num_partitions + 1);
    std::vector<uint8_t> bytes(4 * K + 64);
    std::vector<uint64_t> rec_off(K + 2);
    uint64_t nsk = 0, nb = 0;
    check(tg_partitions_host(ctx, &p, h.bases, h.n_bases, h.off.data(), h.off.size() - 1,
                             part_off.data(), bytes.data(), bytes.size(), rec_off.data(), rec_off.size(),
                             &nsk, &nb));
    std::filesystem::create_directories(cfg.work_dir);
    for (uint32_t q = 0; q < cfg.num_partitions; ++q) {
        const std::string path = cfg.work_dir + "hash_" + std::to_string(q) + ".superkmers";
        std::ofstream o(path, std::ios::binary);
        if (!o) {
            std::cerr << "tuna: error: cannot open partition file for writing: " << path << "\n";
            return 1;
        }
        const uint64_t b = rec_off[part_off[q]], e = rec_off[part_off[q + 1]];
        o.write((const char*)bytes.data() + b, (std::streamsize)(e - b));
    }
    return 0;
}

// Table statistics per reference partition, printed and written as
// src/count.hpp:478-565 (emit_debug_stats) does.  A partition's tables are its
// work items' LDS tables: fixed in size (init_sz = table_cap, no resizes).
// n_unique counts the keys held in them; keys that went to the HBM spill
// table (n_overflow, the last CSV column) are reported in total only.
// Minimizer coverage (src/count.hpp:665-705): k-mers per distinct minimizer,
// summarised and written as debug_min_coverage.csv.
int print_min_coverage(const Config& cfg, tg_ctx* ctx, const tg_params& p, const uint8_t* bases,
                       uint64_t n_bases, const std::vector<uint64_t>& off)
{
    // distinct coverage values of N k-mers are fewer than sqrt(2N) + 1 <= 2^20 here
    std::vector<uint32_t> cov(1u << 20);
    std::vector<uint64_t> cnt(1u << 20);
    const int64_t got = tg_debug_min_coverage(ctx, &p, bases, n_bases, off.data(), off.size() - 1, cov.data(),
                                              cnt.data(), cov.size());
    if (got < 0) {
        std::cerr << "tuna: error: " << tg_last_error() << "\n";
        return 1;
    }
    if (got == 0) return 0;
    cov.resize(got);
    cnt.resize(got);
    uint64_t minimizers = 0, covered = 0;
    for (int64_t i = 0; i < got; ++i) {
        minimizers += cnt[i];
        covered += (uint64_t)cov[i] * cnt[i];
    }
    std::ostringstream o;
    o << "\n[debug] minimizer coverage (k-mers sharing one minimizer):\n"
      << "  unique minimizers tracked : " << minimizers << "\n"
      << "  total k-mers covered      : " << covered << "\n"
      << "  avg k-mers / minimizer    : " << std::fixed << std::setprecision(2)
      << (minimizers ? (double)covered / (double)minimizers : 0.0) << "\n"
      << "  max k-mers / minimizer    : " << (got ? cov[got - 1] : 0u) << "\n";
    const std::string path = cfg.work_dir + "debug_min_coverage.csv";
    std::ofstream csv(path);
    if (csv) {
        csv << "coverage,n_minimizers,total_kmers\n";
        for (int64_t i = 0; i < got; ++i) csv << cov[i] << "," << cnt[i] << "," << (uint64_t)cov[i] * cnt[i] << "\n";
        o << "[debug] minimizer coverage CSV written to: " << path << "\n";
    } else {
        o << "[debug] warning: could not write CSV to " << path << "\n";
    }
    std::cerr << o.str();
    return 0;
}

int print_table_stats(const Config& cfg, tg_ctx* ctx)
{
    const uint64_t n = cfg.num_partitions;
    std::vector<uint64_t> ins(n), uni(n), sp(n), cap(n);
    const int64_t np = tg_debug_partitions(ctx, ins.data(), uni.data(), sp.data(), cap.data(), n);
    if (np < 0) {
        std::cerr << "tuna: error: " << tg_last_error() << "\n";
        return 1;
    }
    std::ostringstream o;
    o << "\n[debug] per-partition table stats (first 20):\n"
      << "  part   init_sz   table_cap   n_inserted    n_unique  load%  n_resizes\n";
    auto lf = [&](uint64_t i) { return cap[i] ? (double)uni[i] / (double)cap[i] : 0.0; };
    for (int64_t i = 0; i < np && i < 20; ++i)
        o << "  " << std::setw(5) << i << "  " << std::setw(8) << cap[i] << "  " << std::setw(9) << cap[i]
          << "  " << std::setw(11) << ins[i] << "  " << std::setw(10) << uni[i] << "  " << std::fixed
          << std::setprecision(1) << std::setw(5) << 100.0 * lf(i) << "%  0\n";
    if (np > 20) o << "  ... (" << (np - 20) << " more partitions)\n";
    double s_lf = 0, lo_lf = 2, hi_lf = -1, s_or = 0, lo_or = 1e18, hi_or = -1e18;
    uint64_t s_u = 0, lo_u = UINT64_MAX, hi_u = 0, s_c = 0, lo_c = UINT64_MAX, hi_c = 0, spilled = 0;
    int64_t valid = 0, nonempty = 0;
    for (int64_t i = 0; i < np; ++i) {
        spilled += sp[i];
        if (!cap[i]) continue;
        ++valid;
        s_lf += lf(i);
        lo_lf = std::min(lo_lf, lf(i));
        hi_lf = std::max(hi_lf, lf(i));
        s_u += uni[i];
        lo_u = std::min(lo_u, uni[i]);
        hi_u = std::max(hi_u, uni[i]);
        s_c += cap[i];
        lo_c = std::min(lo_c, cap[i]);
        hi_c = std::max(hi_c, cap[i]);
        if (uni[i]) {
            ++nonempty;
            const double r = (double)cap[i] / (double)uni[i];
            s_or += r;
            lo_or = std::min(lo_or, r);
            hi_or = std::max(hi_or, r);
        }
    }
    if (valid) {
        o << "\n[debug] aggregate table stats (" << np << " partitions):\n" << std::fixed;
        o << "  load_factor:    mean=" << std::setprecision(3) << s_lf / valid << "  min=" << lo_lf
          << "  max=" << hi_lf << "\n";
        if (nonempty)
            o << "  oversize_ratio: mean=" << std::setprecision(1) << s_or / nonempty << "  min=" << lo_or
              << "  max=" << hi_or << "  (init_sz / n_unique)\n";
        o << "  n_unique/part:  mean=" << std::setprecision(0) << (double)s_u / valid << "  min=" << lo_u
          << "  max=" << hi_u << "\n";
        o << "  init_sz/part:   mean=" << (double)s_c / valid << "  min=" << lo_c << "  max=" << hi_c << "\n";
        o << "dbg_load_mean: " << std::setprecision(6) << s_lf / valid << "\n"
          << "dbg_load_min: " << lo_lf << "\n"
          << "dbg_load_max: " << hi_lf << "\n"
          << "dbg_oversize_mean: " << std::setprecision(2) << (nonempty ? s_or / nonempty : 0.0) << "\n"
          << "dbg_unique_mean: " << std::setprecision(0) << (double)s_u / valid << "\n"
          << "dbg_n_resizes: 0\n"
          << "dbg_parts_resized: 0\n";
    }
    if (spilled)
        o << "[debug] k-mers sent to the HBM spill table: " << spilled
          << " (their distinct keys are in unique_kmers, not in n_unique per partition)\n";
    std::error_code ec;
    std::filesystem::create_directories(cfg.work_dir, ec);
    const std::string path = cfg.work_dir + "debug_table_stats.csv";
    std::ofstream csv(path);
    if (csv) {
        csv << "partition_id,init_sz,table_cap,n_inserted,n_unique,load_factor,n_resizes,resize_s,n_overflow\n";
        for (int64_t i = 0; i < np; ++i)
            csv << i << "," << cap[i] << "," << cap[i] << "," << ins[i] << "," << uni[i] << "," << std::fixed
                << std::setprecision(6) << lf(i) << ",0," << 0.0 << "," << sp[i] << "\n";
        o << "[debug] table stats CSV: " << path << "\n";
    } else {
        o << "[debug] warning: could not write table stats CSV to " << path << "\n";
    }
    o << "\n[debug] resize summary:\n"
      << "  partitions with resizes : 0 / " << np << "\n"
      << "  total resize events     : 0\n"
      << "  total resize time       : 0.000s\n"
      << "  overflow-triggered      : 0\n"
      << "  load-triggered          : 0\n";
    std::cerr << o.str();
    return 0;
}

int run(Config& cfg)
{
    const auto t_start = std::chrono::steady_clock::now();
    tg_ctx* ctx = nullptr;
    check(tg_ctx_create(0, &ctx));
    std::unique_ptr<tg_ctx, void (*)(tg_ctx*)> hold(ctx, tg_ctx_destroy);
    if (cfg.debug_stats) {
        // per-kernel times, and table statistics per partition (the plain
        // launch sequence records them; the counts are the same either way)
        tg_set_profiling(ctx, 1);
        check(tg_set_plain_sequence(ctx, 1));
        check(tg_set_debug_stats(ctx, 1));
    }

    if (!cfg.hide_progress)
        std::cerr << "[1/2] partitioning  (" << cfg.num_threads << " thread"
                  << (cfg.num_threads > 1 ? "s" : "") << ", in-memory, gfx950) ...\n";
    const auto t_part = std::chrono::steady_clock::now();
    tg_params p{cfg.k, cfg.l, cfg.num_partitions, cfg.ci, cfg.cx};
    // -tp, -kt and -dbg work from a host copy of the bases; a plain count
    // reads the files straight to the device, the upload overlapping the read
    const bool host_copy = cfg.partition_only || cfg.keep_tmp || cfg.debug_stats;
    HostChunks h;
    if (host_copy) read_inputs(cfg, h);
    if (cfg.partition_only) {
        uint64_t nsk = 0;
        uint32_t nsp = 0;
        // counts only (no output buffers): TG_ECAP just reports that none were copied
        const int rc = tg_superkmers_host(ctx, &p, h.bases, h.n_bases, h.off.data(),
                                          h.off.size() - 1, nullptr, nullptr, nullptr, nullptr, 0, &nsk, &nsp);
        if (rc != TG_OK && rc != TG_ECAP) check(rc);
        if (cfg.keep_tmp && write_partition_files(cfg, ctx, p, h)) return 1;
        const double t1 = since(t_part);
        uint64_t K = 0;
        for (size_t i = 0; i + 1 < h.off.size(); ++i) {
            const uint64_t L = h.off[i + 1] - h.off[i];
            if (L >= cfg.k) K += L - cfg.k + 1;
        }
        if (!cfg.hide_progress) std::cerr << "      " << h.seqs << " seqs  " << K << " k-mers  " << fmt_s(t1) << "\n";
        std::cerr << "phase1: " << t1 << "s\n"
                  << "superkmers: " << nsk << "\n";
        if (!cfg.hide_progress) std::cerr << "done  (partition only)\n";
        return 0;
    }
    tg_stats st{};
    if (host_copy) {
        check(tg_count_host(ctx, &p, h.bases, h.n_bases, h.off.data(), h.off.size() - 1, &st));
    } else {
        std::vector<const char*> paths;
        for (const auto& f : cfg.input_files) paths.push_back(f.c_str());
        check(tg_count_files(ctx, &p, paths.data(), (uint32_t)paths.size(), std::max(1u, cfg.num_threads), &st));
        h.seqs = st.n_chunks;
    }
    // phase split on the wall clock: parse + device phase 1 | device phase 2 + output
    const double t_all = since(t_part);
    const double t_phase1 = std::max(0.0, t_all - st.ms_phase2 * 1e-3);
    if (!cfg.hide_progress) {
        std::cerr << "      " << h.seqs << " seqs  " << st.n_kmers << " k-mers  " << fmt_s(t_phase1) << "\n";
        std::cerr << "[2/2] counting  (gfx950, in-memory) ...\n";
    }
    const auto t2 = std::chrono::steady_clock::now();
    // the counts are formatted on the device (TSV lines or KFF records) and
    // only their bytes cross to the host, overlapping the file writes
    if (tg_result_write(ctx, cfg.k, cfg.output_file.c_str(), cfg.output_kff ? 1 : 0) != TG_OK) {
        std::cerr << "tuna: error: " << tg_last_error() << "\n";
        return 1;
    }
    if (cfg.keep_tmp && write_partition_files(cfg, ctx, p, h)) return 1;
    const double t_phase2 = st.ms_phase2 * 1e-3 + since(t2);
    if (!cfg.hide_progress)
        std::cerr << "      " << st.n_kmers << " k-mers in  " << st.n_out << " unique out  " << fmt_s(t_phase2) << "\n";
    std::cerr << "phase1: " << t_phase1 << "s\n"
              << "phase2: " << t_phase2 << "s\n"
              << "superkmers: " << st.n_superkmers << "\n"
              << "n_parts: " << cfg.num_partitions << "\n"
              << "unique_kmers: " << st.n_out << "\n";
    if (cfg.debug_stats) {
        char names[4096];
        double ms[64];
        uint64_t cnt[64];
        const int nk = tg_kernel_times(ctx, names, sizeof(names), ms, cnt, 64);
        const char* q = names;
        std::cerr << "[dbg] device phase1 " << st.ms_phase1 << " ms, phase2 " << st.ms_phase2
                  << " ms, sub-partitions " << st.n_subparts << ", work items " << st.n_lds_groups
                  << ", spilled k-mers " << st.n_hbm_parts << "\n";
        for (int i = 0; i < nk && i < 64; ++i) {
            std::cerr << "[dbg] kernel " << q << " " << ms[i] << " ms x" << cnt[i] << "\n";
            q += strlen(q) + 1;
        }
        if (print_table_stats(cfg, ctx)) return 1;
        if (print_min_coverage(cfg, ctx, p, h.bases, h.n_bases, h.off)) return 1;
    }
    if (!cfg.hide_progress) std::cerr << "total: " << fmt_s(since(t_start)) << "\n";
    return 0;
}

}  // namespace

int main(int argc, char* argv[])
{
    Config cfg;
    if (!parse_args(argc, argv, cfg)) {
        print_usage(argv[0]);
        return 1;
    }
    if (cfg.k > 64 && (cfg.partition_only || cfg.keep_tmp || cfg.debug_stats)) {
        // k > 64 is counted without superkmer records (src/pipeline.hpp:545-551
        // is the reference's message form for combinations a build lacks)
        std::cerr << "tuna: error: unsupported combination k=" << cfg.k << " with -tp, -kt or -dbg\n"
                  << "  Partition files and debug statistics are built for k in [2,64].\n";
        return 1;
    }
    if (cfg.num_partitions == 0) {
        std::vector<const char*> fp;
        for (const auto& f : cfg.input_files) fp.push_back(f.c_str());
        cfg.num_partitions = tg_auto_partitions(fp.data(), (uint32_t)fp.size(), cfg.k, cfg.num_threads);
    }
    cfg.num_partitions = round_pow2(cfg.num_partitions);
    if (cfg.work_dir.empty()) {
        std::filesystem::path parent = std::filesystem::path(cfg.output_file).parent_path();
        if (parent.empty()) parent = std::filesystem::current_path();
        cfg.work_dir = (parent / "tuna_tmp").string();
    }
    if (cfg.work_dir.back() != '/') cfg.work_dir += '/';
    if (!cfg.hide_progress) {
        std::cerr << "tuna  k=" << cfg.k << "  m=" << cfg.l << "  n=" << cfg.num_partitions
                  << "  t=" << cfg.num_threads;
        if (cfg.ci > 1 || cfg.cx != std::numeric_limits<uint64_t>::max())
            std::cerr << "  ci=" << cfg.ci << "  cx=" << cfg.cx;
        std::cerr << "\n";
    }
    int rc = 1;
    try {
        rc = run(cfg);
    } catch (const std::exception& e) {
        std::cerr << "tuna: error: " << e.what() << "\n";
        rc = 1;
    } catch (...) {
        std::cerr << "tuna: error: unexpected failure\n";
        rc = 1;
    }
    return rc;
}
