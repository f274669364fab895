// host_io.cpp — host input path (ACGT chunking of FASTA/FASTQ) and the
// seeded synthetic workload generators.
//
// Chunking follows the reference configuration HELICASE_ACTG
// (src/seq_source.hpp:33-39): headers ignored, sequence lines joined,
// every byte outside [ACGTacgt] ends a chunk.
#include "../../include/tuna_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// 1 for A C G T a c g t (include/helicase_cpp/simd.hpp:305-307)
struct AcgtTable {
    uint8_t t[256];
    AcgtTable()
    {
        memset(t, 0, sizeof(t));
        for (const char* p = "ACGTacgt"; *p; ++p) t[(uint8_t)*p] = 1;
    }
};
const AcgtTable ACGT;

inline uint64_t lf_at_or_after(const uint8_t* d, uint64_t len, uint64_t from)
{
    if (from >= len) return len;
    const void* q = memchr(d + from, '\n', len - from);
    return q ? (uint64_t)((const uint8_t*)q - d) : len;
}

// FASTA: include/helicase_cpp/lexer.hpp:46-71 (a '>' opens a header that runs
// through the next '\n'), parser_fasta.hpp:233-316 (bytes before the first
// header are skipped; '\n' inside sequence is transparent; any other
// non-ACGT byte or a header closes the chunk).
int64_t parse_fasta(const uint8_t* d, uint64_t len, uint8_t* ob, uint64_t* off, uint64_t cap)
{
    const void* first = memchr(d, '>', len);
    if (!first) return 0;
    uint64_t i = (uint64_t)((const uint8_t*)first - d);
    uint64_t nb = 0, nc = 0;
    bool in_chunk = false;
    off[0] = 0;
    while (i < len) {
        const uint8_t c = d[i];
        if (c == '>') {
            if (in_chunk) {
                if (nc + 1 >= cap) return TG_ECAP;
                off[++nc] = nb;
                in_chunk = false;
            }
            i = lf_at_or_after(d, len, i) + 1;
            continue;
        }
        if (ACGT.t[c]) {
            // copy the run up to the next non-ACGT byte
            uint64_t j = i;
            while (j < len && ACGT.t[d[j]]) ++j;
            memcpy(ob + nb, d + i, j - i);
            nb += j - i;
            in_chunk = true;
            i = j;
            continue;
        }
        if (c != '\n' && in_chunk) {
            if (nc + 1 >= cap) return TG_ECAP;
            off[++nc] = nb;
            in_chunk = false;
        }
        ++i;
    }
    if (in_chunk) {
        if (nc + 1 >= cap) return TG_ECAP;
        off[++nc] = nb;
    }
    return (int64_t)nc;
}

// FASTQ: include/helicase_cpp/parser_fastq.hpp:112-285.  Four-line records,
// line 2 split at non-ACGT.  The 64-byte block lexer has two boundary
// behaviours that this scanner keeps: a header line's first byte is stepped
// over and, when that byte is a '\n' at block offset 63, the line feed is
// lost with the block (:115, increment_pos); a DNA run that reaches end of
// input exactly at a multiple of 64 bytes is not delivered (:181).
int64_t parse_fastq(const uint8_t* d, uint64_t len, uint8_t* ob, uint64_t* off, uint64_t cap)
{
    uint64_t pos = 0, nb = 0, nc = 0;
    off[0] = 0;
    for (;;) {
        // header line
        if (pos + 1 >= len) return (int64_t)nc;
        uint64_t lf = lf_at_or_after(d, len, (pos % 64 == 63) ? pos + 1 : pos);
        if (lf >= len) return (int64_t)nc;
        pos = lf + 1;
        // sequence line
        for (;;) {
            uint64_t s = pos;
            while (s < len && !ACGT.t[d[s]] && d[s] != '\n') ++s;
            if (s >= len) return (int64_t)nc;
            if (d[s] == '\n') {
                pos = s + 1;
                break;
            }
            uint64_t r = s;
            while (r < len && ACGT.t[d[r]]) ++r;
            if (r == len && (len & 63u) == 0) return (int64_t)nc;
            memcpy(ob + nb, d + s, r - s);
            nb += r - s;
            if (nc + 1 >= cap) return TG_ECAP;
            off[++nc] = nb;
            if (r < len && d[r] == '\n') {
                pos = r + 1;
                break;
            }
            pos = r;
        }
        // '+' line, quality line
        for (int l = 0; l < 2; ++l) {
            lf = lf_at_or_after(d, len, pos);
            if (lf >= len) return (int64_t)nc;
            pos = lf + 1;
        }
    }
}

// ------------------------------------------------------------ synthetic ----
// xoshiro256** seeded by splitmix64 (Blackman & Vigna).
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double unif() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = unif(), u2 = unif();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t prob_thresh(double p)
{
    if (p <= 0) return 0;
    if (p >= 1) return ~0ull;
    return (uint64_t)(p * 18446744073709551616.0);
}

const char BASES[4] = {'A', 'C', 'G', 'T'};

// base with GC fraction gc: C and G gc/2 each, A and T (1-gc)/2 each
inline char gc_base(Rng& r, double gc)
{
    if (gc == 0.5) return BASES[r.next() >> 62];
    const double u = r.unif();
    if (u < gc * 0.5) return 'C';
    if (u < gc) return 'G';
    if (u < gc + (1 - gc) * 0.5) return 'A';
    return 'T';
}

inline char comp(char c)
{
    switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    default: return 'A';
    }
}

void random_genome(Rng& r, uint64_t n, double gc, std::string& g)
{
    g.resize(n);
    for (uint64_t i = 0; i < n; ++i) g[i] = gc_base(r, gc);
}

// Repetitive genome: interspersed copies of `fam` random families of `flen`
// bases (each copy with `div` i.i.d. substitutions) make up frac_rep of the
// length, tandem arrays of 2-12 bp units over 200-2000 bp make up frac_tan,
// the rest is random sequence; segments are shuffled and separated by random
// stretches.
void repeat_genome(Rng& r, uint64_t n, double gc, uint32_t fam, uint32_t flen, double div,
                   double frac_rep, double frac_tan, std::string& g)
{
    std::vector<std::string> fams(fam);
    for (auto& f : fams) random_genome(r, flen, gc, f);
    std::vector<std::string> segs;
    uint64_t rep_target = (uint64_t)(frac_rep * (double)n), tan_target = (uint64_t)(frac_tan * (double)n);
    uint64_t rep = 0, tan = 0;
    const uint64_t dth = prob_thresh(div);
    while (rep + flen <= rep_target) {
        std::string s = fams[r.below(fam)];
        for (auto& ch : s)
            if (r.next() < dth) ch = BASES[(((ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : 3) + 1 + r.below(3)) & 3)];
        rep += s.size();
        segs.push_back(std::move(s));
    }
    while (tan < tan_target) {
        const uint32_t ul = 2 + (uint32_t)r.below(11);
        uint64_t al = 200 + r.below(1801);
        if (tan + al > tan_target) al = tan_target - tan;
        std::string unit;
        random_genome(r, ul, gc, unit);
        std::string s(al, 'A');
        for (uint64_t i = 0; i < al; ++i) s[i] = unit[i % ul];
        tan += al;
        segs.push_back(std::move(s));
    }
    for (uint64_t i = segs.size(); i > 1; --i) std::swap(segs[i - 1], segs[r.below(i)]);
    const uint64_t used = rep + tan;
    const uint64_t rnd = n > used ? n - used : 0;
    // gaps: random cut points of the random-sequence budget
    std::vector<uint64_t> cuts(segs.size());
    for (auto& cpos : cuts) cpos = r.below(rnd + 1);
    std::sort(cuts.begin(), cuts.end());
    g.clear();
    g.reserve(n);
    uint64_t prev = 0;
    std::string tmp;
    for (size_t i = 0; i < segs.size(); ++i) {
        random_genome(r, cuts[i] - prev, gc, tmp);
        g += tmp;
        prev = cuts[i];
        g += segs[i];
    }
    random_genome(r, rnd - prev, gc, tmp);
    g += tmp;
}

// Reads of length rl at uniform positions and strands, substitution rate sub,
// N rate pn (one draw per base: [0,sub) substitutes, [sub, sub+pn) gives N).
void sample_reads_fq(Rng& r, const std::string& g, uint64_t nreads, uint32_t rl, double sub,
                     double pn, std::string& out, uint64_t id0)
{
    const uint64_t ts = prob_thresh(sub), tn = prob_thresh(sub + pn);
    std::string rd(rl, 'A'), q(rl, '?');
    const uint64_t span = g.size() >= rl ? g.size() - rl + 1 : 1;
    char hdr[64];
    for (uint64_t i = 0; i < nreads; ++i) {
        const uint64_t p = r.below(span);
        const bool rc = r.next() >> 63;
        for (uint32_t j = 0; j < rl; ++j) rd[j] = rc ? comp(g[p + rl - 1 - j]) : g[p + j];
        for (uint32_t j = 0; j < rl; ++j) {
            const uint64_t u = r.next();
            if (u < ts) {
                const int b = rd[j] == 'A' ? 0 : rd[j] == 'C' ? 1 : rd[j] == 'G' ? 2 : 3;
                rd[j] = BASES[(b + 1 + (int)r.below(3)) & 3];
            } else if (u < tn) {
                rd[j] = 'N';
            }
        }
        const int hl = snprintf(hdr, sizeof(hdr), "@r%llu\n", (unsigned long long)(id0 + i));
        out.append(hdr, (size_t)hl);
        out += rd;
        out += "\n+\n";
        out += q;
        out += '\n';
    }
}

void fasta(const std::string& name, const std::string& g, std::string& out)
{
    out += '>';
    out += name;
    out += '\n';
    for (uint64_t i = 0; i < g.size(); i += 80) {
        out.append(g, i, 80);
        out += '\n';
    }
}

std::string g_synth;

double spec_num(const std::string& spec, const char* key, double dflt)
{
    const std::string k = std::string(key) + "=";
    size_t p = 0;
    while ((p = spec.find(k, p)) != std::string::npos) {
        if (p == 0 || spec[p - 1] == ' ' || spec[p - 1] == ',') return atof(spec.c_str() + p + k.size());
        p += k.size();
    }
    return dflt;
}
std::string spec_str(const std::string& spec, const char* key, const char* dflt)
{
    const std::string k = std::string(key) + "=";
    size_t p = spec.find(k);
    if (p == std::string::npos) return dflt;
    p += k.size();
    size_t e = spec.find_first_of(" ,", p);
    return spec.substr(p, e == std::string::npos ? std::string::npos : e - p);
}

}  // namespace

extern "C" {

int64_t tg_parse_buffer(const uint8_t* data, uint64_t len, uint8_t* out_bases, uint64_t* out_off,
                        uint64_t off_cap)
{
    if (len == 0) return TG_EINPUT;
    if (off_cap < 1) return TG_ECAP;
    return data[0] == '@' ? parse_fastq(data, len, out_bases, out_off, off_cap)
                          : parse_fasta(data, len, out_bases, out_off, off_cap);
}

// Workloads (all sizes in bases; seed selects the stream):
//   kind=genome_fa   len, seed, gc                     one FASTA record
//   kind=reads_fq    genome, seed, gc, reads|cov, rl, sub, nrate
//   kind=repeat_fq   genome, seed, gc, cov, rl, sub, fam, flen, div, rep, tan
//   kind=samples_fq  samples, seed, median, sigma, lo, hi, cov, rl, sub, index
//                    (the index-th sample of a log-normal family)
int64_t tg_synth(const char* spec_c)
{
    const std::string spec = spec_c ? spec_c : "";
    const std::string kind = spec_str(spec, "kind", "genome_fa");
    const uint64_t seed = (uint64_t)spec_num(spec, "seed", 1);
    const double gc = spec_num(spec, "gc", 0.5);
    g_synth.clear();
    Rng r(seed);
    std::string g;
    if (kind == "genome_fa") {
        random_genome(r, (uint64_t)spec_num(spec, "len", 4600000), gc, g);
        fasta("synthetic_seed" + std::to_string(seed), g, g_synth);
    } else if (kind == "reads_fq" || kind == "repeat_fq") {
        const uint64_t G = (uint64_t)spec_num(spec, "genome", 5000000);
        if (kind == "reads_fq") random_genome(r, G, gc, g);
        else
            repeat_genome(r, G, gc, (uint32_t)spec_num(spec, "fam", 50), (uint32_t)spec_num(spec, "flen", 300),
                          spec_num(spec, "div", 0.05), spec_num(spec, "rep", 0.35), spec_num(spec, "tan", 0.10), g);
        const uint32_t rl = (uint32_t)spec_num(spec, "rl", 150);
        uint64_t nreads = (uint64_t)spec_num(spec, "reads", 0);
        if (!nreads) nreads = (uint64_t)std::llround(spec_num(spec, "cov", 50) * (double)G / rl);
        g_synth.reserve(nreads * (2ull * rl + 16));
        Rng rr(seed * 0x100000001b3ull + 7);
        sample_reads_fq(rr, g, nreads, rl, spec_num(spec, "sub", 0.005), spec_num(spec, "nrate", 0.0), g_synth, 0);
    } else if (kind == "samples_fq") {
        const uint64_t idx = (uint64_t)spec_num(spec, "index", 0);
        Rng lr(seed);
        double len = 0;
        const double med = spec_num(spec, "median", 300000), sg = spec_num(spec, "sigma", 1.2);
        const double lo = spec_num(spec, "lo", 10000), hi = spec_num(spec, "hi", 8000000);
        for (uint64_t i = 0; i <= idx; ++i) len = med * std::exp(sg * lr.normal());
        len = std::min(std::max(len, lo), hi);
        Rng sr(seed * 1000003ull + idx * 7919ull + 11);
        random_genome(sr, (uint64_t)len, gc, g);
        const uint32_t rl = (uint32_t)spec_num(spec, "rl", 150);
        const uint64_t nreads = (uint64_t)std::llround(spec_num(spec, "cov", 10) * len / rl);
        sample_reads_fq(sr, g, nreads, rl, spec_num(spec, "sub", 0.005), spec_num(spec, "nrate", 0.0), g_synth, 0);
    } else {
        return TG_EINVAL;
    }
    return (int64_t)g_synth.size();
}

int tg_synth_copy(uint8_t* dst, uint64_t cap)
{
    if (!dst || cap < g_synth.size()) return TG_ECAP;
    memcpy(dst, g_synth.data(), g_synth.size());
    return TG_OK;
}

void tg_synth_free(void)
{
    std::string().swap(g_synth);
}

}  // extern "C"
