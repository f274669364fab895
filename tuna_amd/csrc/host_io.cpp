// host_io.cpp — host input path (ACGT chunking of FASTA/FASTQ) and the
// seeded synthetic workload generators.
//
// Chunking follows the reference configuration HELICASE_ACTG
// (src/seq_source.hpp:33-39): headers ignored, sequence lines joined,
// every byte outside [ACGTacgt] ends a chunk.
#include "../../include/tuna_gpu.h"
#include "host_io.h"
#include "kff.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/resource.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <cstdio>
#include <fstream>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
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

// Output of a scan: DNA runs and chunk ends, in order.  The direct sink
// writes into the caller's arrays (TG_ECAP when the offsets do not fit); the
// section sink keeps a section's output for the merge of a parallel parse.
struct DirectSink {
    uint8_t* ob;
    uint64_t* off;
    uint64_t cap, nb = 0, nc = 0;
    bool full = false;
    bool bases(const uint8_t* p, uint64_t n)
    {
        memcpy(ob + nb, p, n);
        nb += n;
        return true;
    }
    bool end()
    {
        if (nc + 1 >= cap) {
            full = true;
            return false;
        }
        off[++nc] = nb;
        return true;
    }
};

// A section of a parallel parse is scanned twice: once to count its bases
// and chunk ends, then again to write them at offsets known from the counts.
struct CountSink {
    uint64_t nb = 0, ne = 0;
    bool bases(const uint8_t*, uint64_t n)
    {
        nb += n;
        return true;
    }
    bool end()
    {
        ++ne;
        return true;
    }
};

struct PlaceSink {
    uint8_t* ob;     // the section's first base
    uint64_t* off;   // the section's first chunk-end slot
    uint64_t base;   // the section's first base, as a global offset
    uint64_t nb = 0, ne = 0;
    bool bases(const uint8_t* p, uint64_t n)
    {
        memcpy(ob + nb, p, n);
        nb += n;
        return true;
    }
    bool end()
    {
        off[ne++] = base + nb;
        return true;
    }
};

// FASTA: include/helicase_cpp/lexer.hpp:46-71 (a '>' opens a header that runs
// through the next '\n'), parser_fasta.hpp:233-316 (bytes before the first
// header are skipped; '\n' inside sequence is transparent; any other
// non-ACGT byte or a header closes the chunk).  Scans [i, stop) with the chunk
// state `in_chunk`; a run is cut at `stop` (the next section continues it).
// `closed` reports a chunk-closing byte ('>' or non-ACGT, non-'\n') seen
// before the first base, which a parallel merge needs.
template <class Sink>
bool fasta_scan(const uint8_t* d, uint64_t len, uint64_t i, uint64_t stop, Sink& sink, bool& in_chunk,
                bool& closed)
{
    bool seen_base = false;
    while (i < stop) {
        const uint8_t c = d[i];
        if (c == '>') {
            if (!seen_base) closed = true;
            if (in_chunk) {
                if (!sink.end()) return false;
                in_chunk = false;
            }
            i = lf_at_or_after(d, len, i) + 1;
            continue;
        }
        if (ACGT.t[c]) {
            uint64_t j = i;
            while (j < stop && ACGT.t[d[j]]) ++j;
            if (!sink.bases(d + i, j - i)) return false;
            in_chunk = true;
            seen_base = true;
            i = j;
            continue;
        }
        if (c != '\n') {
            if (!seen_base) closed = true;
            if (in_chunk) {
                if (!sink.end()) return false;
                in_chunk = false;
            }
        }
        ++i;
    }
    return true;
}

int64_t parse_fasta(const uint8_t* d, uint64_t len, uint8_t* ob, uint64_t* off, uint64_t cap)
{
    const void* first = memchr(d, '>', len);
    if (!first) return 0;
    DirectSink k{ob, off, cap};
    off[0] = 0;
    bool in_chunk = false, closed = false;
    fasta_scan(d, len, (uint64_t)((const uint8_t*)first - d), len, k, in_chunk, closed);
    if (k.full) return TG_ECAP;
    if (in_chunk && !k.end()) return TG_ECAP;
    return (int64_t)k.nc;
}

// FASTQ: include/helicase_cpp/parser_fastq.hpp:112-285.  Four-line records,
// line 2 split at non-ACGT.  The 64-byte block lexer has two boundary
// behaviours that this scanner keeps: a header line's first byte is stepped
// over and, when that byte is a '\n' at block offset 63, the line feed is
// lost with the block (:115, increment_pos); a DNA run that reaches end of
// input exactly at a multiple of 64 bytes is not delivered (:181).
//
// Scans the records from `pos` (a record start) up to the first record that
// starts at or after `stop`.  `quirk` reports a header whose first byte is a
// '\n' at block offset 63: that record spans five lines.
// (fastq_scan_at leaves `pos` at the record start where it stopped, so a
// scan over a growing buffer can resume there.)
template <class Sink>
void fastq_scan_at(const uint8_t* d, uint64_t len, uint64_t& pos, uint64_t stop, Sink& sink, bool& quirk)
{
    for (;;) {
        // header line
        if (pos >= stop || pos + 1 >= len) return;
        if (pos % 64 == 63 && d[pos] == '\n') quirk = true;
        uint64_t lf = lf_at_or_after(d, len, (pos % 64 == 63) ? pos + 1 : pos);
        if (lf >= len) return;
        pos = lf + 1;
        // sequence line
        for (;;) {
            uint64_t s = pos;
            while (s < len && !ACGT.t[d[s]] && d[s] != '\n') ++s;
            if (s >= len) return;
            if (d[s] == '\n') {
                pos = s + 1;
                break;
            }
            uint64_t r = s;
            while (r < len && ACGT.t[d[r]]) ++r;
            if (r == len && (len & 63u) == 0) return;
            if (!sink.bases(d + s, r - s) || !sink.end()) return;
            if (r < len && d[r] == '\n') {
                pos = r + 1;
                break;
            }
            pos = r;
        }
        // '+' line, quality line
        for (int l = 0; l < 2; ++l) {
            lf = lf_at_or_after(d, len, pos);
            if (lf >= len) return;
            pos = lf + 1;
        }
    }
}

template <class Sink>
void fastq_scan(const uint8_t* d, uint64_t len, uint64_t pos, uint64_t stop, Sink& sink, bool& quirk)
{
    fastq_scan_at(d, len, pos, stop, sink, quirk);
}

int64_t parse_fastq(const uint8_t* d, uint64_t len, uint8_t* ob, uint64_t* off, uint64_t cap)
{
    DirectSink k{ob, off, cap};
    off[0] = 0;
    bool quirk = false;
    fastq_scan(d, len, 0, len, k, quirk);
    return k.full ? TG_ECAP : (int64_t)k.nc;
}

// ---- parallel parse: the same output from S sections on S threads ----
template <class F>
void run_sections(uint32_t S, F&& f)
{
    std::vector<std::thread> th;
    th.reserve(S);
    for (uint32_t q = 0; q < S; ++q) th.emplace_back(f, q);
    for (auto& t : th) t.join();
}

// Placement of the sections' outputs: base offset and first chunk-end slot of
// each.  A FASTA chunk may continue across a cut (`carry`); when a closing
// byte opens the next section, the carried chunk ends at that section's first
// base, in an extra slot before the section's own.  Returns the chunk count.
struct SectionPlan {
    std::vector<uint64_t> bbase, ebase;
    std::vector<uint8_t> close_carry;  // section q closes the carried chunk first
    bool close_last = false;           // the text ends inside a chunk
    uint64_t n_ends = 0;
};

SectionPlan plan_sections(const std::vector<CountSink>& cnt, const std::vector<uint8_t>& open_end,
                          const std::vector<uint8_t>& closed_first, bool fasta)
{
    const size_t S = cnt.size();
    SectionPlan p;
    p.bbase.assign(S + 1, 0);
    p.ebase.assign(S + 1, 0);
    p.close_carry.assign(S, 0);
    bool carry = false;
    uint64_t e = 0;
    for (size_t q = 0; q < S; ++q) {
        p.bbase[q + 1] = p.bbase[q] + cnt[q].nb;
        if (fasta && carry && closed_first[q]) {
            p.close_carry[q] = 1;
            ++e;
            carry = false;
        }
        p.ebase[q] = e;
        e += cnt[q].ne;
        if (fasta && (cnt[q].ne || cnt[q].nb || closed_first[q])) carry = open_end[q];
    }
    p.ebase[S] = e;
    if (fasta && carry) {
        p.close_last = true;
        ++e;
    }
    p.n_ends = e;
    return p;
}

int64_t parse_fastq_mt(const uint8_t* d, uint64_t len, uint8_t* ob, uint64_t* off, uint64_t cap, uint32_t S)
{
    std::vector<uint64_t> st(S + 1), nl(S, 0), first(S + 1, len + 1);
    for (uint32_t q = 0; q <= S; ++q) st[q] = len * q / S;
    run_sections(S, [&](uint32_t q) {
        uint64_t n = 0;
        for (const uint8_t* p = d + st[q]; p < d + st[q + 1];) {
            const void* x = memchr(p, '\n', (size_t)(d + st[q + 1] - p));
            if (!x) break;
            ++n;
            p = (const uint8_t*)x + 1;
        }
        nl[q] = n;
    });
    // first record start of each section: the first line start at or after
    // the section start whose line index is a multiple of 4
    uint64_t lines = 0;
    for (uint32_t q = 0; q < S; ++q) {
        uint64_t p = st[q], idx = lines;
        if (p > 0 && d[p - 1] != '\n') {
            p = lf_at_or_after(d, len, p) + 1;
            idx += 1;
        }
        while (p < st[q + 1] && (idx & 3u)) {
            p = lf_at_or_after(d, len, p) + 1;
            ++idx;
        }
        first[q] = p < st[q + 1] ? p : len + 1;
        lines += nl[q];
    }
    // a section's records run up to the next section's first record
    for (int q = (int)S - 1; q >= 0; --q)
        if (first[q] > len) first[q] = first[q + 1];
    std::vector<CountSink> cnt(S);
    std::vector<uint8_t> quirk(S, 0);
    run_sections(S, [&](uint32_t q) {
        if (first[q] > len) return;
        bool qk = false;
        fastq_scan(d, len, first[q], first[q + 1], cnt[q], qk);
        quirk[q] = qk;
    });
    for (uint32_t q = 0; q < S; ++q)
        if (quirk[q]) return parse_fastq(d, len, ob, off, cap);  // records do not follow the lines
    const std::vector<uint8_t> none(S, 0);
    const SectionPlan pl = plan_sections(cnt, none, none, false);
    if (pl.n_ends >= cap) return TG_ECAP;
    off[0] = 0;
    run_sections(S, [&](uint32_t q) {
        if (first[q] > len) return;
        PlaceSink k{ob + pl.bbase[q], off + 1 + pl.ebase[q], pl.bbase[q]};
        bool qk = false;
        fastq_scan(d, len, first[q], first[q + 1], k, qk);
    });
    return (int64_t)pl.n_ends;
}

int64_t parse_fasta_mt(const uint8_t* d, uint64_t len, uint8_t* ob, uint64_t* off, uint64_t cap, uint32_t S)
{
    const void* fp = memchr(d, '>', len);
    if (!fp) return 0;
    const uint64_t f0 = (uint64_t)((const uint8_t*)fp - d);
    std::vector<uint64_t> st(S + 1);
    for (uint32_t q = 0; q <= S; ++q) st[q] = std::max(f0, len * q / S);
    // a section starts inside a header when the last '>' before it comes
    // after the last '\n' before it: last of each per section, then maxima
    std::vector<int64_t> lnl(S, -1), lgt(S, -1);
    run_sections(S, [&](uint32_t q) {
        const size_t n = (size_t)(st[q + 1] - st[q]);
        const void* a = n ? memrchr(d + st[q], '\n', n) : nullptr;
        const void* b = n ? memrchr(d + st[q], '>', n) : nullptr;
        if (a) lnl[q] = (int64_t)((const uint8_t*)a - d);
        if (b) lgt[q] = (int64_t)((const uint8_t*)b - d);
    });
    std::vector<uint8_t> in_hdr(S, 0);
    {
        int64_t nl_before = -1, gt_before = -1;
        for (uint32_t q = 0; q < S; ++q) {
            in_hdr[q] = gt_before > nl_before;
            nl_before = std::max(nl_before, lnl[q]);
            gt_before = std::max(gt_before, lgt[q]);
        }
    }
    std::vector<uint64_t> from(S, 0);
    std::vector<uint8_t> skip_hdr(S, 0);
    for (uint32_t q = 0; q < S; ++q) {
        from[q] = st[q];
        if (st[q] < st[q + 1] && st[q] > f0 && in_hdr[q]) {
            // the rest of the header is skipped (a closing event for a chunk
            // carried into this section)
            skip_hdr[q] = 1;
            from[q] = lf_at_or_after(d, len, st[q]) + 1;
        }
    }
    std::vector<CountSink> cnt(S);
    std::vector<uint8_t> open_end(S, 0), closed_first(S, 0);
    run_sections(S, [&](uint32_t q) {
        if (st[q] >= st[q + 1]) return;
        bool in_chunk = false, cl = false;
        if (from[q] < st[q + 1]) fasta_scan(d, len, from[q], st[q + 1], cnt[q], in_chunk, cl);
        open_end[q] = in_chunk;
        closed_first[q] = skip_hdr[q] || cl;
    });
    const SectionPlan pl = plan_sections(cnt, open_end, closed_first, true);
    if (pl.n_ends >= cap) return TG_ECAP;
    off[0] = 0;
    run_sections(S, [&](uint32_t q) {
        if (pl.close_carry[q]) off[1 + pl.ebase[q] - 1] = pl.bbase[q];
        if (st[q] >= st[q + 1] || from[q] >= st[q + 1]) return;
        PlaceSink k{ob + pl.bbase[q], off + 1 + pl.ebase[q], pl.bbase[q]};
        bool in_chunk = false, cl = false;
        fasta_scan(d, len, from[q], st[q + 1], k, in_chunk, cl);
    });
    if (pl.close_last) off[pl.n_ends] = pl.bbase[S];
    return (int64_t)pl.n_ends;
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

// thread-local message shared with the engine (tg_last_error)
void set_host_error(const std::string& m);

namespace {

// Cache size of a level from sysfs, as the reference's sysfs_cache_size
// (src/setup.hpp:16-46): cpu0 index entries, unified or data, K/M suffixes.
size_t cache_size(int level, size_t fallback)
{
    for (int idx = 0; idx <= 5; ++idx) {
        const std::string base = "/sys/devices/system/cpu/cpu0/cache/index" + std::to_string(idx);
        std::ifstream lf(base + "/level");
        int lv = 0;
        if (!lf || !(lf >> lv) || lv != level) continue;
        std::ifstream tf(base + "/type");
        std::string type;
        if (tf && (tf >> type) && type == "Instruction") continue;
        std::ifstream sf(base + "/size");
        std::string sz;
        if (!sf || !(sf >> sz) || sz.empty()) continue;
        char* end = nullptr;
        size_t v = strtoul(sz.c_str(), &end, 10);
        if (*end == 'K' || *end == 'k') v <<= 10;
        else if (*end == 'M' || *end == 'm') v <<= 20;
        if (v) return v;
    }
    return fallback;
}

void put_be(std::string& o, uint64_t v, int bytes)
{
    for (int i = bytes - 1; i >= 0; --i) o += (char)((v >> (8 * i)) & 0xFF);
}

}  // namespace

namespace tgi {

std::string kff_header(uint32_t k, uint64_t n)
{
    std::string h;
    h += "KFF";
    h += (char)1;                                   // version 1.0
    h += (char)0;
    h += (char)((0 << 6) | (1 << 4) | (2 << 2) | 3);  // A=0 C=1 G=2 T=3
    h += (char)1;                                   // unique
    h += (char)1;                                   // canonical
    put_be(h, 0, 4);                                // no metadata
    // global variables: k, max = 1, data_size = 4
    h += 'v';
    put_be(h, 3, 8);
    for (const auto& kv : {std::make_pair("k", (uint64_t)k), std::make_pair("max", (uint64_t)1),
                           std::make_pair("data_size", (uint64_t)4)}) {
        h += kv.first;
        h += '\0';
        put_be(h, kv.second, 8);
    }
    // raw section: one block per k-mer (max = 1: no per-block count)
    h += 'r';
    put_be(h, n, 8);
    return h;
}

std::string kff_footer()
{
    std::string t;
    t += 'v';
    put_be(t, 2, 8);
    t += "first_index";
    t += '\0';
    put_be(t, 0, 8);
    t += "footer_size";
    t += '\0';
    const uint64_t footer = 1 + 8 + 12 + 8 + 12 + 8 + 3;
    put_be(t, footer, 8);
    t += "KFF";
    return t;
}

}  // namespace tgi

extern "C" {

int tg_map_file(const char* path, const uint8_t** data, uint64_t* len)
{
    if (!path || !data || !len) return TG_EINVAL;
    *data = nullptr;
    *len = 0;
    const std::string p(path);
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        set_host_error("Cannot open file: " + p);
        return TG_EINPUT;
    }
    struct stat st {};
    if (fstat(fd, &st) != 0 || st.st_size == 0) {
        close(fd);
        set_host_error("Empty file: " + p);
        return TG_EINPUT;
    }
    const uint64_t n = (uint64_t)st.st_size;
    void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) {
        set_host_error("cannot map file: " + p);
        return TG_EINPUT;
    }
    (void)madvise(m, n, MADV_WILLNEED);
    *data = (const uint8_t*)m;
    *len = n;
    return TG_OK;
}

void tg_unmap_file(const uint8_t* data, uint64_t len)
{
    if (data && len) munmap((void*)data, len);
}

int tg_read_file(const char* path, uint8_t** data, uint64_t* len)
{
    if (!path || !data || !len) return TG_EINVAL;
    *data = nullptr;
    *len = 0;
    const std::string p(path);
    const bool gz = p.size() > 3 && p.compare(p.size() - 3, 3, ".gz") == 0;
    if (!gz) {
        FILE* f = fopen(path, "rb");
        if (!f) {
            set_host_error("Cannot open file: " + p);
            return TG_EINPUT;
        }
        struct stat st {};
        fstat(fileno(f), &st);
        const uint64_t n = (uint64_t)st.st_size;
        if (n == 0) {
            fclose(f);
            set_host_error("Empty file: " + p);
            return TG_EINPUT;
        }
        uint8_t* b = (uint8_t*)malloc(n);
        if (!b) {
            fclose(f);
            set_host_error("out of host memory reading " + p);
            return TG_ENOMEM;
        }
        const size_t got = fread(b, 1, n, f);
        fclose(f);
        if (got != n) {
            free(b);
            set_host_error("short read: " + p);
            return TG_EINPUT;
        }
        *data = b;
        *len = n;
        return TG_OK;
    }
    gzFile g = gzopen(path, "rb");
    if (!g) {
        set_host_error("Cannot open: " + p);
        return TG_EINPUT;
    }
    gzbuffer(g, 1u << 20);
    uint64_t cap = 1u << 24, n = 0;
    uint8_t* b = (uint8_t*)malloc(cap);
    for (;;) {
        if (n == cap) {
            cap *= 2;
            uint8_t* nb = (uint8_t*)realloc(b, cap);
            if (!nb) {
                free(b);
                gzclose(g);
                set_host_error("out of host memory reading " + p);
                return TG_ENOMEM;
            }
            b = nb;
        }
        const uint64_t want = std::min<uint64_t>(cap - n, 1u << 30);
        const int r = gzread(g, b + n, (unsigned)want);
        if (r <= 0) break;
        n += (uint64_t)r;
    }
    gzclose(g);
    if (n == 0) {
        free(b);
        set_host_error("Cannot read: " + p);
        return TG_EINPUT;
    }
    *data = b;
    *len = n;
    return TG_OK;
}

void tg_free(void* p) { free(p); }

// src/setup.hpp:82-196, restated: n = max(n_L3, n_throughput) capped by the
// fd limit and the writer-cache cap, floor 16.  Every host-dependent input is
// an argument here; tg_auto_partitions reads them from this host.
uint32_t tg_auto_partitions_with(const uint64_t* sizes, const uint8_t* gz, uint32_t n_files, uint32_t k,
                                 uint32_t threads, uint64_t l2_bytes, uint64_t l3_bytes, uint64_t fd_limit)
{
    const uint64_t max_parts = fd_limit > 32 ? fd_limit - 32 : 16;
    const uint64_t writer_bytes = 32;  // sizeof the reference's SuperkmerWriter
    const uint64_t ratio = l2_bytes / writer_bytes;
    uint64_t writer_cap = 4096;
    while ((writer_cap << 1) <= ratio) writer_cap <<= 1;
    writer_cap = std::min<uint64_t>(writer_cap, 32768);
    uint64_t total = 0, max_single = 0;
    for (uint32_t i = 0; i < n_files; ++i) {
        const uint64_t sz = sizes[i] * (gz && gz[i] ? 6u : 1u);
        total += sz;
        max_single = std::max(max_single, sz);
    }
    uint64_t est_bytes = total;
    if (n_files > 1 && total / n_files < (50ull << 20)) est_bytes = max_single;
    const uint64_t est_unique = est_bytes * 76 / 100;
    const uint64_t words = (k + 31) / 32;
    const uint64_t per_slot = words * 8 + 4 + 2;
    const uint64_t per_kmer = (per_slot * 5 + 3) / 4;
    const uint64_t l3_usable = std::max<uint64_t>(l3_bytes * 3 / 4, 1);
    const uint64_t num = (uint64_t)std::max<uint32_t>(threads, 1) * est_unique * per_kmer;
    const uint64_t n_l3_raw = (num + l3_usable - 1) / l3_usable;
    uint64_t n_l3 = 16;
    while (n_l3 < n_l3_raw) n_l3 <<= 1;
    const uint64_t n_tp_raw = std::max<uint64_t>(1, total >> 21);
    uint64_t n_tp = 16;
    while (n_tp < n_tp_raw) n_tp <<= 1;
    uint64_t n = std::max(n_l3, n_tp);
    n = std::min(n, std::min(max_parts, writer_cap));
    n = std::max<uint64_t>(n, 16);
    return (uint32_t)n;
}

uint32_t tg_auto_partitions(const char* const* files, uint32_t n_files, uint32_t k, uint32_t threads)
{
    struct rlimit rl {};
    uint64_t fd_limit = 1024;  // also when the limit is infinite, as the reference does
    if (getrlimit(RLIMIT_NOFILE, &rl) == 0 && rl.rlim_cur != RLIM_INFINITY) fd_limit = rl.rlim_cur;
    std::vector<uint64_t> sizes(n_files, 0);
    std::vector<uint8_t> gz(n_files, 0);
    for (uint32_t i = 0; i < n_files; ++i) {
        struct stat st {};
        const std::string f(files[i]);
        gz[i] = f.size() >= 3 && f.compare(f.size() - 3, 3, ".gz") == 0;
        if (stat(files[i], &st) == 0) sizes[i] = (uint64_t)st.st_size;  // a missing file adds nothing
    }
    return tg_auto_partitions_with(sizes.data(), gz.data(), n_files, k, threads, cache_size(2, 256u << 10),
                                   cache_size(3, 30u << 20), fd_limit);
}

int64_t tg_format_tsv(uint32_t k, const uint64_t* keys, const uint32_t* counts, uint64_t n, char* out,
                      uint64_t cap)
{
    static const char B[4] = {'A', 'C', 'G', 'T'};
    const uint32_t nw = (k + 31) / 32;  // Kmer<k> words
    uint64_t o = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (o + k + 12 > cap) return TG_ECAP;
        const uint64_t* w = keys + i * nw;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t bit = 2u * (k - 1u - j);
            out[o + j] = B[(w[bit >> 6] >> (bit & 63u)) & 3u];
        }
        o += k;
        out[o++] = '\t';
        char tmp[12];
        int t = 0;
        uint32_t v = counts[i];
        do {
            tmp[t++] = (char)('0' + v % 10u);
            v /= 10u;
        } while (v);
        while (t) out[o++] = tmp[--t];
        out[o++] = '\n';
    }
    return (int64_t)o;
}

int tg_write_kff(const char* path, uint32_t k, const uint64_t* keys, const uint32_t* counts, uint64_t n)
{
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_host_error(std::string("cannot open output file: ") + path);
        return TG_EINPUT;
    }
    const std::string h = tgi::kff_header(k, n);
    fwrite(h.data(), 1, h.size(), f);
    const uint32_t kb = (k + 3) / 4, nw = (k + 31) / 32;
    std::vector<uint8_t> buf;
    buf.reserve(1u << 20);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t* w = keys + i * nw;
        const size_t at = buf.size();
        buf.resize(at + kb + 4, 0);
        for (uint32_t j = 0; j < k; ++j) {  // first base in the top bits of byte 0
            const uint32_t bit = 2u * (k - 1u - j);
            buf[at + j / 4] |= (uint8_t)(((w[bit >> 6] >> (bit & 63u)) & 3u) << (6u - 2u * (j % 4u)));
        }
        const uint32_t c = counts[i];
        buf[at + kb] = (uint8_t)(c >> 24);
        buf[at + kb + 1] = (uint8_t)(c >> 16);
        buf[at + kb + 2] = (uint8_t)(c >> 8);
        buf[at + kb + 3] = (uint8_t)c;
        if (buf.size() >= (1u << 20)) {
            fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    if (!buf.empty()) fwrite(buf.data(), 1, buf.size(), f);
    const std::string t = tgi::kff_footer();
    fwrite(t.data(), 1, t.size(), f);
    const bool ok = ferror(f) == 0;
    fclose(f);
    if (!ok) {
        set_host_error(std::string("failed while writing output file: ") + path);
        return TG_EINPUT;
    }
    return TG_OK;
}

int64_t tg_parse_buffer(const uint8_t* data, uint64_t len, uint8_t* out_bases, uint64_t* out_off,
                        uint64_t off_cap)
{
    if (len == 0) return TG_EINPUT;
    if (off_cap < 1) return TG_ECAP;
    return data[0] == '@' ? parse_fastq(data, len, out_bases, out_off, off_cap)
                          : parse_fasta(data, len, out_bases, out_off, off_cap);
}

int64_t tg_parse_buffer_mt(const uint8_t* data, uint64_t len, uint8_t* out_bases, uint64_t* out_off,
                           uint64_t off_cap, uint32_t threads)
{
    if (len == 0) return TG_EINPUT;
    if (off_cap < 1) return TG_ECAP;
    const uint32_t S = (uint32_t)std::min<uint64_t>(std::max<uint32_t>(threads, 1u), std::max<uint64_t>(len / 64, 1));
    if (S <= 1) return tg_parse_buffer(data, len, out_bases, out_off, off_cap);
    return data[0] == '@' ? parse_fastq_mt(data, len, out_bases, out_off, off_cap, S)
                          : parse_fasta_mt(data, len, out_bases, out_off, off_cap, S);
}

// Workloads (all sizes in bases; seed selects the stream):
//   kind=genome_fa   len, seed, gc                     one FASTA record
//   kind=reads_fq    genome, seed, gc, reads|cov, rl, sub, nrate
//   kind=repeat_fq   genome, seed, gc, cov, rl, sub, fam, flen, div, rep, tan
//   kind=samples_fq  samples, seed, median, sigma, lo, hi, cov, rl, sub, index
//                    (the index-th sample of a log-normal family)
}  // extern "C"

namespace {

int64_t synth_into(const char* spec_c, std::string& out)
{
    const std::string spec = spec_c ? spec_c : "";
    const std::string kind = spec_str(spec, "kind", "genome_fa");
    const uint64_t seed = (uint64_t)spec_num(spec, "seed", 1);
    const double gc = spec_num(spec, "gc", 0.5);
    out.clear();
    Rng r(seed);
    std::string g;
    if (kind == "genome_fa") {
        random_genome(r, (uint64_t)spec_num(spec, "len", 4600000), gc, g);
        fasta("synthetic_seed" + std::to_string(seed), g, out);
    } else if (kind == "reads_fq" || kind == "repeat_fq") {
        const uint64_t G = (uint64_t)spec_num(spec, "genome", 5000000);
        if (kind == "reads_fq") random_genome(r, G, gc, g);
        else
            repeat_genome(r, G, gc, (uint32_t)spec_num(spec, "fam", 50), (uint32_t)spec_num(spec, "flen", 300),
                          spec_num(spec, "div", 0.05), spec_num(spec, "rep", 0.35), spec_num(spec, "tan", 0.10), g);
        const uint32_t rl = (uint32_t)spec_num(spec, "rl", 150);
        uint64_t nreads = (uint64_t)spec_num(spec, "reads", 0);
        if (!nreads) nreads = (uint64_t)std::llround(spec_num(spec, "cov", 50) * (double)G / rl);
        out.reserve(nreads * (2ull * rl + 16));
        Rng rr(seed * 0x100000001b3ull + 7);
        sample_reads_fq(rr, g, nreads, rl, spec_num(spec, "sub", 0.005), spec_num(spec, "nrate", 0.0), out, 0);
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
        sample_reads_fq(sr, g, nreads, rl, spec_num(spec, "sub", 0.005), spec_num(spec, "nrate", 0.0), out, 0);
    } else {
        return TG_EINVAL;
    }
    return (int64_t)out.size();
}

}  // namespace

extern "C" {

int64_t tg_synth(const char* spec_c) { return synth_into(spec_c, g_synth); }

int64_t tg_synth_alloc(const char* spec_c, uint8_t** data)
{
    if (!data) return TG_EINVAL;
    std::string t;
    const int64_t n = synth_into(spec_c, t);
    if (n < 0) return n;
    *data = static_cast<uint8_t*>(malloc(t.size() ? t.size() : 1));
    if (!*data) return TG_ENOMEM;
    memcpy(*data, t.data(), t.size());
    return n;
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

// ---- streamed input files ---------------------------------------------------
// The reference reads a gzip file through zlib's gzread, one decompressor
// feeding the parser (src/seq_source.hpp:48-120, :196-204).  Here a gzip file
// is inflated on one thread while the calling thread parses the text that has
// arrived, section by section, and hands each section's bases on (to the
// device upload) as soon as they are final; a BGZF file (gzip members that
// record their own sizes) is inflated on all threads at once; a plain file is
// mapped and parsed on all threads.  Every route gives the bases and chunk
// ends that tg_parse_buffer gives for the file's whole text.
namespace {

// A private anonymous mapping whose pages are committed on first touch: a
// buffer that grows in place up to its reserved size.
struct Reserved {
    uint8_t* p = nullptr;
    uint64_t cap = 0;
    bool reserve(uint64_t n)
    {
        void* q = mmap(nullptr, n, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (q == MAP_FAILED) return false;
        p = static_cast<uint8_t*>(q);
        cap = n;
        return true;
    }
    ~Reserved()
    {
        if (p) munmap(p, cap);
    }
};

// Bases appended to a reserved buffer, chunk ends to a vector.
struct GrowSink {
    uint8_t* ob;
    uint64_t cap;
    std::vector<uint64_t>& off;
    uint64_t nb = 0;
    bool full = false;
    bool bases(const uint8_t* p, uint64_t n)
    {
        if (nb + n > cap) {
            full = true;
            return false;
        }
        memcpy(ob + nb, p, n);
        nb += n;
        return true;
    }
    bool end()
    {
        off.push_back(nb);
        return true;
    }
};

inline uint32_t le32(const uint8_t* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

// BGZF (the blocked gzip of the SAM/BAM specification, section 4.1): every
// member has FEXTRA with a subfield 'B','C' of two bytes holding the member's
// size minus one, and ends with its uncompressed size (ISIZE).  Fills the
// member starts (and the file end) and returns true when the whole file is
// such members.
bool bgzf_members(const uint8_t* d, uint64_t len, std::vector<uint64_t>& starts)
{
    starts.clear();
    uint64_t p = 0;
    while (p < len) {
        if (len - p < 26 || d[p] != 0x1f || d[p + 1] != 0x8b || d[p + 2] != 8 || !(d[p + 3] & 4)) return false;
        const uint64_t xlen = d[p + 10] | (d[p + 11] << 8);
        uint64_t q = p + 12, bsize = 0;
        const uint64_t e = q + xlen;
        if (e > len) return false;
        while (q + 4 <= e) {
            const uint32_t slen = d[q + 2] | (d[q + 3] << 8);
            if (d[q] == 'B' && d[q + 1] == 'C' && slen == 2 && q + 6 <= e) bsize = (uint64_t)(d[q + 4] | (d[q + 5] << 8)) + 1;
            q += 4 + slen;
        }
        if (bsize < 26 || p + bsize > len) return false;
        starts.push_back(p);
        p += bsize;
    }
    starts.push_back(len);
    return starts.size() > 1;
}

// One gzip member into dst; true when it inflated to exactly `want` bytes.
bool inflate_member(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t want)
{
    z_stream z{};
    if (inflateInit2(&z, 16 + MAX_WBITS) != Z_OK) return false;
    z.next_in = const_cast<Bytef*>(src);
    z.avail_in = (uInt)n;
    z.next_out = dst;
    z.avail_out = (uInt)want;
    uint8_t spare;
    int r = inflate(&z, Z_FINISH);
    if (r == Z_BUF_ERROR && z.avail_out == 0) {  // output full: the member must end here
        z.next_out = &spare;
        z.avail_out = 1;
        r = inflate(&z, Z_FINISH);
    }
    const bool ok = r == Z_STREAM_END && z.total_out == want;
    inflateEnd(&z);
    return ok;
}

// Progress of the inflating thread.
struct Flow {
    uint64_t piece = 4u << 20;  // output per inflate call
    std::atomic<uint64_t> produced{0};
    std::atomic<int> state{0};  // 0 running, 1 done, 2 out of room
    std::atomic<bool> stop{false};
};

// Every gzip member of [src, src+n) into out, publishing progress.  As
// gzread does, a member may be followed by further members (concatenated
// files), bytes after the last member that do not start a member are ignored,
// and a truncated or damaged stream ends the text where it stops inflating.
void inflate_stream(const uint8_t* src, uint64_t n, Reserved& out, Flow& f)
{
    z_stream z{};
    if (inflateInit2(&z, 16 + MAX_WBITS) != Z_OK) {
        f.state.store(1);
        return;
    }
    uint64_t fed = 0, made = 0;
    int rc = 1;
    while (!f.stop.load(std::memory_order_relaxed)) {
        if (z.avail_in == 0 && fed < n) {
            const uint64_t take = std::min<uint64_t>(n - fed, 1u << 30);
            z.next_in = const_cast<Bytef*>(src + fed);
            z.avail_in = (uInt)take;
            fed += take;
        }
        const uint64_t room = std::min<uint64_t>(out.cap - made, f.piece);
        if (room == 0) {
            rc = 2;
            break;
        }
        z.next_out = out.p + made;
        z.avail_out = (uInt)room;
        const int r = inflate(&z, Z_NO_FLUSH);
        made += room - z.avail_out;
        f.produced.store(made, std::memory_order_release);
        if (r == Z_STREAM_END) {
            const uint64_t at = fed - z.avail_in;  // first byte after the member
            if (at + 1 < n && src[at] == 0x1f && src[at + 1] == 0x8b) {
                inflateReset(&z);
                continue;
            }
            break;
        }
        if (r == Z_OK) continue;
        if (r == Z_BUF_ERROR && (z.avail_in != 0 || fed < n)) continue;
        break;  // input ended inside a member, or damaged data
    }
    inflateEnd(&z);
    f.state.store(rc, std::memory_order_release);
}

// Waits until the text holds at least `want` more bytes than `have`, or the
// inflation ended; returns the text length then.
uint64_t wait_text(Flow& f, uint64_t have, uint64_t want, bool& final_)
{
    for (;;) {
        const int st = f.state.load(std::memory_order_acquire);
        if (st != 0) {
            final_ = true;
            return f.produced.load(std::memory_order_acquire);
        }
        const uint64_t n = f.produced.load(std::memory_order_acquire);
        if (n >= have + want) {
            final_ = false;
            return n;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
}

// One past the `back`-th last line feed in [from, end) (the last one counts
// 1); `from` when there are fewer.
uint64_t after_lf_back(const uint8_t* d, uint64_t from, uint64_t end, int back)
{
    uint64_t p = end;
    while (p > from) {
        const void* q = memrchr(d + from, '\n', (size_t)(p - from));
        if (!q) return from;
        p = (uint64_t)((const uint8_t*)q - d);
        if (--back == 0) return p + 1;
    }
    return from;
}

// Parses the growing text section by section.  A section ends where all
// output before it is final: FASTA just after a line feed (a header that
// started before it has ended there; a chunk running across it continues in
// the next section), FASTQ after the sixth-last line feed (each record that
// starts before it has its four lines, or five for the lexer's block-boundary
// case, within the text).  The last section ends with the text, so the
// whole-text rules (a DNA run reaching the end of a text whose length is a
// multiple of 64) apply as in tg_parse_buffer.
int parse_growing(const uint8_t* d, Flow& f, uint64_t section, GrowSink& sink, const tgi::SectionFn& ready,
                  uint64_t& total)
{
    uint64_t have = 0, done_to = 0, reported = 0, pos = 0;
    bool final_ = false, in_chunk = false, quirk = false, started = false;
    int fmt = 0;  // 1 FASTA, 2 FASTQ
    for (;;) {
        have = wait_text(f, have, section, final_);
        if (!fmt && have) fmt = d[0] == '@' ? 2 : 1;
        if (fmt) {
            const uint64_t stop = final_ ? have : after_lf_back(d, fmt == 1 ? done_to : pos, have, fmt == 1 ? 1 : 6);
            if (fmt == 1 && stop > done_to) {
                if (!started) {  // bytes before the first header are skipped
                    const void* q = memchr(d + done_to, '>', (size_t)(stop - done_to));
                    started = q != nullptr;
                    pos = started ? (uint64_t)((const uint8_t*)q - d) : stop;
                }
                bool closed = false;
                if (started && !fasta_scan(d, have, pos, stop, sink, in_chunk, closed)) return TG_ENOMEM;
                pos = stop;
                done_to = stop;
            } else if (fmt == 2 && stop > pos) {
                fastq_scan_at(d, have, pos, stop, sink, quirk);
                if (sink.full) return TG_ENOMEM;
            }
            if (final_ && fmt == 1 && in_chunk) sink.end();
            if (sink.nb > reported) {
                const int rc = ready(sink.ob + reported, sink.nb - reported, reported, false);
                if (rc) return rc;
                reported = sink.nb;
            }
        }
        if (final_) break;
    }
    total = have;
    return f.state.load() == 2 ? TG_ENOMEM : TG_OK;
}

// A whole text parsed on `threads` threads into `ob`; fills off.
int parse_whole(const uint8_t* d, uint64_t len, uint8_t* ob, uint32_t threads, std::vector<uint64_t>& off,
                const std::string& path)
{
    std::vector<uint64_t> o;
    for (uint64_t cap = len / 64 + 1024;;) {
        o.resize(cap);
        const int64_t n = tg_parse_buffer_mt(d, len, ob, o.data(), cap, threads);
        if (n == TG_ECAP && cap < len / 2 + 2) {
            cap = len / 2 + 2;
            continue;
        }
        if (n < 0) {
            set_host_error("cannot parse input: " + path);
            return TG_EINPUT;
        }
        off.assign(o.begin(), o.begin() + n + 1);
        return TG_OK;
    }
}

}  // namespace

namespace tgi {

int parse_file_sections(const char* path, uint32_t threads, uint64_t section, const SectionFn& ready,
                        std::vector<uint64_t>& off, uint64_t* n_bases)
{
    const std::string p(path);
    const bool gz = p.size() > 3 && p.compare(p.size() - 3, 3, ".gz") == 0;
    off.assign(1, 0);
    *n_bases = 0;
    threads = std::max(threads, 1u);
    if (!section) section = 32ull << 20;
    if (gz && access(path, R_OK) != 0) {
        set_host_error("Cannot open: " + p);  // gzopen's message (src/seq_source.hpp:90)
        return TG_EINPUT;
    }
    const uint8_t* data = nullptr;
    uint64_t len = 0;
    int rc = tg_map_file(path, &data, &len);
    if (rc) {
        if (gz && rc == TG_EINPUT) set_host_error("Cannot read: " + p);
        return rc;
    }
    struct Unmap {
        const uint8_t* d;
        uint64_t n;
        ~Unmap() { tg_unmap_file(d, n); }
    } unmap{data, len};
    // gzread reads a file that does not open with the gzip magic as it is
    const bool inflate = gz && len >= 2 && data[0] == 0x1f && data[1] == 0x8b;
    if (!inflate) {
        // parsed into page-locked memory when it can be had (one upload, no
        // staging), else into ordinary memory
        uint8_t* pin = nullptr;
        Reserved plain;
        if (tg_host_alloc(std::max<uint64_t>(len, 1), (void**)&pin) != TG_OK) {
            pin = nullptr;
            if (!plain.reserve(len + 64)) {
                set_host_error("out of host memory reading " + p);
                return TG_ENOMEM;
            }
        }
        struct Free {
            uint8_t* q;
            ~Free()
            {
                if (q) tg_host_free(q);
            }
        } fr{pin};
        uint8_t* ob = pin ? pin : plain.p;
        if ((rc = parse_whole(data, len, ob, threads, off, p))) return rc;
        *n_bases = off.back();
        return *n_bases ? ready(ob, *n_bases, 0, pin != nullptr) : TG_OK;
    }
    std::vector<uint64_t> members;
    if (bgzf_members(data, len, members)) {
        const size_t nm = members.size() - 1;
        std::vector<uint64_t> at(nm + 1, 0);
        for (size_t i = 0; i < nm; ++i) at[i + 1] = at[i] + le32(data + members[i + 1] - 4);
        const uint64_t tl = at[nm];
        Reserved text, bases;
        if (!text.reserve(tl + 64) || !bases.reserve(tl + 64)) {
            set_host_error("out of host memory reading " + p);
            return TG_ENOMEM;
        }
        std::atomic<size_t> next{0};
        std::atomic<size_t> bad{nm};
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < std::min<uint64_t>(threads, nm); ++t)
            th.emplace_back([&] {
                for (size_t i; (i = next.fetch_add(1)) < nm;)
                    if (!inflate_member(data + members[i], members[i + 1] - members[i], text.p + at[i], at[i + 1] - at[i])) {
                        size_t b = bad.load();
                        while (i < b && !bad.compare_exchange_weak(b, i)) {}
                    }
            });
        for (auto& t : th) t.join();
        // as gzread does, the text ends where a damaged member begins
        const uint64_t good = at[bad.load()];
        if (good == 0) {
            set_host_error("Cannot read: " + p);
            return TG_EINPUT;
        }
        if ((rc = parse_whole(text.p, good, bases.p, threads, off, p))) return rc;
        *n_bases = off.back();
        return *n_bases ? ready(bases.p, *n_bases, 0, false) : TG_OK;
    }
    // one thread inflating, this one parsing what has arrived
    Reserved text, bases;
    const uint64_t room = std::min<uint64_t>(std::max<uint64_t>(len * 1100, 1ull << 26), 1ull << 42);
    if (!text.reserve(room + 64) || !bases.reserve(room + 64)) {
        set_host_error("out of host memory reading " + p);
        return TG_ENOMEM;
    }
    Flow f;
    f.piece = std::min<uint64_t>(std::max<uint64_t>(section / 8, 16), 4u << 20);
    std::thread inflater([&] { inflate_stream(data, len, text, f); });
    GrowSink sink{bases.p, bases.cap, off};
    uint64_t total = 0;
    rc = parse_growing(text.p, f, section, sink, ready, total);
    f.stop.store(true);
    inflater.join();
    if (rc == TG_ENOMEM) set_host_error("out of host memory reading " + p);
    if (rc) return rc;
    if (total == 0) {
        set_host_error("Cannot read: " + p);
        return TG_EINPUT;
    }
    *n_bases = sink.nb;
    return TG_OK;
}

}  // namespace tgi

extern "C" {

int64_t tg_parse_file(const char* path, uint32_t threads, uint64_t section_bytes, uint8_t** bases,
                      uint64_t* n_bases, uint64_t** off)
{
    if (!path || !bases || !n_bases || !off) return TG_EINVAL;
    *bases = nullptr;
    *off = nullptr;
    *n_bases = 0;
    std::vector<uint64_t> o;
    std::vector<uint8_t> b;
    uint64_t n = 0;
    const int rc = tgi::parse_file_sections(path, threads, section_bytes,
                                            [&](const uint8_t* q, uint64_t cnt, uint64_t at, bool) {
                                                if (at != b.size()) return (int)TG_EINVAL;
                                                b.insert(b.end(), q, q + cnt);
                                                return (int)TG_OK;
                                            },
                                            o, &n);
    if (rc) return rc;
    *bases = (uint8_t*)malloc(std::max<uint64_t>(n, 1));
    *off = (uint64_t*)malloc(o.size() * 8);
    if (!*bases || !*off) {
        free(*bases);
        free(*off);
        *bases = nullptr;
        *off = nullptr;
        return TG_ENOMEM;
    }
    if (n) memcpy(*bases, b.data(), n);
    memcpy(*off, o.data(), o.size() * 8);
    *n_bases = n;
    return (int64_t)(o.size() - 1);
}

}  // extern "C"
