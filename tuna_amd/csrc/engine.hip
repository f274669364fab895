// engine.hip — host orchestration of the gfx950 counting path and its C ABI.
//
// One job = Phase 1 (pack, strips, superkmers, routing) + Phase 2 (counting
// work items, spill drain).  It replaces the in-memory pipeline of
// src/pipeline.hpp:111-178 (partition_kmers_mem + count_and_write_mem).
// Device buffers live in the context and are reused across jobs.
#include "engine_impl.h"
#include "host_io.h"
#include "tg_device.h"

#include <sys/stat.h>

#include <chrono>
#include <functional>
#include <unordered_map>

namespace tgi {
thread_local std::string g_err;
}

void set_host_error(const std::string& m) { tgi::g_err = m; }

using namespace tgi;

namespace {

// ---------------------------------------------------------------------------
struct Phase1 {
    uint32_t k = 0, m = 0, W = 0, S = 0, rw = 4;
    uint32_t n = 0, log2n = 0, nsp = 0, log2s = 0, sb = 0;
    uint64_t n_strips = 0, n_kmers = 0, n_sk = 0, n_rec = 0;
    uint32_t* packed = nullptr;
    uint32_t* sorted = nullptr;  // partition-ordered records (rw words each)
    uint64_t* part_off = nullptr;
    uint64_t* part_km = nullptr;
    uint32_t ntiles = 0;
    uint64_t n_flagged = 0;
};

// Phase 1: pack, strips, superkmers into bump-allocated runs, routing.
// slim = true keeps (start, meta) records for the export entry points; the
// counting path uses fat records that carry their bases.
int run_phase1(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t N,
               const uint64_t* d_off, uint64_t nch, Phase1& o, bool slim = false, uint32_t nsp_min = 0)
{
    hipStream_t st = c->stream;
    if (p->k > 64)
        return fail(TG_EINVAL, "k = %u: superkmer records, partition exports and debug statistics are built "
                               "for k <= 64 (k > 64 is counted directly)", p->k);
    o.k = p->k;
    o.m = p->m;
    o.W = p->k - p->m + 1;
    o.rw = slim ? 2u : tg::fat_rw(p->k);
    o.n = round_pow2(p->n_partitions ? p->n_partitions : 16u);
    if (o.n > (1u << tg::HSEL_BITS))
        return fail(TG_EINVAL, "n_partitions %u exceeds the %u partitions of the GPU engine", o.n,
                    1u << tg::HSEL_BITS);
    o.log2n = ilog2(o.n);
    // strip length: short chunks are one strip; long chunks are cut so that
    // the look-back (256-k k-mers) stays a small fraction
    o.S = 2048;

    // -- pack
    const uint64_t n_words = (N + 15) / 16 + 8;
    int rc = ensure_t(c, "packed", n_words, &o.packed);
    if (rc) return rc;
    {
        Prof pr(c, "pack");
        tg::launch_pack(st, d_bases, N, o.packed, n_words);
    }
    // -- strips
    uint64_t *nstr, *nkm, *nstr_x, *nkm_x, *tmp, *strip_long;
    // strips: at most N/S + one per chunk
    const uint64_t long_words = tg::fill_list_words(N / o.S + nch + 1);
    if ((rc = ensure_t(c, "nstr", nch + 1, &nstr)) || (rc = ensure_t(c, "nkm", nch + 1, &nkm)) ||
        (rc = ensure_t(c, "nstr_x", nch + 1, &nstr_x)) || (rc = ensure_t(c, "nkm_x", nch + 1, &nkm_x)) ||
        (rc = ensure_t(c, "scan_tmp", tg::scan_tmp_words(nch + 1) + 64, &tmp)) ||
        (rc = ensure_t(c, "strip_long", long_words, &strip_long)))
        return rc;
    {
        Prof pr(c, "strips");
        tg::launch_chunk_counts(st, d_off, nch, o.k, o.S, nstr, nkm, strip_long);
        tg::launch_scan_u64(st, nstr, nch, nstr_x, tmp);
        tg::launch_scan_u64(st, nkm, nch, nkm_x, tmp);
    }
    HIPOK(hipGetLastError());
    uint64_t tot[2];
    HIPOK(hipMemcpyAsync(&tot[0], nstr_x + nch, 8, hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(&tot[1], nkm_x + nch, 8, hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    o.n_strips = tot[0];
    o.n_kmers = tot[1];
    if (o.n_strips >= (1ull << 31)) return fail(TG_EINVAL, "too many strips");
    const uint64_t n_blocks = (o.n_strips + tg::SK_BLOCK - 1) / tg::SK_BLOCK;
    uint32_t *strip_chunk, *strip_tb, *flag_list, *flags, *run_start, *run_cnt;
    uint64_t *run_off, *strip_slot;
    unsigned long long* sk_count;
    const uint64_t ns1 = o.n_strips + 1;
    if ((rc = ensure_t(c, "strip_chunk", ns1, &strip_chunk)) ||
        (rc = ensure_t(c, "strip_tb", ns1, &strip_tb)) ||
        (rc = ensure_t(c, "strip_slot", ns1, &strip_slot)) ||
        (rc = ensure_t(c, "flag_list", ns1, &flag_list)) || (rc = ensure_t(c, "sk_flags", 64, &flags)) ||
        (rc = ensure_t(c, "sk_count", 8, &sk_count)))
        return rc;
    {
        Prof pr(c, "strips");
        tg::launch_fill_strips(st, nch, nstr_x, nkm_x, nkm, o.S, strip_chunk, strip_tb, strip_slot, strip_long,
                               long_words);
    }
    // record space: the slot space of the strips (a strip of K k-mers has at
    // most K records), then fix-up regions of SK_BLOCK*S records per block
    uint64_t fixb = tg::FIX_BLOCKS;
    for (;;) {
        const uint64_t fix_recs = fixb * tg::SK_BLOCK * o.S;
        const uint64_t rec_cap = o.n_kmers + fix_recs;
        const uint64_t run_cap = n_blocks + fixb + 1;
        if (rec_cap >= (1ull << 32)) return fail(TG_EINVAL, "input too large for one job's record space");
        uint32_t *recs, *metas;
        if ((rc = ensure_t(c, "recs", rec_cap * o.rw, &recs)) || (rc = ensure_t(c, "metas", rec_cap, &metas)) ||
            (rc = ensure_t(c, "run_start", run_cap, &run_start)) ||
            (rc = ensure_t(c, "run_cnt", run_cap, &run_cnt)) ||
            (rc = ensure_t(c, "run_off", run_cap + 1, &run_off)))
            return rc;
        {
            Prof pr(c, "strips");
            tg::launch_fill_u32(st, flags, 0u, 64);
            tg::launch_fill_u64(st, (uint64_t*)sk_count, 0ull, 1);
        }
        tg::SkArgs a{};
        a.packed = o.packed;
        a.off = d_off;
        a.strip_chunk = strip_chunk;
        a.strip_tb = strip_tb;
        a.strip_slot = strip_slot;
        a.list = nullptr;
        a.n = (uint32_t)o.n_strips;
        a.k = o.k;
        a.m = o.m;
        a.W = o.W;
        a.S = o.S;
        a.capk = 256u - o.k;
        a.piece = slim ? 0u : 16u * (o.rw - 1u) - o.k + 1u;
        a.lookback = a.capk + 1u;
        a.log2n = o.log2n;
        a.xb = tg::HSEL_BITS - o.log2n;
        a.err = flags + 1;
        a.recs = recs;
        a.metas = metas;
        a.fix_base = (uint32_t)o.n_kmers;
        a.run_start = run_start;
        a.run_cnt = run_cnt;
        a.run_base = 0;
        a.sk_count = sk_count;
        a.flag_list = flag_list;
        a.flag_count = flags;
        {
            Prof pr(c, "superkmers");
            tg::launch_superkmers(st, a, o.rw);
        }
        HIPOK(hipGetLastError());
        uint32_t nflag = 0;
        HIPOK(hipMemcpyAsync(&nflag, flags, 4, hipMemcpyDeviceToHost, st));
        HIPOK(hipStreamSynchronize(st));
        o.n_flagged = nflag;
        const uint64_t fb = (nflag + tg::SK_BLOCK - 1) / tg::SK_BLOCK;
        if (fb > fixb) {
            // more flagged strips than fix-up regions: size them and redo
            fixb = fb;
            continue;
        }
        if (nflag) {
            // strips whose first superkmer began before the look-back: redo
            // them from the chunk start (long low-complexity runs)
            a.list = flag_list;
            a.n = nflag;
            a.lookback = tg::LOOKBACK_FULL;
            a.run_base = (uint32_t)n_blocks;
            Prof pr(c, "superkmers_fixup");
            tg::launch_superkmers(st, a, o.rw);
        }
        HIPOK(hipGetLastError());
        const uint32_t n_runs = (uint32_t)(n_blocks + fb);
        uint64_t* run_cnt64;
        if ((rc = ensure_t(c, "run_cnt64", n_runs + 1, &run_cnt64)) ||
            (rc = ensure_t(c, "scan_tmp3", tg::scan_tmp_words(n_runs + 1) + 64, &tmp)))
            return rc;
        {
            Prof pr(c, "route_scan");
            tg::launch_u32_to_u64(st, run_cnt, n_runs, run_cnt64);
            tg::launch_scan_u64(st, run_cnt64, n_runs, run_off, tmp);
        }
        HIPOK(hipMemcpyAsync(&o.n_rec, run_off + n_runs, 8, hipMemcpyDeviceToHost, st));
        HIPOK(hipMemcpyAsync(&o.n_sk, sk_count, 8, hipMemcpyDeviceToHost, st));
        HIPOK(hipStreamSynchronize(st));
        // -- routing: tiles over the concatenated run space
        uint64_t tile = 16384;
        while ((o.n_rec + tile - 1) / tile > tg::ROUTE_MAX_TILES) tile <<= 1;
        o.ntiles = (uint32_t)((o.n_rec + tile - 1) / tile);
        // sub-partitions: ~16K k-mers each (N/16384), at least the reference n
        // and nsp_min, at least 2, at most 2^14 (LDS histograms).  With more
        // than 2^14 reference partitions a sub-partition holds several of them
        // (its top bits); the exports split them on the host.
        {
            uint64_t want = 2;
            while (want < o.n_kmers / 16384 && want < (1u << 14)) want <<= 1;
            want = std::max<uint64_t>({want, o.n, nsp_min});
            o.nsp = (uint32_t)std::min<uint64_t>(want, 1u << 14);
            o.log2s = o.nsp >= o.n ? ilog2(o.nsp) - o.log2n : 0;
            o.sb = tg::HSEL_BITS - ilog2(o.nsp);
        }
        uint32_t *hsk, *hkm, *route_tmp;
        uint64_t* part_sk;
        const uint64_t msz = (uint64_t)std::max<uint32_t>(o.ntiles, 1) * o.nsp;
        if ((rc = ensure_t(c, "hsk", msz, &hsk)) || (rc = ensure_t(c, "hkm", msz, &hkm)) ||
            (rc = ensure_t(c, "part_sk", o.nsp + 1, &part_sk)) ||
            (rc = ensure_t(c, "part_km", o.nsp + 1, &o.part_km)) ||
            (rc = ensure_t(c, "part_off", o.nsp + 1, &o.part_off)) ||
            (rc = ensure_t(c, "sorted", (o.n_rec + 1) * o.rw, &o.sorted)) ||
            (rc = ensure_t(c, "route_tmp", (o.n_rec + 1) * o.rw, &route_tmp)) ||
            (rc = ensure_t(c, "scan_tmp2", tg::scan_tmp_words(o.nsp + 1) + 64, &tmp)))
            return rc;
        tg::RouteArgs r{};
        r.recs = recs;
        r.metas = metas;
        r.run_start = run_start;
        r.run_cnt = run_cnt;
        r.run_off = run_off;
        r.n_runs = n_runs;
        r.n_rec = o.n_rec;
        r.tile = (uint32_t)tile;
        r.nsp = o.nsp;
        r.sb = o.sb;
        r.k = o.k;
        r.hsk = hsk;
        r.hkm = hkm;
        r.part_off = o.part_off;
        r.dst = o.sorted;
        r.tmp_meta = nullptr;  // (the plain sequence's items filter split buckets)
        r.split_off = nullptr;
        r.split_bits = tg::route_split_bits(o.nsp);
        r.split_km = nullptr;
        r.split_scale = nullptr;
        r.split_flags = nullptr;
        {
            Prof pr(c, "route_hist");
            tg::launch_route_hist(st, r, o.ntiles, o.rw);
        }
        {
            Prof pr(c, "route_scan");
            if (o.ntiles) tg::launch_route_colscan(st, r, o.ntiles, part_sk, o.part_km);
            else {
                tg::launch_fill_u64(st, part_sk, 0, o.nsp);
                tg::launch_fill_u64(st, o.part_km, 0, o.nsp);
            }
            tg::launch_scan_u64(st, part_sk, o.nsp, o.part_off, tmp);
        }
        {
            Prof pr(c, "route_scatter");
            tg::launch_route_scatter(st, r, o.ntiles, route_tmp, o.rw);
        }
        HIPOK(hipGetLastError());
        return TG_OK;
    }
}

// Work items over partitions [p_begin, p_end): consecutive sub-partitions
// merged while their estimated distinct k-mers (k-mers x rho) fit one LDS table
// at ~60% load; larger ones split by minimizer-hash bits.
// part_shift: with item_part, groups stop at reference partitions (sub-partition
// >> part_shift) and item_part gets each item's partition
void build_items(const Phase1& o, const std::vector<uint64_t>& km, const std::vector<uint64_t>& off,
                 uint32_t cap, double rho, const std::vector<uint8_t>* skip, uint32_t p_begin,
                 uint32_t p_end, uint32_t stride, std::vector<uint4>& items,
                 std::vector<uint32_t>* item_part = nullptr, uint32_t part_shift = 0)
{
    const double target = 0.75 * cap;
    const uint32_t rmax = 1u << o.sb;
    uint32_t p = p_begin;
    while (p < p_end) {
        if (skip && (*skip)[p]) {
            p += 1;
            continue;
        }
        const double est = (double)km[p] * rho;
        if (est > target) {
            uint32_t r = 1;
            while (r < rmax && est / r > target) r <<= 1;
            for (uint32_t j = 0; j < r; ++j) {
                items.push_back(make_uint4((uint32_t)off[p], (uint32_t)off[p + 1], r - 1, j));
                if (item_part) item_part->push_back(p >> part_shift);
            }
            p += stride;
            continue;
        }
        uint32_t q = p;
        double acc = 0;
        const uint32_t q_end = stride > 1 ? std::min(p + 1, p_end) : p_end;
        while (q < q_end && !(skip && (*skip)[q])) {
            const double e = (double)km[q] * rho;
            if (e > target || (q > p && acc + e > target)) break;
            if (item_part && (q >> part_shift) != (p >> part_shift)) break;
            acc += e;
            ++q;
        }
        if (q == p) q = p + 1;
        if (off[q] > off[p]) {
            items.push_back(make_uint4((uint32_t)off[p], (uint32_t)off[q], 0u, 0u));
            if (item_part) item_part->push_back(p >> part_shift);
        }
        p = stride > 1 ? p + stride : q;
    }
}

int launch_items(tg_ctx* c, const char* buf, const std::vector<uint4>& items, tg::CountArgs& a,
                 uint32_t nw, uint32_t** item_uniq, uint32_t** item_sp = nullptr, bool sample = false)
{
    if (items.empty()) return TG_OK;
    int rc;
    uint4* d_items;
    if ((rc = ensure_t(c, buf, items.size() + 1, &d_items))) return rc;
    HIPOK(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice,
                         c->stream));
    if (item_uniq) {
        if ((rc = ensure_t(c, "item_uniq", items.size() + 1, item_uniq))) return rc;
        a.item_uniq = *item_uniq;
    } else {
        a.item_uniq = nullptr;
    }
    a.item_sp = nullptr;
    if (item_sp) {
        if ((rc = ensure_t(c, "item_sp", items.size() + 1, item_sp))) return rc;
        a.item_sp = *item_sp;
    }
    a.items = d_items;
    a.n_items = (uint32_t)items.size();
    Prof pr(c, sample ? "count_sample" : "count");
    tg::launch_count(c->stream, a, nw);
    return TG_OK;
}

int run_phase2(tg_ctx* c, const tg_params* p, const Phase1& o, tg_stats* st_out)
{
    hipStream_t st = c->stream;
    const uint32_t nw = p->k > 32 ? 2u : 1u;
    const uint32_t cap = nw == 1 ? tg::LDS_CAP_NW1 : tg::LDS_CAP_NW2;
    std::vector<uint64_t> km(o.nsp), off(o.nsp + 1);
    HIPOK(hipMemcpyAsync(km.data(), o.part_km, o.nsp * 8ull, hipMemcpyDeviceToHost, st));
    HIPOK(hipMemcpyAsync(off.data(), o.part_off, (o.nsp + 1) * 8ull, hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    int rc;
    uint64_t* out_keys;
    uint32_t* out_cnt;
    unsigned long long* counters;
    uint32_t* sp_err;
    if ((rc = ensure_t(c, "out_keys", (o.n_kmers + 1) * nw, &out_keys)) ||
        (rc = ensure_t(c, "out_cnt", o.n_kmers + 1, &out_cnt)) ||
        (rc = ensure_t(c, "counters", 8, &counters)) || (rc = ensure_t(c, "sp_err", 8, &sp_err)))
        return rc;
    uint64_t sp_slots = 1ull << 16;
    while (sp_slots < o.n_kmers / 64) sp_slots <<= 1;
    for (int attempt = 0;; ++attempt) {
        uint64_t* sp_keys;
        uint32_t* sp_cnt;
        if ((rc = ensure_t(c, "sp_keys", sp_slots * nw, &sp_keys)) ||
            (rc = ensure_t(c, "sp_cnt", sp_slots, &sp_cnt)))
            return rc;
        tg::CountArgs a{};
        a.rec = o.sorted;
        a.k = p->k;
        a.ci = p->ci;
        a.cx = p->cx;
        a.probe_max = 64;
        a.per_len = p->k + (p->k - p->m + 1);
        a.sp_keys = sp_keys;
        a.sp_cnt = sp_cnt;
        a.sp_mask = sp_slots - 1;
        a.sp_err = sp_err;
        a.out_keys = out_keys;
        a.out_cnt = out_cnt;
        a.counters = counters;
        a.out_cap = o.n_kmers + 1;
        {
            Prof pr(c, "count_init");
            if (nw == 1) tg::launch_fill_u64(st, sp_keys, ~0ull, sp_slots);
            tg::launch_fill_u32(st, sp_cnt, 0u, sp_slots);
            tg::launch_fill_u64(st, (uint64_t*)counters, 0ull, 8);
            tg::launch_fill_u32(st, sp_err, 0u, 1);
        }
        // Calibration (after src/count.hpp:905-938, which sizes tables from the
        // first partition's distinct count): count a spread sample of
        // sub-partitions first, estimate distinct/k-mers, size the rest.
        std::vector<uint8_t> done(o.nsp, 0);
        std::vector<uint4> items;
        double rho = 1.0;
        const uint32_t stride = o.nsp >= 256 ? o.nsp / 64 : 0;
        // debug statistics per reference partition (src/count.hpp:42-60)
        const bool dbg = c->debug;
        if (dbg && o.n > o.nsp)
            return fail(TG_EINVAL, "debug statistics support at most %u partitions (-n %u)", o.nsp, o.n);
        uint32_t pshift = 0;
        while ((o.n << pshift) < o.nsp) ++pshift;
        std::vector<uint64_t> d_ins(dbg ? o.n : 0), d_uni(dbg ? o.n : 0), d_sp(dbg ? o.n : 0),
            d_cap(dbg ? o.n : 0);
        std::vector<uint32_t> ipart;
        auto dbg_take = [&](const std::vector<uint32_t>& uq, uint32_t* dev_sp) -> int {
            std::vector<uint32_t> sp(ipart.size());
            if (!sp.empty())
                HIPOK(hipMemcpyAsync(sp.data(), dev_sp, sp.size() * 4, hipMemcpyDeviceToHost, st));
            HIPOK(hipStreamSynchronize(st));
            for (size_t i = 0; i < ipart.size(); ++i) {
                d_uni[ipart[i]] += uq[i] - sp[i];
                d_sp[ipart[i]] += sp[i];
                d_cap[ipart[i]] += cap;
            }
            return TG_OK;
        };
        if (stride) {
            build_items(o, km, off, cap, 1.0, nullptr, 0, o.nsp, stride, items, dbg ? &ipart : nullptr, pshift);
            uint32_t *item_uniq = nullptr, *item_sp = nullptr;
            if ((rc = launch_items(c, "items_cal", items, a, nw, &item_uniq, dbg ? &item_sp : nullptr, true)))
                return rc;
            std::vector<uint32_t> uq(items.size());
            if (!items.empty())
                HIPOK(hipMemcpyAsync(uq.data(), item_uniq, items.size() * 4, hipMemcpyDeviceToHost, st));
            HIPOK(hipStreamSynchronize(st));
            if (dbg && (rc = dbg_take(uq, item_sp))) return rc;
            uint64_t su = 0, sk = 0;
            for (size_t i = 0; i < uq.size(); ++i) su += uq[i];
            for (uint32_t q = 0; q < o.nsp; q += stride) {
                sk += km[q];
                done[q] = 1;
            }
            if (sk) rho = std::min(1.0, std::max(0.02, 1.15 * (double)su / (double)sk));
        }
        items.clear();
        ipart.clear();
        build_items(o, km, off, cap, rho, stride ? &done : nullptr, 0, o.nsp, 1, items, dbg ? &ipart : nullptr,
                    pshift);
        uint32_t *main_uniq = nullptr, *main_sp = nullptr;
        if ((rc = launch_items(c, "items", items, a, nw, dbg ? &main_uniq : nullptr, dbg ? &main_sp : nullptr)))
            return rc;
        {
            Prof pr(c, "spill_drain");
            tg::launch_spill_drain(st, a, nw, sp_slots);
        }
        HIPOK(hipGetLastError());
        unsigned long long hc[4];
        uint32_t err = 0;
        HIPOK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
        HIPOK(hipMemcpyAsync(&err, sp_err, 4, hipMemcpyDeviceToHost, st));
        HIPOK(hipStreamSynchronize(st));
        if (err) {
            if (attempt >= 4) return fail(TG_ENOMEM, "spill table overflow after %d retries", attempt);
            sp_slots <<= 2;
            continue;
        }
        if (dbg) {
            std::vector<uint32_t> uq(items.size());
            if (!items.empty())
                HIPOK(hipMemcpyAsync(uq.data(), main_uniq, items.size() * 4, hipMemcpyDeviceToHost, st));
            if ((rc = dbg_take(uq, main_sp))) return rc;
            for (uint32_t q = 0; q < o.nsp; ++q) d_ins[q >> pshift] += km[q];
            c->dbg_inserted = d_ins;
            c->dbg_unique = d_uni;
            c->dbg_spilled = d_sp;
            c->dbg_cap = d_cap;
        }
        c->n_out = hc[0];
        c->nw = nw;
        c->d_keys = out_keys;
        c->d_cnt = out_cnt;
        if (st_out) {
            st_out->n_unique = hc[1];
            st_out->n_out = hc[0];
            st_out->n_lds_groups = (uint32_t)items.size();
            st_out->n_hbm_parts = (uint32_t)std::min<unsigned long long>(hc[2], 0xFFFFFFFFull);
        }
        break;
    }
    return TG_OK;
}

int count_device(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t N,
                 const uint64_t* d_off, uint64_t nch, tg_stats* st)
{
    int rc = check_params(p, N, nch);
    if (rc || (rc = check_pass(N, nch))) return rc;
    HIPOK(hipSetDevice(c->device));
    tg_stats s{};
    s.n_bases = N;
    s.n_chunks = nch;
    HIPOK(hipEventRecord(c->ev_a, c->stream));
    Phase1 o;
    if ((rc = run_phase1(c, p, d_bases, N, d_off, nch, o))) return rc;
    HIPOK(hipEventRecord(c->ev_b, c->stream));
    if ((rc = run_phase2(c, p, o, &s))) return rc;
    HIPOK(hipEventRecord(c->ev_c, c->stream));
    HIPOK(hipEventSynchronize(c->ev_c));
    float m1 = 0, m2 = 0;
    hipEventElapsedTime(&m1, c->ev_a, c->ev_b);
    hipEventElapsedTime(&m2, c->ev_b, c->ev_c);
    s.ms_phase1 = m1;
    s.ms_phase2 = m2;
    s.n_kmers = o.n_kmers;
    s.n_superkmers = o.n_sk;
    s.n_parts = o.n;
    s.n_subparts = o.nsp;
    collect_prof(c);
    if (st) *st = s;
    return TG_OK;
}

// Slices of a job: maximal runs of whole chunks of at most `limit` bases (a
// chunk longer than that is a slice of its own).  cuts[i]..cuts[i+1] are the
// chunks of slice i.
int plan_slices(const uint64_t* off, uint64_t nch, uint64_t limit, std::vector<uint64_t>& cuts)
{
    cuts.assign(1, 0);
    while (cuts.back() < nch) {
        const uint64_t c0 = cuts.back();
        // last chunk end within limit of the slice start
        uint64_t c1 = (uint64_t)(std::upper_bound(off + c0 + 1, off + nch + 1, off[c0] + limit) - off) - 1;
        if (c1 <= c0) {
            c1 = c0 + 1;
            if (off[c1] - off[c0] >= DEFAULT_SLICE_BASES)
                return fail(TG_EINVAL, "an ACGT run of %llu bases exceeds the %llu-base limit of one pass",
                            (unsigned long long)(off[c1] - off[c0]), (unsigned long long)DEFAULT_SLICE_BASES);
        }
        cuts.push_back(c1);
    }
    return TG_OK;
}

// Counts the slice of chunks [c0, c1) (bases [b0, b1)) with the parameters q
// (no count filter) into the context's result.
using SliceFn = std::function<int(uint64_t c0, uint64_t c1, uint64_t b0, uint64_t b1, const tg_params* q,
                                  tg_stats* ss)>;

// A job of more bases than the slice limit: every slice is counted as a job of
// its own with no count filter, its pairs are added into the device merge
// table, and the filtered table becomes the job's result.  Exact: the k-mers
// of a job are the k-mers of its chunks, and every chunk is in one slice.
// h_off: the chunk offsets on the host.
int count_slices(tg_ctx* c, const tg_params* p, uint64_t N, uint64_t nch, const uint64_t* h_off, uint64_t limit,
                 const SliceFn& count_slice, tg_stats* out)
{
    hipStream_t st = c->stream;
    if (h_off[nch] != N || h_off[0] != 0)
        return fail(TG_EINVAL, "chunk offsets span [%llu, %llu), not the %llu bases of the input",
                    (unsigned long long)h_off[0], (unsigned long long)h_off[nch], (unsigned long long)N);
    std::vector<uint64_t> cuts;
    int rc = plan_slices(h_off, nch, limit, cuts);
    if (rc) return rc;
    const uint32_t nw = p->k > 32 ? 2u : 1u;
    tg_params q = *p;
    q.ci = 1;
    q.cx = ~0ull;
    // the merge table: grown (rehashed) so that it stays at most half full
    Buf& tk = c->bufs["merge_keys"];
    Buf& tc = c->bufs["merge_cnt"];
    unsigned long long* claimed;
    uint32_t* err;
    if ((rc = ensure_t(c, "merge_ctr", 8, &claimed))) return rc;
    err = reinterpret_cast<uint32_t*>(claimed + 4);
    tg::launch_fill_u64(st, (uint64_t*)claimed, 0ull, 8);
    // a table left by an earlier job is released first (its memory is reused)
    if (tk.p) HIPOK(hipFree(tk.p));
    if (tc.p) HIPOK(hipFree(tc.p));
    tk = Buf{};
    tc = Buf{};
    uint64_t slots = 0, occ = 0;
    auto table = [&](uint64_t* k, uint32_t* n, uint64_t sl) {
        return tg::MergeTab{k, n, sl - 1, claimed, err};
    };
    tg_stats sum{};
    double ms_merge = 0;
    for (size_t i = 0; i + 1 < cuts.size(); ++i) {
        const uint64_t c0 = cuts[i], c1 = cuts[i + 1], b0 = h_off[c0], b1 = h_off[c1];
        tg_stats ss{};
        if ((rc = count_slice(c0, c1, b0, b1, &q, &ss))) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        const uint64_t n = c->n_out;
        uint64_t want = 1ull << 20;
        while (want < 2 * (occ + n)) want <<= 1;
        if (want > slots) {
            uint64_t* nk = nullptr;
            uint32_t* nc = nullptr;
            if (hipMalloc(&nk, want * nw * 8ull) != hipSuccess || hipMalloc(&nc, want * 4ull) != hipSuccess) {
                (void)hipGetLastError();
                if (nk) hipFree(nk);
                return fail(TG_ENOMEM, "merge table of %llu slots does not fit in device memory",
                            (unsigned long long)want);
            }
            {
                Prof pr(c, "merge");
                const tg::MergeTab fresh = table(nk, nc, want);
                tg::launch_merge_clear(st, fresh, nw);
                tg::launch_fill_u64(st, (uint64_t*)claimed, 0ull, 1);
                if (slots) tg::launch_merge_rehash(st, fresh, table((uint64_t*)tk.p, (uint32_t*)tc.p, slots), nw);
            }
            HIPOK(hipStreamSynchronize(st));
            if (tk.p) hipFree(tk.p);
            if (tc.p) hipFree(tc.p);
            tk.p = nk;
            tk.cap = want * nw * 8ull;
            tc.p = nc;
            tc.cap = want * 4ull;
            slots = want;
        }
        {
            Prof pr(c, "merge");
            tg::launch_merge_insert(st, table((uint64_t*)tk.p, (uint32_t*)tc.p, slots), nw, c->d_keys, c->d_cnt, n);
        }
        unsigned long long hc[5];
        HIPOK(hipMemcpyAsync(hc, claimed, sizeof(hc), hipMemcpyDeviceToHost, st));
        HIPOK(hipStreamSynchronize(st));
        if (reinterpret_cast<uint32_t*>(&hc[4])[0])
            return fail(TG_EDEVICE, "merge table overflow (%llu slots, %llu used)", (unsigned long long)slots,
                        (unsigned long long)hc[0]);
        occ = hc[0];
        ms_merge += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        sum.n_kmers += ss.n_kmers;
        sum.n_superkmers += ss.n_superkmers;
        sum.n_subparts += ss.n_subparts;
        sum.n_lds_groups += ss.n_lds_groups;
        sum.n_hbm_parts += ss.n_hbm_parts;
        sum.ms_phase1 += ss.ms_phase1;
        sum.ms_phase2 += ss.ms_phase2;
    }
    // the job's result: the table's keys whose sums pass ci/cx
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t* rk;
    uint32_t* rcnt;
    if ((rc = ensure_t(c, "sliced_keys", (occ + 1) * nw, &rk)) || (rc = ensure_t(c, "sliced_cnt", occ + 1, &rcnt)))
        return rc;
    uint64_t n_out = 0;
    if (slots) {
        Prof pr(c, "merge");
        tg::launch_fill_u64(st, (uint64_t*)claimed + 1, 0ull, 1);
        tg::launch_merge_drain(st, table((uint64_t*)tk.p, (uint32_t*)tc.p, slots), nw, p->ci, p->cx, rk, rcnt,
                               claimed + 1);
    }
    HIPOK(hipMemcpyAsync(&n_out, claimed + 1, 8, hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    ms_merge += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    drop_single_view(c);
    c->n_out = n_out;
    c->nw = nw;
    c->d_keys = rk;
    c->d_cnt = rcnt;
    collect_prof(c);
    sum.n_bases = N;
    sum.n_chunks = nch;
    sum.n_unique = occ;
    sum.n_out = n_out;
    sum.n_parts = round_pow2(p->n_partitions ? p->n_partitions : 16u);
    sum.ms_phase2 += ms_merge;
    if (out) *out = sum;
    return TG_OK;
}

// Slices of a device-resident input (h_off: its chunk offsets on the host, or
// nullptr to read them from d_off).
int count_sliced(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t N, const uint64_t* d_off,
                 uint64_t nch, const uint64_t* h_off, uint64_t limit, tg_stats* out)
{
    hipStream_t st = c->stream;
    std::vector<uint64_t> own;
    if (!h_off) {
        own.resize(nch + 1);
        HIPOK(hipMemcpyAsync(own.data(), d_off, (nch + 1) * 8ull, hipMemcpyDeviceToHost, st));
        HIPOK(hipStreamSynchronize(st));
        h_off = own.data();
    }
    auto slice = [&](uint64_t c0, uint64_t c1, uint64_t b0, uint64_t b1, const tg_params* q, tg_stats* ss) -> int {
        if (!c->plain) return count_range_batched(c, q, d_bases, N, d_off, nch, c0, c1, b0, b1, ss);
        // the plain sequence takes offsets relative to its bases
        std::vector<uint64_t> rel(c1 - c0 + 1);
        for (uint64_t j = c0; j <= c1; ++j) rel[j - c0] = h_off[j] - b0;
        uint64_t* d_rel;
        int rc;
        if ((rc = ensure_t(c, "slice_off", rel.size(), &d_rel))) return rc;
        HIPOK(hipMemcpyAsync(d_rel, rel.data(), rel.size() * 8ull, hipMemcpyHostToDevice, st));
        return count_device(c, q, d_bases + b0, b1 - b0, d_rel, c1 - c0, ss);
    };
    return count_slices(c, p, N, nch, h_off, limit, slice, out);
}

int upload_inputs(tg_ctx* c, const uint8_t* bases, uint64_t N, const uint64_t* off, uint64_t nch,
                  uint8_t** d_b, uint64_t** d_o);

// Slices of an input in host memory: each slice's bases and chunk offsets are
// uploaded into device staging and counted there, so the device holds one
// slice (and the merge table) however large the input is.
int count_host_sliced(tg_ctx* c, const tg_params* p, const uint8_t* h_bases, uint64_t N, const uint64_t* h_off,
                      uint64_t nch, uint64_t limit, tg_stats* out)
{
    hipStream_t st = c->stream;
    auto slice = [&](uint64_t c0, uint64_t c1, uint64_t b0, uint64_t b1, const tg_params* q, tg_stats* ss) -> int {
        std::vector<uint64_t> rel(c1 - c0 + 1);
        for (uint64_t j = c0; j <= c1; ++j) rel[j - c0] = h_off[j] - b0;
        uint8_t* d_b;
        uint64_t* d_o;
        int rc;
        if ((rc = upload_inputs(c, h_bases + b0, b1 - b0, rel.data(), c1 - c0, &d_b, &d_o))) return rc;
        HIPOK(hipStreamSynchronize(st));  // rel leaves scope
        if (c->plain) return count_device(c, q, d_b, b1 - b0, d_o, c1 - c0, ss);
        return count_one_batched(c, q, d_b, b1 - b0, d_o, c1 - c0, ss);
    };
    return count_slices(c, p, N, nch, h_off, limit, slice, out);
}

// k > 64: the superkmers are counted (main pass, then the flagged strips with
// full look-back), every k-mer goes into one open-addressed HBM table whose
// slots hold a key's position in the packed stream and its count (a power of
// two of at least 1.5 slots per k-mer, doubled while a key finds no slot and
// memory allows), and the table is drained through the ci/cx filter: per-block
// totals, their scan, then ceil(k/32) words per written key in slot order.
// Indices are 64-bit, so one pass takes the whole job.
int count_wide(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t N, const uint64_t* d_off,
               uint64_t nch, tg_stats* st)
{
    int rc = check_params(p, N, nch);
    if (rc) return rc;
    const uint32_t k = p->k, m = p->m, W = k - m + 1;
    if (W > tg::WIDE_WMAX) return fail(TG_EINVAL, "k - m + 1 = %u exceeds %u", W, tg::WIDE_WMAX);
    if (N >= (1ull << 40) - 1u) return fail(TG_EINVAL, "k > 64 counts inputs of fewer than 2^40 bases");
    HIPOK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const uint32_t nw = k <= 128 ? 4u : 8u, onw = (k + 31) / 32;
    tg_stats sum{};
    sum.n_bases = N;
    sum.n_chunks = nch;
    sum.n_parts = round_pow2(p->n_partitions ? p->n_partitions : 16u);
    HIPOK(hipEventRecord(c->ev_a, s));
    // pack and strips
    const uint32_t S = 2048;
    const uint64_t n_words = (N + 15) / 16 + 8;
    uint32_t* packed;
    const uint32_t SI = 256;  // strips of the insert (below)
    uint64_t *nstr, *nkm, *nstr_x, *nkm_x, *tmp, *strip_long;
    const uint64_t long_words = tg::fill_list_words(N / std::min(S, SI) + nch + 1);
    if ((rc = ensure_t(c, "packed", n_words, &packed)) || (rc = ensure_t(c, "nstr", nch + 1, &nstr)) ||
        (rc = ensure_t(c, "nkm", nch + 1, &nkm)) || (rc = ensure_t(c, "nstr_x", nch + 1, &nstr_x)) ||
        (rc = ensure_t(c, "nkm_x", nch + 1, &nkm_x)) ||
        (rc = ensure_t(c, "scan_tmp", tg::scan_tmp_words(nch + 1) + 64, &tmp)) ||
        (rc = ensure_t(c, "strip_long", long_words, &strip_long)))
        return rc;
    {
        Prof pr(c, "pack");
        tg::launch_pack(s, d_bases, N, packed, n_words);
    }
    {
        Prof pr(c, "strips");
        tg::launch_chunk_counts(s, d_off, nch, k, S, nstr, nkm, strip_long);
        tg::launch_scan_u64(s, nstr, nch, nstr_x, tmp);
        tg::launch_scan_u64(s, nkm, nch, nkm_x, tmp);
    }
    HIPOK(hipGetLastError());
    uint64_t tot[2];
    HIPOK(hipMemcpyAsync(&tot[0], nstr_x + nch, 8, hipMemcpyDeviceToHost, s));
    HIPOK(hipMemcpyAsync(&tot[1], nkm_x + nch, 8, hipMemcpyDeviceToHost, s));
    HIPOK(hipStreamSynchronize(s));
    const uint64_t n_strips = tot[0], n_kmers = tot[1];
    if (n_strips >= (1ull << 31)) return fail(TG_EINVAL, "too many strips");
    uint32_t *strip_chunk, *strip_tb, *flag_list, *ovf_list, *flags;
    uint64_t* strip_slot;
    unsigned long long* sk_count;
    const uint64_t ns1 = n_strips + 1;
    if ((rc = ensure_t(c, "strip_chunk", ns1, &strip_chunk)) || (rc = ensure_t(c, "strip_tb", ns1, &strip_tb)) ||
        (rc = ensure_t(c, "strip_slot", ns1, &strip_slot)) || (rc = ensure_t(c, "flag_list", ns1, &flag_list)) ||
        (rc = ensure_t(c, "ovf_list", ns1, &ovf_list)) || (rc = ensure_t(c, "sk_flags", 64, &flags)) ||
        (rc = ensure_t(c, "sk_count", 8, &sk_count)))
        return rc;
    {
        Prof pr(c, "strips");
        tg::launch_fill_strips(s, nch, nstr_x, nkm_x, nkm, S, strip_chunk, strip_tb, strip_slot, strip_long,
                               long_words);
        tg::launch_fill_u32(s, flags, 0u, 64);
        tg::launch_fill_u64(s, (uint64_t*)sk_count, 0ull, 1);
    }
    // superkmers, counted: the cap of the reference's header type
    // (2k - m <= 255: one byte, 255 bases; else two bytes, 65535 bases)
    const uint32_t hdr_max = 2u * k - m <= 255u ? 255u : 65535u;
    tg::SkArgs a{};
    a.packed = packed;
    a.off = d_off;
    a.strip_chunk = strip_chunk;
    a.strip_tb = strip_tb;
    a.strip_slot = strip_slot;
    a.n = (uint32_t)n_strips;
    a.k = k;
    a.m = m;
    a.W = W;
    a.S = S;
    a.capk = hdr_max + 1u - k;
    a.piece = 0;
    a.lookback = a.capk + 1u;
    a.err = flags;
    a.sk_count = sk_count;
    a.flag_list = flag_list;
    a.flag_count = flags + 1;
    a.dq_cap = c->wide_dq;
    a.ovf_list = ovf_list;
    a.ovf_count = flags + 3;
    // The superkmer pass only totals superkmers, and nothing in the table work
    // reads it: it runs on a second stream beside the insert (LDS-bound with
    // few waves per CU beside a latency-bound insert), unless kernels are timed
    hipStream_t ss = s;
    if (!c->profiling) {
        if (!c->side && (hipStreamCreate(&c->side) != hipSuccess || hipEventCreate(&c->ev_side) != hipSuccess))
            return fail(TG_EDEVICE, "HIP stream setup failed");
        ss = c->side;
        HIPOK(hipEventRecord(c->ev_side, s));
        HIPOK(hipStreamWaitEvent(ss, c->ev_side, 0));
    }
    {
        Prof pr(c, "superkmers");
        tg::launch_superkmers_wide(ss, a);
    }
    HIPOK(hipGetLastError());
    HIPOK(hipEventRecord(c->ev_b, ss));
    // every return from here on leaves the second stream idle (its kernel
    // reads buffers the next job writes)
    struct SideJoin {
        hipStream_t st;
        ~SideJoin() { hipStreamSynchronize(st); }
    } join{ss};
    // strips of the insert, in tables of their own: 256 k-mers, so that a long
    // chunk gives many lanes (a lane's k-mers are one chain of dependent table
    // probes)
    uint64_t *inst, *inst_x, *inkm;
    if ((rc = ensure_t(c, "inst", nch + 1, &inst)) || (rc = ensure_t(c, "inst_x", nch + 1, &inst_x)) ||
        (rc = ensure_t(c, "inkm", nch + 1, &inkm)))
        return rc;
    {
        Prof pr(c, "strips");
        tg::launch_chunk_counts(s, d_off, nch, k, SI, inst, inkm, strip_long);
        tg::launch_scan_u64(s, inst, nch, inst_x, tmp);
    }
    HIPOK(hipGetLastError());
    uint64_t n_ins = 0;
    HIPOK(hipMemcpyAsync(&n_ins, inst_x + nch, 8, hipMemcpyDeviceToHost, s));
    HIPOK(hipStreamSynchronize(s));
    if (n_ins >= (1ull << 31)) return fail(TG_EINVAL, "too many strips");
    uint32_t *istrip_chunk, *istrip_tb;
    uint64_t* istrip_slot;
    if ((rc = ensure_t(c, "istrip_chunk", n_ins + 1, &istrip_chunk)) ||
        (rc = ensure_t(c, "istrip_tb", n_ins + 1, &istrip_tb)) ||
        (rc = ensure_t(c, "istrip_slot", n_ins + 1, &istrip_slot)))
        return rc;
    {
        Prof pr(c, "strips");
        tg::launch_fill_strips(s, nch, inst_x, nkm_x, inkm, SI, istrip_chunk, istrip_tb, istrip_slot, strip_long,
                               long_words);
    }
    // the k-mer table: 16-byte slots {tag, count}, a power of two of at least
    // 1.5 slots per k-mer
    size_t free_b = 0, total_b = 0;
    HIPOK(hipMemGetInfo(&free_b, &total_b));
    constexpr uint64_t slot_bytes = 16;
    uint64_t slots = 1ull << 16;
    while (slots < n_kmers + n_kmers / 2) slots <<= 1;
    uint64_t* tab = nullptr;
    uint64_t *blk, *blk_x;
    for (int attempt = 0;; ++attempt) {
        // keep a quarter of the free memory for the result
        if (slots * slot_bytes > free_b / 4 * 3 + c->bufs["wide_tab"].cap)
            return fail(TG_ENOMEM, "the k-mer table for k = %u needs %llu slots of %llu bytes; device memory "
                                   "holds fewer", k, (unsigned long long)slots, (unsigned long long)slot_bytes);
        const uint32_t nb = tg::wide_drain_blocks(slots);
        if ((rc = ensure_t(c, "wide_tab", 2 * slots, &tab)) || (rc = ensure_t(c, "wide_blk", 2 * nb, &blk)) ||
            (rc = ensure_t(c, "wide_blk_x", 2 * (nb + 1), &blk_x)) ||
            (rc = ensure_t(c, "scan_tmp", tg::scan_tmp_words(nb + 1) + 64, &tmp)))
            return rc;
        tg::WideArgs w{};
        w.packed = packed;
        w.off = d_off;
        w.strip_chunk = istrip_chunk;
        w.strip_tb = istrip_tb;
        w.n_strips = (uint32_t)n_ins;
        w.n_dev = inst_x + nch;
        w.S = SI;
        w.k = k;
        w.tab = tab;
        w.mask = slots - 1;
        w.err = flags + 8;
        {
            Prof pr(c, "table_clear");
            tg::launch_fill_u64(s, tab, 0ull, 2 * slots);
            tg::launch_fill_u32(s, flags + 8, 0u, 1);
        }
        {
            Prof pr(c, "count");
            tg::launch_wide_insert(s, w, nw);
        }
        HIPOK(hipGetLastError());
        uint32_t err = 0;
        HIPOK(hipMemcpyAsync(&err, flags + 8, 4, hipMemcpyDeviceToHost, s));
        HIPOK(hipStreamSynchronize(s));
        if (err) {
            if (attempt >= 3) return fail(TG_ENOMEM, "k-mer table overflow after %d retries", attempt);
            slots <<= 1;
            continue;
        }
        // drain: per-block totals (written, distinct), their scans, the keys
        uint64_t hc[2] = {0, 0};
        {
            Prof pr(c, "drain");
            tg::launch_wide_drain_count(s, w, p->ci, p->cx, blk, blk + nb);
            tg::launch_scan_u64(s, blk, nb, blk_x, tmp);
            tg::launch_scan_u64(s, blk + nb, nb, blk_x + nb + 1, tmp);
        }
        HIPOK(hipGetLastError());
        HIPOK(hipMemcpyAsync(&hc[0], blk_x + nb, 8, hipMemcpyDeviceToHost, s));
        HIPOK(hipMemcpyAsync(&hc[1], blk_x + 2 * nb + 1, 8, hipMemcpyDeviceToHost, s));
        HIPOK(hipStreamSynchronize(s));
        uint64_t* out_keys;
        uint32_t* out_cnt;
        if ((rc = ensure_t(c, "wide_out_keys", (hc[0] + 1) * onw, &out_keys)) ||
            (rc = ensure_t(c, "wide_out_cnt", hc[0] + 1, &out_cnt)))
            return rc;
        {
            Prof pr(c, "job_gather");
            tg::launch_wide_drain_write(s, w, nw, p->ci, p->cx, onw, blk_x, out_keys, out_cnt);
        }
        HIPOK(hipGetLastError());
        // the superkmer total: after the main pass, the strips whose deque
        // overflowed counted again by the window kernel with the same
        // look-back (a look-back without a cut joins the flagged strips),
        // then the flagged strips counted again from their chunk start
        HIPOK(hipStreamWaitEvent(s, c->ev_b, 0));
        uint32_t nf[3] = {0, 0, 0};  // flagged, (unused), overflowed
        HIPOK(hipMemcpyAsync(nf, flags + 1, sizeof(nf), hipMemcpyDeviceToHost, s));
        HIPOK(hipStreamSynchronize(s));
        if (nf[2]) {
            a.list = ovf_list;
            a.n = nf[2];
            {
                Prof pr(c, "superkmers_fixup");
                tg::launch_superkmers_wide(s, a);
            }
            HIPOK(hipGetLastError());
            HIPOK(hipMemcpyAsync(nf, flags + 1, 4, hipMemcpyDeviceToHost, s));
            HIPOK(hipStreamSynchronize(s));
        }
        if (nf[0]) {
            a.list = flag_list;
            a.n = nf[0];
            a.lookback = tg::LOOKBACK_FULL;
            a.flag_count = flags + 2;  // nothing is flagged with full look-back
            Prof pr(c, "superkmers_fixup");
            tg::launch_superkmers_wide(s, a);
        }
        HIPOK(hipGetLastError());
        unsigned long long nsk = 0;
        HIPOK(hipMemcpyAsync(&nsk, sk_count, 8, hipMemcpyDeviceToHost, s));
        HIPOK(hipEventRecord(c->ev_c, s));
        HIPOK(hipEventSynchronize(c->ev_c));
        c->n_out = hc[0];
        c->nw = onw;
        c->d_keys = out_keys;
        c->d_cnt = out_cnt;
        sum.n_kmers = n_kmers;
        sum.n_superkmers = nsk;
        sum.n_unique = hc[1];
        sum.n_out = hc[0];
        break;
    }
    float m1 = 0, m2 = 0;
    hipEventElapsedTime(&m1, c->ev_a, c->ev_b);
    hipEventElapsedTime(&m2, c->ev_b, c->ev_c);
    sum.ms_phase1 = m1;
    sum.ms_phase2 = m2;
    collect_prof(c);
    if (st) *st = sum;
    return TG_OK;
}

// One job: in one pass, or in slices when it is larger than the slice limit.
int count_job(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t N, const uint64_t* d_off,
              uint64_t nch, const uint64_t* h_off, tg_stats* st)
{
    if (p && p->k > 64) return count_wide(c, p, d_bases, N, d_off, nch, st);
    const uint64_t limit = slice_limit(c);
    if (N > limit) {
        c->dbg_inserted.clear();
        c->dbg_unique.clear();
        c->dbg_spilled.clear();
        c->dbg_cap.clear();
        const bool dbg = c->debug;
        c->debug = false;  // per-partition statistics are kept for one-pass jobs only
        const int rc = count_sliced(c, p, d_bases, N, d_off, nch, h_off, limit, st);
        c->debug = dbg;
        return rc;
    }
    if (c->plain) return count_device(c, p, d_bases, N, d_off, nch, st);
    return count_one_batched(c, p, d_bases, N, d_off, nch, st);
}

int upload_inputs(tg_ctx* c, const uint8_t* bases, uint64_t N, const uint64_t* off, uint64_t nch,
                  uint8_t** d_b, uint64_t** d_o)
{
    int rc;
    if ((rc = ensure_t(c, "in_bases", N + 64, d_b)) || (rc = ensure_t(c, "in_off", nch + 1, d_o)))
        return rc;
    if (N) HIPOK(hipMemcpyAsync(*d_b, bases, N, hipMemcpyHostToDevice, c->stream));
    HIPOK(hipMemcpyAsync(*d_o, off, (nch + 1) * 8ull, hipMemcpyHostToDevice, c->stream));
    return TG_OK;
}

// A superkmer of a host input, as the exports need it: the global index of its
// first base and its slim meta word (length, minimizer offset, routing field).
struct SlimRec {
    uint64_t start;
    uint32_t meta;
};

// Phase 1 (slim records) over a host input, in slices of whole chunks when it
// is larger than the slice limit.  Every slice uses the same sub-partition
// count (that of the whole input), so routing fields mean the same in all.
int superkmers_all(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t N, const uint64_t* off,
                   uint64_t nch, std::vector<SlimRec>& out, Phase1& info)
{
    int rc;
    if (off[nch] != N || off[0] != 0)
        return fail(TG_EINVAL, "chunk offsets span [%llu, %llu), not the %llu bases of the input",
                    (unsigned long long)off[0], (unsigned long long)off[nch], (unsigned long long)N);
    std::vector<uint64_t> cuts;
    if ((rc = plan_slices(off, nch, slice_limit(c), cuts))) return rc;
    uint64_t kmers = 0;
    for (uint64_t i = 0; i < nch; ++i)
        if (off[i + 1] - off[i] >= p->k) kmers += off[i + 1] - off[i] - p->k + 1;
    uint64_t want = 2;
    while (want < kmers / 16384 && want < (1u << 14)) want <<= 1;
    out.clear();
    std::vector<uint64_t> rel;
    for (size_t i = 0; i + 1 < cuts.size(); ++i) {
        const uint64_t c0 = cuts[i], c1 = cuts[i + 1], b0 = off[c0];
        rel.resize(c1 - c0 + 1);
        for (uint64_t j = c0; j <= c1; ++j) rel[j - c0] = off[j] - b0;
        uint8_t* d_b;
        uint64_t* d_o;
        if ((rc = check_pass(off[c1] - b0, c1 - c0)) ||
            (rc = upload_inputs(c, bases + b0, off[c1] - b0, rel.data(), c1 - c0, &d_b, &d_o)))
            return rc;
        Phase1 o;
        if ((rc = run_phase1(c, p, d_b, off[c1] - b0, d_o, c1 - c0, o, true, (uint32_t)want))) return rc;
        std::vector<uint2> recs(o.n_sk);
        if (o.n_sk) HIPOK(hipMemcpy(recs.data(), o.sorted, o.n_sk * sizeof(uint2), hipMemcpyDeviceToHost));
        out.reserve(out.size() + recs.size());
        for (const uint2& r : recs) out.push_back(SlimRec{b0 + r.x, r.y});
        info = o;
    }
    info.n = round_pow2(p->n_partitions ? p->n_partitions : 16u);
    info.log2n = ilog2(info.n);
    if (cuts.size() < 2) {  // no chunks: no pass ran
        info.nsp = (uint32_t)std::max<uint64_t>(want, std::min<uint64_t>(info.n, 1u << 14));
        info.sb = tg::HSEL_BITS - ilog2(info.nsp);
    }
    info.n_sk = out.size();
    info.n_kmers = kmers;
    collect_prof(c);
    return TG_OK;
}

}  // namespace

// =================================================================== C ABI ==
extern "C" {

const char* tg_last_error(void) { return g_err.c_str(); }

int tg_ctx_create(int device, tg_ctx** out)
{
    if (!out) return fail(TG_EINVAL, "null output pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(TG_EDEVICE, "no HIP device available");
    }
    if (device < 0 || device >= n) return fail(TG_EINVAL, "device %d out of range (%d devices)", device, n);
    tg_ctx* c = new tg_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev_a) != hipSuccess || hipEventCreate(&c->ev_b) != hipSuccess ||
        hipEventCreate(&c->ev_c) != hipSuccess) {
        delete c;
        return fail(TG_EDEVICE, "HIP context setup failed");
    }
    noalloc_setup(device);
    *out = c;
    return TG_OK;
}

void tg_ctx_destroy(tg_ctx* c)
{
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    destroy_batch(c->batch);
    for (auto& kv : c->bufs)
        if (kv.second.p) hipFree(kv.second.p);
    for (auto& pe : c->pending) {
        hipEventDestroy(pe.second.first);
        hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    hipEventDestroy(c->ev_a);
    hipEventDestroy(c->ev_b);
    hipEventDestroy(c->ev_c);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
        hipEventDestroy(c->ev_side);
    }
    hipStreamDestroy(c->stream);
    delete c;
}

int tg_set_profiling(tg_ctx* c, int on)
{
    if (!c) return fail(TG_EINVAL, "null context");
    c->profiling = on != 0;
    return TG_OK;
}

void tg_reset_kernel_times(tg_ctx* c)
{
    if (c) c->ktimes.clear();
}

int tg_kernel_times(tg_ctx* c, char* names, uint64_t cap, double* ms, uint64_t* launches, int n)
{
    if (!c) return fail(TG_EINVAL, "null context");
    collect_prof(c);
    int i = 0;
    uint64_t used = 0;
    for (auto& kv : c->ktimes) {
        if (i < n) {
            if (ms) ms[i] = kv.second.ms;
            if (launches) launches[i] = kv.second.n;
            if (names) {
                const uint64_t L = kv.first.size() + 1;
                if (used + L <= cap) {
                    memcpy(names + used, kv.first.c_str(), L);
                    used += L;
                }
            }
        }
        ++i;
    }
    return i;
}

int tg_dev_alloc(tg_ctx* c, uint64_t bytes, void** dptr)
{
    if (!c || !dptr) return fail(TG_EINVAL, "null argument");
    HIPOK(hipSetDevice(c->device));
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(TG_ENOMEM, "hipMalloc of %llu bytes failed", (unsigned long long)bytes);
    }
    return TG_OK;
}

int tg_dev_free(tg_ctx* c, void* dptr)
{
    if (!c) return fail(TG_EINVAL, "null context");
    HIPOK(hipFree(dptr));
    return TG_OK;
}

int tg_h2d(tg_ctx* c, void* dst, const void* src, uint64_t bytes)
{
    if (!c) return fail(TG_EINVAL, "null context");
    HIPOK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_d2h(tg_ctx* c, void* dst, const void* src, uint64_t bytes)
{
    if (!c) return fail(TG_EINVAL, "null context");
    HIPOK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_sync(tg_ctx* c)
{
    if (!c) return fail(TG_EINVAL, "null context");
    HIPOK(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_count_device(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                    const uint64_t* d_off, uint64_t n_chunks, tg_stats* st)
{
    if (!c) return fail(TG_EINVAL, "null context");
    int rc = check_params(p, n_bases, n_chunks);
    if (rc) return rc;
    HIPOK(hipSetDevice(c->device));
    return count_job(c, p, d_bases, n_bases, d_off, n_chunks, nullptr, st);
}

int tg_set_plain_sequence(tg_ctx* c, int on)
{
    if (!c) return fail(TG_EINVAL, "null context");
    c->plain = on != 0;
    return TG_OK;
}

int tg_set_wide_deque(tg_ctx* c, uint32_t entries)
{
    if (!c) return fail(TG_EINVAL, "null context");
    if (entries < 1 || entries > tg::WDQ_MAX) return fail(TG_EINVAL, "deque entries must be in [1, %u]", tg::WDQ_MAX);
    c->wide_dq = entries;
    return TG_OK;
}

int tg_set_debug_stats(tg_ctx* c, int on)
{
    if (!c) return fail(TG_EINVAL, "null context");
    c->debug = on != 0;
    c->dbg_inserted.clear();
    c->dbg_unique.clear();
    c->dbg_spilled.clear();
    c->dbg_cap.clear();
    return TG_OK;
}

int64_t tg_debug_partitions(tg_ctx* c, uint64_t* inserted, uint64_t* lds_unique, uint64_t* spilled,
                            uint64_t* table_cap, uint64_t n)
{
    if (!c) return fail(TG_EINVAL, "null context");
    const uint64_t np = c->dbg_inserted.size();
    if (np == 0) return fail(TG_EINVAL, "no debug statistics: enable them and count with the plain sequence");
    if (n < np) return fail(TG_ECAP, "room for %llu partitions, %llu needed", (unsigned long long)n,
                            (unsigned long long)np);
    for (uint64_t i = 0; i < np; ++i) {
        if (inserted) inserted[i] = c->dbg_inserted[i];
        if (lds_unique) lds_unique[i] = c->dbg_unique[i];
        if (spilled) spilled[i] = c->dbg_spilled[i];
        if (table_cap) table_cap[i] = c->dbg_cap[i];
    }
    return (int64_t)np;
}

// A job whose input is in host memory: one that needs slices (k <= 64, more
// bases than the slice limit) is uploaded slice by slice; any other is
// uploaded whole.
int count_host_job(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t n_bases, const uint64_t* off,
                   uint64_t n_chunks, tg_stats* st)
{
    const uint64_t limit = slice_limit(c);
    if (p->k <= 64 && n_bases > limit) {
        c->dbg_inserted.clear();
        c->dbg_unique.clear();
        c->dbg_spilled.clear();
        c->dbg_cap.clear();
        const bool dbg = c->debug;
        c->debug = false;  // per-partition statistics are kept for one-pass jobs only
        const int rc = count_host_sliced(c, p, bases, n_bases, off, n_chunks, limit, st);
        c->debug = dbg;
        return rc;
    }
    uint8_t* d_b;
    uint64_t* d_o;
    int rc;
    if ((rc = upload_inputs(c, bases, n_bases, off, n_chunks, &d_b, &d_o))) return rc;
    return count_job(c, p, d_b, n_bases, d_o, n_chunks, off, st);
}

// Bytes of input a file count keeps on the device: the set budget, else half
// the free device memory.
uint64_t input_budget(tg_ctx* c)
{
    if (c->input_budget) return c->input_budget;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return free_b / 2;
}

int tg_count_host(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                  const uint64_t* off, uint64_t n_chunks, tg_stats* st)
{
    if (!c) return fail(TG_EINVAL, "null context");
    int rc = check_params(p, n_bases, n_chunks);
    if (rc) return rc;
    HIPOK(hipSetDevice(c->device));
    return count_host_job(c, p, bases, n_bases, off, n_chunks, st);
}

int tg_set_input_budget(tg_ctx* c, uint64_t bytes)
{
    if (!c) return fail(TG_EINVAL, "null context");
    c->input_budget = bytes;
    return TG_OK;
}

// Input files read into device memory as they are parsed: the bases of each
// section the host reader finalizes are staged through two pinned buffers
// and uploaded while the reader goes on, so inflating, parsing and the
// upload overlap; the count then runs on the device-resident bases.
// src/pipeline.hpp:64-178 (read, partition and count the input files).
int tg_count_files(tg_ctx* c, const tg_params* p, const char* const* paths, uint32_t n_files, uint32_t threads,
                   tg_stats* st)
{
    if (!c || !p || (!paths && n_files)) return fail(TG_EINVAL, "null argument");
    int rc = check_params(p, 0, 0);
    if (rc) return rc;
    if (n_files == 0) return fail(TG_EINVAL, "no input files");
    HIPOK(hipSetDevice(c->device));
    // a first guess of the bases, regrown (keeping what arrived) when short
    uint64_t guess = 64;
    for (uint32_t i = 0; i < n_files; ++i) {
        struct stat sb {};
        const std::string f(paths[i] ? paths[i] : "");
        const bool gz = f.size() > 3 && f.compare(f.size() - 3, 3, ".gz") == 0;
        if (stat(f.c_str(), &sb) == 0) guess += (uint64_t)sb.st_size * (gz ? 3u : 1u);
    }
    if (p->k <= 64 && guess > input_budget(c)) {
        // larger than the device keeps: the bases are gathered in host memory
        // and counted slice by slice
        std::vector<uint8_t> hb;
        std::vector<uint64_t> off{0}, fo;
        hb.reserve(guess);
        uint64_t base0 = 0;
        for (uint32_t i = 0; i < n_files; ++i) {
            auto ready = [&](const uint8_t* q, uint64_t n, uint64_t at, bool) -> int {
                const uint64_t g = base0 + at;
                if (hb.size() < g + n) hb.resize(g + n);
                memcpy(hb.data() + g, q, n);
                return TG_OK;
            };
            uint64_t nb = 0;
            if ((rc = tgi::parse_file_sections(paths[i], threads, 0, ready, fo, &nb))) return rc;
            off.reserve(off.size() + fo.size() - 1);
            for (size_t j = 1; j < fo.size(); ++j) off.push_back(base0 + fo[j]);
            base0 += nb;
        }
        const uint64_t nch = off.size() - 1;
        if ((rc = check_params(p, base0, nch))) return rc;
        return count_host_job(c, p, hb.data(), base0, off.data(), nch, st);
    }
    Buf& dst = c->bufs["file_bases"];
    auto room = [&](uint64_t need, uint64_t used) -> int {
        if (dst.cap >= need) return TG_OK;
        const uint64_t want = std::max<uint64_t>(need + need / 2, dst.cap + dst.cap / 2);
        void* q = nullptr;
        if (hipMalloc(&q, want) != hipSuccess) {
            (void)hipGetLastError();
            return fail(TG_ENOMEM, "device allocation of %llu bytes for the input bases failed",
                        (unsigned long long)want);
        }
        if (used) HIPOK(hipMemcpyAsync(q, dst.p, used, hipMemcpyDeviceToDevice, c->stream));
        HIPOK(hipStreamSynchronize(c->stream));
        if (dst.p) HIPOK(hipFree(dst.p));
        dst.p = q;
        dst.cap = want;
        return TG_OK;
    };
    if ((rc = room(guess, 0))) return rc;
    constexpr uint64_t STAGE = 32ull << 20;
    struct Stage {
        uint8_t* buf[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        hipStream_t s;
        ~Stage()
        {
            (void)hipStreamSynchronize(s);
            for (int i = 0; i < 2; ++i) {
                if (buf[i]) (void)hipHostFree(buf[i]);
                if (done[i]) (void)hipEventDestroy(done[i]);
            }
        }
    } stage;
    stage.s = c->stream;
    for (int i = 0; i < 2; ++i) {
        HIPOK(hipHostMalloc((void**)&stage.buf[i], STAGE, hipHostMallocDefault));
        HIPOK(hipEventCreateWithFlags(&stage.done[i], hipEventDisableTiming));
    }
    int cur = 0;
    uint64_t base0 = 0;
    std::vector<uint64_t> off{0}, fo;
    for (uint32_t i = 0; i < n_files; ++i) {
        auto ready = [&](const uint8_t* q, uint64_t n, uint64_t at, bool pinned) -> int {
            const uint64_t g = base0 + at;
            int r = room(g + n + 64, g);
            if (r) return r;
            uint8_t* d = static_cast<uint8_t*>(dst.p) + g;
            if (pinned) {  // page-locked already: one copy, done before the memory goes
                HIPOK(hipMemcpyAsync(d, q, n, hipMemcpyHostToDevice, c->stream));
                HIPOK(hipStreamSynchronize(c->stream));
                return TG_OK;
            }
            for (uint64_t done = 0; done < n;) {
                const uint64_t piece = std::min(n - done, STAGE);
                HIPOK(hipEventSynchronize(stage.done[cur]));
                memcpy(stage.buf[cur], q + done, piece);
                HIPOK(hipMemcpyAsync(d + done, stage.buf[cur], piece, hipMemcpyHostToDevice, c->stream));
                HIPOK(hipEventRecord(stage.done[cur], c->stream));
                cur ^= 1;
                done += piece;
            }
            return TG_OK;
        };
        uint64_t nb = 0;
        if ((rc = tgi::parse_file_sections(paths[i], threads, 0, ready, fo, &nb))) return rc;
        off.reserve(off.size() + fo.size() - 1);
        for (size_t j = 1; j < fo.size(); ++j) off.push_back(base0 + fo[j]);
        base0 += nb;
    }
    HIPOK(hipStreamSynchronize(c->stream));
    const uint64_t N = base0, nch = off.size() - 1;
    if ((rc = check_params(p, N, nch))) return rc;
    uint64_t* d_o;
    if ((rc = ensure_t(c, "in_off", nch + 1, &d_o))) return rc;
    HIPOK(hipMemcpyAsync(d_o, off.data(), (nch + 1) * 8ull, hipMemcpyHostToDevice, c->stream));
    return count_job(c, p, static_cast<const uint8_t*>(dst.p), N, d_o, nch, off.data(), st);
}

uint64_t tg_result_size(tg_ctx* c) { return c ? c->n_out : 0; }

int tg_result_copy(tg_ctx* c, uint64_t* keys, uint32_t* counts, uint64_t cap)
{
    if (!c) return fail(TG_EINVAL, "null context");
    const uint64_t n = std::min<uint64_t>(cap, c->n_out);
    if (n == 0) return TG_OK;
    if (keys) HIPOK(hipMemcpyAsync(keys, c->d_keys, n * c->nw * 8ull, hipMemcpyDeviceToHost, c->stream));
    if (counts) HIPOK(hipMemcpyAsync(counts, c->d_cnt, n * 4ull, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_pack_host(tg_ctx* c, const uint8_t* bases, uint64_t n_bases, uint32_t* out_words)
{
    if (!c || (!bases && n_bases) || !out_words) return fail(TG_EINVAL, "null argument");
    HIPOK(hipSetDevice(c->device));
    uint8_t* d_b;
    uint32_t* d_w;
    int rc;
    const uint64_t nw = (n_bases + 15) / 16;
    if ((rc = ensure_t(c, "in_bases", n_bases + 64, &d_b)) || (rc = ensure_t(c, "packed", nw + 8, &d_w)))
        return rc;
    if (n_bases) HIPOK(hipMemcpyAsync(d_b, bases, n_bases, hipMemcpyHostToDevice, c->stream));
    tg::launch_pack(c->stream, d_b, n_bases, d_w, nw);
    if (nw) HIPOK(hipMemcpyAsync(out_words, d_w, nw * 4, hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_superkmers_host(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                       const uint64_t* off, uint64_t n_chunks, uint64_t* start, uint32_t* len,
                       uint32_t* minpos, uint32_t* pid, uint64_t cap, uint64_t* n_sk,
                       uint32_t* n_subparts)
{
    if (!c) return fail(TG_EINVAL, "null context");
    int rc = check_params(p, n_bases, n_chunks);
    if (rc) return rc;
    HIPOK(hipSetDevice(c->device));
    std::vector<SlimRec> recs;
    Phase1 o;
    if ((rc = superkmers_all(c, p, bases, n_bases, off, n_chunks, recs, o))) return rc;
    if (n_sk) *n_sk = recs.size();
    if (n_subparts) *n_subparts = o.nsp;
    const uint64_t n = std::min<uint64_t>(recs.size(), cap);
    const uint32_t shift = tg::REC_HSEL_SHIFT + o.sb;
    for (uint64_t i = 0; i < n; ++i) {
        if (start) start[i] = recs[i].start;
        if (len) len[i] = recs[i].meta & tg::REC_LEN_MASK;
        if (minpos) minpos[i] = (recs[i].meta >> tg::REC_MINPOS_SHIFT) & 63u;
        if (pid) pid[i] = recs[i].meta >> shift;
    }
    return recs.size() > cap ? fail(TG_ECAP, "superkmer buffer too small (%llu needed)",
                                    (unsigned long long)recs.size())
                             : TG_OK;
}

int64_t tg_debug_min_coverage(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                              const uint64_t* off, uint64_t n_chunks, uint32_t* coverage,
                              uint64_t* n_minimizers, uint64_t cap)
{
    if (!c) return fail(TG_EINVAL, "null context");
    int rc = check_params(p, n_bases, n_chunks);
    if (rc) return rc;
    HIPOK(hipSetDevice(c->device));
    std::vector<SlimRec> recs;
    Phase1 o;
    if ((rc = superkmers_all(c, p, bases, n_bases, off, n_chunks, recs, o))) return rc;
    // k-mers per minimizer, keyed by the canonical ntHash of the m-mer at the
    // superkmer's minimizer position, as count_partition does under -dbg
    // (src/count.hpp:99-103,141-144 with Kmer_Window::init_packed_with_min,
    // include/kache-hash/Streaming_Kmer_Hash_Table.hpp:585-598)
    const uint32_t m = p->m;
    std::unordered_map<uint64_t, uint32_t> per_min;
    per_min.reserve(recs.size() / 2 + 16);
    for (const SlimRec& r : recs) {
        const uint32_t len = r.meta & tg::REC_LEN_MASK, mp = (r.meta >> tg::REC_MINPOS_SHIFT) & 63u;
        uint64_t f = 0, rv = 0;
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t ch = bases[r.start + mp + i];
            const uint32_t kc = ((ch >> 2) ^ (ch >> 1)) & 3u, nt = kc ^ (kc >> 1);
            f ^= tg::rotl64(tg::nt_seed(nt), m - 1u - i);
            rv ^= tg::rotl64(tg::nt_seed(nt ^ 2u), i);
        }
        per_min[f ^ rv] += len - p->k + 1u;
    }
    std::map<uint32_t, uint64_t> hist;
    for (const auto& e : per_min) ++hist[e.second];
    uint64_t i = 0;
    for (const auto& e : hist) {
        if (i < cap) {
            if (coverage) coverage[i] = e.first;
            if (n_minimizers) n_minimizers[i] = e.second;
        }
        ++i;
    }
    if (i > cap) return fail(TG_ECAP, "room for %llu coverage values, %llu needed", (unsigned long long)cap,
                             (unsigned long long)i);
    return (int64_t)i;
}

int tg_partitions_host(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                       const uint64_t* off, uint64_t n_chunks, uint64_t* part_off,
                       uint8_t* rec_bytes, uint64_t bytes_cap, uint64_t* rec_off,
                       uint64_t rec_cap, uint64_t* n_sk, uint64_t* n_bytes)
{
    if (!c) return fail(TG_EINVAL, "null context");
    int rc = check_params(p, n_bases, n_chunks);
    if (rc) return rc;
    HIPOK(hipSetDevice(c->device));
    std::vector<SlimRec> recs;
    Phase1 o;
    if ((rc = superkmers_all(c, p, bases, n_bases, off, n_chunks, recs, o))) return rc;
    if (n_sk) *n_sk = recs.size();
    // group by reference partition P = H & (n-1), the top log2(n) bits of the
    // routing field (counting sort; order inside a partition is unspecified,
    // as it is in the reference's partition files)
    // (the shift is 32 when n = 1: taken on 64 bits, where it gives 0)
    const uint32_t n = o.n, pshift = tg::REC_HSEL_SHIFT + (tg::HSEL_BITS - o.log2n);
    auto part = [&](const SlimRec& r) { return (uint64_t)r.meta >> pshift; };
    std::vector<uint64_t> cnt(n + 1, 0);
    for (const SlimRec& r : recs) ++cnt[part(r) + 1];
    for (uint32_t q = 0; q < n; ++q) cnt[q + 1] += cnt[q];
    if (part_off)
        for (uint32_t q = 0; q <= n; ++q) part_off[q] = cnt[q];
    std::vector<uint64_t> order(recs.size());
    for (uint64_t i = 0; i < recs.size(); ++i) order[cnt[part(recs[i])]++] = i;
    // serialise in the reference record format (src/superkmer_io.hpp:154-172);
    // hdr_t is one byte for every k <= 64 (2k-m <= 127)
    uint64_t used = 0;
    for (uint64_t i = 0; i < recs.size(); ++i) {
        const SlimRec& r = recs[order[i]];
        const uint32_t L = r.meta & tg::REC_LEN_MASK;
        const uint32_t mp = (r.meta >> tg::REC_MINPOS_SHIFT) & 63u;
        const uint64_t need = 2 + (L + 3) / 4;
        if (rec_off && i < rec_cap) rec_off[i] = used;
        if (rec_bytes && used + need <= bytes_cap) {
            uint8_t* d = rec_bytes + used;
            d[0] = (uint8_t)L;
            d[1] = (uint8_t)mp;
            memset(d + 2, 0, (L + 3) / 4);
            for (uint32_t j = 0; j < L; ++j) {
                const uint8_t ch = bases[r.start + j];
                d[2 + j / 4] |= (uint8_t)((((ch >> 2) ^ (ch >> 1)) & 3u) << (6u - 2u * (j % 4u)));
            }
        }
        used += need;
    }
    if (rec_off && recs.size() < rec_cap) rec_off[recs.size()] = used;
    if (n_bytes) *n_bytes = used;
    if (used > bytes_cap || recs.size() + 1 > rec_cap)
        return fail(TG_ECAP, "record buffers too small (%llu bytes, %llu records)",
                    (unsigned long long)used, (unsigned long long)recs.size());
    return TG_OK;
}

}  // extern "C"
