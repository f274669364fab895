// batch.hip — many independent counting jobs per launch sequence.
//
// The reference counts one sample per invocation (src/main.cpp:77 ->
// src/pipeline.hpp:111-178); counting many samples means one full pipeline per
// sample.  Here a batch concatenates the chunks of many jobs and runs ONE
// launch sequence over all of them: every job owns a private range of routing
// buckets, so records, work items, LDS tables and spill entries (tagged with
// the job) never mix two jobs, and the per-job results are gathered into job
// order on the device.  The sequence needs no host round trip: every size the
// host does not know (strips, flagged strips, records, work items) is read by
// the kernels from device memory, grids cover a capacity class, and the whole
// sequence is captured once per class as two hipGraphs (the two phases) and
// replayed per batch.
#include "engine_impl.h"

#include <chrono>

namespace tgi {

namespace {

constexpr uint32_t STRIP = 2048;            // k-mers per strip (as the single-job path)
constexpr uint32_t ITEM_CAP = 1u << tg::HSEL_BITS;  // items of a batch: buckets x splits
constexpr uint32_t MAX_BUCKETS = 1u << 14;  // LDS histogram bound of the routing kernels
constexpr uint32_t GATHER_BLOCKS = 1024;

uint64_t pow2_ceil64(uint64_t x)
{
    uint64_t r = 1;
    while (r < x) r <<= 1;
    return r;
}

// Everything a captured launch sequence bakes in.
struct BatchClass {
    uint32_t k = 0, m = 0, ci = 0, nw = 0;
    uint64_t cx = 0;
    uint64_t cap_bases = 0, cap_ch = 0, cap_strips = 0;
    uint32_t cap_jobs = 0, nsp = 0, xb = 0, max_tiles = 0, fixb = 0, rw = 4;
    uint64_t sp_slots = 0, rec_cap = 0;
    bool operator==(const BatchClass& o) const
    {
        return k == o.k && m == o.m && ci == o.ci && nw == o.nw && cx == o.cx &&
               cap_bases == o.cap_bases && cap_ch == o.cap_ch && cap_strips == o.cap_strips &&
               cap_jobs == o.cap_jobs && nsp == o.nsp && xb == o.xb && max_tiles == o.max_tiles &&
               fixb == o.fixb && sp_slots == o.sp_slots && rw == o.rw && rec_cap == o.rec_cap;
    }
    // every capacity of o fits in this class (same parameters)
    bool covers(const BatchClass& o) const
    {
        return k == o.k && m == o.m && ci == o.ci && nw == o.nw && cx == o.cx &&
               cap_bases >= o.cap_bases && cap_ch >= o.cap_ch && cap_strips >= o.cap_strips &&
               cap_jobs >= o.cap_jobs && nsp >= o.nsp && max_tiles == o.max_tiles &&
               fixb >= o.fixb && sp_slots >= o.sp_slots && rw == o.rw && rec_cap >= o.rec_cap;
    }
};

// Device buffers of one class (fixed addresses: the captured graph uses them).
struct BatchBufs {
    uint8_t* tab = nullptr;  // BatchDesc | job_chunk | job_route
    tg::BatchDesc* desc = nullptr;
    uint32_t* job_chunk = nullptr;
    uint32_t* job_route = nullptr;
    uint32_t* packed = nullptr;
    uint64_t cap_words = 0;
    uint64_t *boff, *nstr, *nkm, *nstr_x, *nkm_x, *scan_tmp;
    uint32_t *chunk_job, *strip_chunk, *strip_tb, *flag_list, *run_start, *run_cnt, *flags;
    uint64_t *strip_slot, *strip_long, *run_cnt64, *run_off;
    uint32_t *recs, *sorted, *route_tmp;  // rw words per record
    uint32_t* route_meta;                 // meta word of every coarse-ordered record
    uint64_t* split_off;                  // fine scatter's split-group starts
    uint32_t* split_flags;                // coarse buckets it orders by split group
    uint32_t* fine_hist;                  // group counts per record range of those buckets
    uint32_t* metas;                      // meta word per record slot
    uint32_t *hsk, *hkm, *bucket_job, *item_job;
    uint64_t *part_sk, *part_km, *part_off, *bucket_item;
    uint4* items;
    uint64_t *seg_pos, *seg_n, *seg_x;
    uint64_t* sp_keys;
    uint32_t* sp_cnt;
    unsigned long long* counters;
    uint8_t* back;             // the readback, packed for one copy
    uint64_t *out_keys, *fin_keys;
    uint32_t *out_cnt, *fin_cnt;
    unsigned long long* job_acc;  // [0,J): distinct, [J,2J): spill reported, [2J,3J): spill cursor,
                                  // [3J,4J): superkmers
    uint64_t *job_tot, *job_item_tot, *job_nsk, *job_nkm, *job_off;
    uint64_t n_blocks = 0, n_runs = 0;
};

// flags block (u32): [0] flagged strips, [1] phase-1 errors (bit 0 fix-up
// capacity, bit 1 partition-buffer capacity, bit 2 offsets), [2] spill
// overflow, [3] work items, [4] item queue
enum { F_FLAGS = 0, F_SKERR = 1, F_SPERR = 2, F_NITEMS = 3, F_ITEMQ = 4, F_WORDS = 16 };

// Readback of a batch (device, packed by tg::launch_pack_back, and its pinned
// host copy): job offsets, superkmers, k-mers, distinct, flags, counters.
inline uint64_t back_bytes(uint32_t cap_jobs) { return 8ull * (4ull * cap_jobs + 1) + 4ull * F_WORDS + 64; }

}  // namespace

// Device-resident results of a call: keys (nw words each) and counts, job j
// at [off[j], off[j] + len[j]).
struct ResultStore {
    uint64_t* keys = nullptr;
    uint32_t* cnt = nullptr;
    // a single job counted in one batch keeps its result where the batch
    // gathered it (valid until the next job on the context): no copy
    uint64_t* view_keys = nullptr;
    uint32_t* view_cnt = nullptr;
    uint64_t cap = 0, n = 0;
    uint32_t nw = 1;
    std::vector<uint64_t> off, len;
};

struct BatchState {
    BatchClass cls;
    bool have_cls = false;
    BatchBufs b{};
    std::vector<void*> allocs;
    // the launch sequence as two captured graphs, phase 1 (pack .. routing)
    // and phase 2 (work items .. readback); the phase-split event is
    // recorded on the stream between their launches, never inside a graph
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint8_t* h_tab = nullptr;  // pinned host copy of the per-batch tables
    uint8_t* h_back = nullptr; // pinned readback
    size_t h_tab_bytes = 0, h_back_bytes = 0;
    uint32_t count_blocks = 0;
    tg_batch_config cfg{};
    double rho = 0.3;          // distinct / k-mers, updated from every batch
    // results: of the last tg_count_many* call, and of the last single job
    ResultStore many, one;
    tg_many_stats last{};
    // device-time split of the last call (phase 1: pack..routing; phase 2:
    // counting and gathering), from events around and inside the sequence
    hipEvent_t ev_start = nullptr, ev_mid = nullptr, ev_end = nullptr;
    double ms_p1 = 0, ms_p2 = 0;
    uint64_t last_items = 0, last_spilled = 0;
};

void destroy_batch(BatchState* s);

namespace {

void free_class(BatchState* s)
{
    for (int i = 0; i < 2; ++i) {
        if (s->exec[i]) hipGraphExecDestroy(s->exec[i]);
        if (s->graph[i]) hipGraphDestroy(s->graph[i]);
        s->exec[i] = nullptr;
        s->graph[i] = nullptr;
    }
    for (void* p : s->allocs) hipFree(p);
    s->allocs.clear();
    s->b = BatchBufs{};
    s->have_cls = false;
}

template <class T>
int balloc(BatchState* s, T** p, uint64_t count)
{
    void* q = nullptr;
    const uint64_t bytes = std::max<uint64_t>(count * sizeof(T), 256);
    if (hipMalloc(&q, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(TG_ENOMEM, "device allocation of %llu bytes for a batch failed",
                    (unsigned long long)bytes);
    }
    s->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return TG_OK;
}

int alloc_class(BatchState* s, const BatchClass& c)
{
    free_class(s);
    BatchBufs& b = s->b;
    int rc;
    const uint64_t tab_bytes = 256 + 4ull * (c.cap_jobs + 1) + 4ull * c.cap_jobs;
    if ((rc = balloc(s, &b.tab, tab_bytes))) return rc;
    b.desc = reinterpret_cast<tg::BatchDesc*>(b.tab);
    b.job_chunk = reinterpret_cast<uint32_t*>(b.tab + 256);
    b.job_route = b.job_chunk + c.cap_jobs + 1;
    if (s->h_tab_bytes < tab_bytes) {
        if (s->h_tab) hipHostFree(s->h_tab);
        s->h_tab = nullptr;
        if (hipHostMalloc((void**)&s->h_tab, tab_bytes, hipHostMallocDefault) != hipSuccess)
            return fail(TG_ENOMEM, "pinned host allocation failed");
        s->h_tab_bytes = tab_bytes;
    }
    const uint64_t back_n = back_bytes(c.cap_jobs);
    if ((rc = balloc(s, &b.back, back_n))) return rc;
    if (s->h_back_bytes < back_n) {
        if (s->h_back) hipHostFree(s->h_back);
        s->h_back = nullptr;
        if (hipHostMalloc((void**)&s->h_back, back_n, hipHostMallocDefault) != hipSuccess)
            return fail(TG_ENOMEM, "pinned host allocation failed");
        s->h_back_bytes = back_n;
    }
    b.cap_words = c.cap_bases / 16 + 16;
    const uint64_t nch1 = c.cap_ch + 1, ns1 = c.cap_strips + 1;
    b.n_blocks = (c.cap_strips + tg::SK_BLOCK - 1) / tg::SK_BLOCK;
    b.n_runs = b.n_blocks + c.fixb;
    const uint64_t slots = c.cap_bases + (uint64_t)c.fixb * tg::SK_BLOCK * STRIP;
    const uint64_t scan_words =
        tg::scan_tmp_words(std::max<uint64_t>({nch1, b.n_runs + 1, c.nsp + 1ull, ITEM_CAP + 1ull,
                                               c.cap_jobs + 1ull})) + 64;
    const uint64_t outs = c.cap_bases + 1;
    if ((rc = balloc(s, &b.packed, b.cap_words)) || (rc = balloc(s, &b.boff, nch1)) ||
        (rc = balloc(s, &b.nstr, nch1)) || (rc = balloc(s, &b.nkm, nch1)) ||
        (rc = balloc(s, &b.nstr_x, nch1)) || (rc = balloc(s, &b.nkm_x, nch1)) ||
        (rc = balloc(s, &b.scan_tmp, scan_words)) || (rc = balloc(s, &b.chunk_job, nch1)) ||
        (rc = balloc(s, &b.strip_chunk, ns1)) || (rc = balloc(s, &b.strip_tb, ns1)) ||
        (rc = balloc(s, &b.strip_slot, ns1)) || (rc = balloc(s, &b.flag_list, ns1)) ||
        (rc = balloc(s, &b.strip_long, tg::fill_list_words(c.cap_strips))) ||
        (rc = balloc(s, &b.flags, F_WORDS)) || (rc = balloc(s, &b.run_start, b.n_runs + 1)) ||
        (rc = balloc(s, &b.run_cnt, b.n_runs + 1)) || (rc = balloc(s, &b.run_cnt64, b.n_runs + 1)) ||
        (rc = balloc(s, &b.run_off, b.n_runs + 2)) ||
        (rc = balloc(s, &b.recs, (slots + 1) * c.rw)) || (rc = balloc(s, &b.metas, slots + 1)) ||
        (rc = balloc(s, &b.sorted, (c.rec_cap + 1) * c.rw)) ||
        (rc = balloc(s, &b.route_tmp, (c.rec_cap + 1) * c.rw)) ||
        (rc = balloc(s, &b.route_meta, c.rec_cap + 1)) ||
        (rc = balloc(s, &b.split_off, (1ull << tg::HSEL_BITS) + 2)) ||
        (rc = balloc(s, &b.split_flags, 256)) ||
        (rc = balloc(s, &b.fine_hist, (uint64_t)tg::FINE_RANGES << tg::HSEL_BITS)) ||
        (rc = balloc(s, &b.hsk, (uint64_t)c.max_tiles * c.nsp)) ||
        (rc = balloc(s, &b.hkm, (uint64_t)c.max_tiles * c.nsp)) ||
        (rc = balloc(s, &b.part_sk, c.nsp + 1)) || (rc = balloc(s, &b.part_km, c.nsp + 1)) ||
        (rc = balloc(s, &b.part_off, c.nsp + 1)) || (rc = balloc(s, &b.bucket_job, c.nsp)) ||
        (rc = balloc(s, &b.bucket_item, c.nsp + 1)) || (rc = balloc(s, &b.items, ITEM_CAP)) ||
        (rc = balloc(s, &b.item_job, ITEM_CAP)) || (rc = balloc(s, &b.seg_pos, ITEM_CAP)) ||
        (rc = balloc(s, &b.seg_n, ITEM_CAP + 1)) || (rc = balloc(s, &b.seg_x, ITEM_CAP + 2)) ||
        (rc = balloc(s, &b.sp_keys, c.sp_slots * (c.nw + 1))) ||
        (rc = balloc(s, &b.sp_cnt, c.sp_slots)) || (rc = balloc(s, &b.counters, 8)) ||
        (rc = balloc(s, &b.out_keys, outs * c.nw)) || (rc = balloc(s, &b.out_cnt, outs)) ||
        (rc = balloc(s, &b.fin_keys, outs * c.nw)) || (rc = balloc(s, &b.fin_cnt, outs)) ||
        (rc = balloc(s, &b.job_acc, 4ull * c.cap_jobs)) ||
        (rc = balloc(s, &b.job_tot, c.cap_jobs + 1)) ||
        (rc = balloc(s, &b.job_item_tot, c.cap_jobs)) || (rc = balloc(s, &b.job_nsk, c.cap_jobs)) ||
        (rc = balloc(s, &b.job_nkm, c.cap_jobs)) || (rc = balloc(s, &b.job_off, c.cap_jobs + 2)))
        return rc;
    s->cls = c;
    s->have_cls = true;
    return TG_OK;
}

// The launch sequence of one batch (no host synchronisation inside).
// prof: the context whose per-kernel timing is on (nullptr while capturing);
// parts: bit 0 phase 1 (pack .. routing), bit 1 phase 2 (work items ..
// readback); with both, ev_mid is recorded between them.
void enqueue_batch(BatchState* s, hipStream_t st, tg_ctx* prof, int parts)
{
    const BatchClass& c = s->cls;
    BatchBufs& b = s->b;
    const tg::BatchJobs jobs{b.job_chunk, b.job_route};
    const uint32_t J = c.cap_jobs;
    if (parts & 1) {
    {
        // clear state (kernels only: the sequence may be captured) and inputs
        Prof pr(prof, "batch_init");
        tg::ZeroRegions z{};
        z.p[0] = b.flags;      z.n[0] = F_WORDS;          z.wide[0] = 0;
        z.p[1] = b.counters;   z.n[1] = 8;                z.wide[1] = 1;
        z.p[2] = b.sp_cnt;     z.n[2] = c.sp_slots;       z.wide[2] = 0;
        z.p[3] = b.seg_n;      z.n[3] = ITEM_CAP + 1;     z.wide[3] = 1;
        z.p[4] = b.job_acc;    z.n[4] = 4ull * J;         z.wide[4] = 1;
        z.count = 5;
        tg::launch_zero_regions(st, z);
        tg::launch_batch_setup(st, b.desc, jobs, J, b.boff, b.chunk_job, c.cap_ch, b.bucket_job, c.nsp,
                               b.flags + F_SKERR);
    }
    {
        Prof pr(prof, "pack");
        tg::launch_pack_desc(st, b.desc, b.packed, b.cap_words);
    }
    {
        Prof pr(prof, "strips");
        tg::launch_chunk_counts(st, b.boff, c.cap_ch, c.k, STRIP, b.nstr, b.nkm, b.strip_long);
        tg::launch_scan_u64(st, b.nstr, c.cap_ch, b.nstr_x, b.scan_tmp);
        tg::launch_scan_u64(st, b.nkm, c.cap_ch, b.nkm_x, b.scan_tmp);
        tg::launch_fill_strips(st, c.cap_ch, b.nstr_x, b.nkm_x, b.nkm, STRIP, b.strip_chunk, b.strip_tb,
                               b.strip_slot, b.strip_long, tg::fill_list_words(c.cap_strips));
    }
    // superkmers: main pass over the strip capacity, fix-up over FIXB blocks
    tg::SkArgs a{};
    a.packed = b.packed;
    a.off = b.boff;
    a.strip_chunk = b.strip_chunk;
    a.strip_tb = b.strip_tb;
    a.strip_slot = b.strip_slot;
    a.list = nullptr;
    a.n = (uint32_t)c.cap_strips;
    a.k = c.k;
    a.m = c.m;
    a.W = c.k - c.m + 1;
    a.S = STRIP;
    a.capk = 256u - c.k;
    a.piece = 16u * (c.rw - 1u) - c.k + 1u;
    a.lookback = a.capk + 1u;
    a.log2n = 0;
    a.xb = c.xb;
    a.chunk_job = b.chunk_job;
    a.job_route = b.job_route;
    a.n_dev = b.nstr_x + c.cap_ch;
    a.fix_dev = 0;
    a.err = b.flags + F_SKERR;
    a.recs = b.recs;
    a.metas = b.metas;
    a.fix_base = (uint32_t)c.cap_bases;
    a.run_start = b.run_start;
    a.run_cnt = b.run_cnt;
    a.run_base = 0;
    a.sk_count = b.job_acc + 3ull * J;
    a.flag_list = b.flag_list;
    a.flag_count = b.flags + F_FLAGS;
    {
        Prof pr(prof, "superkmers");
        tg::launch_superkmers(st, a, c.rw);
    }
    a.list = b.flag_list;
    a.n = c.fixb * tg::SK_BLOCK;
    a.n_dev = nullptr;
    a.fix_dev = 1;
    a.lookback = tg::LOOKBACK_FULL;
    a.run_base = (uint32_t)b.n_blocks;
    {
        Prof pr(prof, "superkmers_fixup");
        tg::launch_superkmers(st, a, c.rw);
    }
    // routing into job buckets
    tg::RouteArgs r{};
    r.recs = b.recs;
    r.metas = b.metas;
    r.run_start = b.run_start;
    r.run_cnt = b.run_cnt;
    r.run_off = b.run_off;
    r.n_runs = (uint32_t)b.n_runs;
    r.nsp = c.nsp;
    r.sb = c.xb;
    r.k = c.k;
    r.hsk = b.hsk;
    r.hkm = b.hkm;
    r.part_off = b.part_off;
    r.dst = b.sorted;
    r.n_rec_dev = b.run_off + b.n_runs;
    r.max_tiles = c.max_tiles;
    r.rec_cap = c.rec_cap;
    r.err = b.flags + F_SKERR;
    r.tmp_meta = b.route_meta;
    r.split_off = b.split_off;
    r.split_bits = tg::route_split_bits(c.nsp);
    r.split_km = b.part_km;
    r.split_scale = &b.desc->scale;
    r.split_flags = b.split_flags;
    r.fine_hist = b.fine_hist;
    {
        Prof pr(prof, "route_scan");
        tg::launch_u32_to_u64(st, b.run_cnt, b.n_runs, b.run_cnt64);
        tg::launch_scan_u64(st, b.run_cnt64, b.n_runs, b.run_off, b.scan_tmp);
    }
    {
        Prof pr(prof, "route_hist");
        tg::launch_route_hist(st, r, c.max_tiles, c.rw);
    }
    {
        Prof pr(prof, "route_scan");
        tg::launch_route_colscan(st, r, c.max_tiles, b.part_sk, b.part_km);
        tg::launch_scan_u64(st, b.part_sk, c.nsp, b.part_off, b.scan_tmp);
    }
    {
        Prof pr(prof, "route_scatter");
        tg::launch_route_scatter(st, r, c.max_tiles, b.route_tmp, c.rw);
    }
    }
    if (parts == 3) (void)hipEventRecord(s->ev_mid, st);
    if (!(parts & 2)) return;
    // work items and counting
    {
        Prof pr(prof, "count_items");
        tg::launch_build_items(st, b.part_km, b.part_off, c.nsp, c.xb, b.bucket_job, &b.desc->scale,
                               b.items, b.item_job, b.bucket_item, b.flags + F_NITEMS, b.part_sk,
                               b.scan_tmp, b.split_off, tg::route_split_bits(c.nsp));
    }
    tg::CountArgs ca{};
    ca.rec = b.sorted;
    ca.items = b.items;
    ca.k = c.k;
    ca.ci = c.ci;
    ca.cx = c.cx;
    ca.probe_max = 64;
    ca.per_len = c.k + (c.k - c.m + 1);
    ca.sp_keys = b.sp_keys;
    ca.sp_cnt = b.sp_cnt;
    ca.sp_mask = c.sp_slots - 1;
    ca.sp_err = b.flags + F_SPERR;
    ca.out_keys = b.out_keys;
    ca.out_cnt = b.out_cnt;
    ca.counters = b.counters;
    ca.out_cap = c.cap_bases + 1;
    ca.item_uniq = nullptr;
    ca.n_items_dev = b.flags + F_NITEMS;
    ca.item_q = b.flags + F_ITEMQ;
    ca.item_job = b.item_job;
    ca.seg_pos = b.seg_pos;
    ca.seg_n = b.seg_n;
    ca.job_uniq = b.job_acc;
    {
        Prof pr(prof, "count");
        tg::launch_count_batch(st, ca, c.nw, s->count_blocks);
    }
    {
        // per-job results in job order
        Prof pr(prof, "job_gather");
        tg::launch_spill_jobs(st, ca, c.nw, c.sp_slots, b.job_acc + J);
        tg::launch_scan_u64(st, b.seg_n, ITEM_CAP, b.seg_x, b.scan_tmp);
        tg::launch_job_totals(st, b.desc, jobs, J, b.bucket_item, b.seg_x, b.job_acc + 3ull * J,
                              b.job_acc + J, b.nkm_x, b.job_tot, b.job_item_tot, b.job_nsk, b.job_nkm);
        tg::launch_scan_u64(st, b.job_tot, J, b.job_off, b.scan_tmp);
        tg::launch_gather_items(st, ca, c.nw, jobs, b.desc, b.bucket_item, b.seg_x, b.job_off, b.fin_keys,
                                b.fin_cnt, GATHER_BLOCKS);
        tg::launch_spill_place(st, ca, b.desc, c.nw, c.sp_slots, b.job_off, b.job_item_tot,
                               b.job_acc + 2ull * J, b.fin_keys, b.fin_cnt);
        tg::launch_pack_back(st, b.desc, J, b.job_off, b.job_nsk, b.job_nkm, b.job_acc, b.flags, F_WORDS,
                             b.counters, b.back);
    }
}

int ensure_graph(BatchState* s, hipStream_t st)
{
    // kernel attributes are set before any capture (set_kernel_attributes)
    for (int part = 0; part < 2; ++part) {
        if (s->exec[part]) continue;
        HIPOK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_batch(s, st, nullptr, part + 1);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (e == hipSuccess) e = hipGraphInstantiate(&s->exec[part], g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (g) hipGraphDestroy(g);
            s->exec[part] = nullptr;
            return fail(TG_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
        }
        s->graph[part] = g;
        s->last.n_captures += part == 0;
    }
    return TG_OK;
}

int grow_store(ResultStore& r, hipStream_t st, uint64_t need, uint32_t nw)
{
    if (r.cap >= need && r.nw == nw) return TG_OK;
    // grown geometrically while the key width stays; a new width keeps
    // nothing, so it is sized from the need alone
    const uint64_t grown = r.nw == nw ? r.cap + r.cap / 2 : 0;
    uint64_t cap = std::max<uint64_t>(need, grown);
    cap = std::max<uint64_t>(cap, 1u << 20);
    uint64_t* k2 = nullptr;
    uint32_t* c2 = nullptr;
    if (hipMalloc(&k2, cap * nw * 8ull) != hipSuccess || hipMalloc(&c2, cap * 4ull) != hipSuccess) {
        (void)hipGetLastError();
        if (k2) hipFree(k2);
        return fail(TG_ENOMEM, "result store of %llu k-mers does not fit in device memory",
                    (unsigned long long)cap);
    }
    if (r.n && r.nw == nw) {
        HIPOK(hipMemcpyAsync(k2, r.keys, r.n * nw * 8ull, hipMemcpyDeviceToDevice, st));
        HIPOK(hipMemcpyAsync(c2, r.cnt, r.n * 4ull, hipMemcpyDeviceToDevice, st));
    }
    HIPOK(hipStreamSynchronize(st));
    if (r.keys) hipFree(r.keys);
    if (r.cnt) hipFree(r.cnt);
    r.keys = k2;
    r.cnt = c2;
    r.cap = cap;
    r.nw = nw;
    return TG_OK;
}

void free_store(ResultStore& r)
{
    if (r.keys) hipFree(r.keys);
    if (r.cnt) hipFree(r.cnt);
    r = ResultStore{};
}

struct JobTab {
    const uint64_t* job_chunk;
    const uint64_t* job_base;
    uint32_t n;
};

uint32_t job_buckets(uint64_t bases, uint32_t kb)
{
    const uint64_t want = std::max<uint64_t>(1, (bases + kb - 1) / kb);
    return (uint32_t)std::min<uint64_t>(pow2_ceil64(want), MAX_BUCKETS);
}

// Greedy batches in job order: bases within the configured budget (a larger
// job goes alone), buckets within MAX_BUCKETS.
void plan_batches(const JobTab& jt, const tg_batch_config& cfg, std::vector<uint32_t>& cuts)
{
    cuts.clear();
    cuts.push_back(0);
    uint64_t bases = 0;
    uint32_t buckets = 0;
    for (uint32_t j = 0; j < jt.n; ++j) {
        const uint64_t nb = jt.job_base[j + 1] - jt.job_base[j];
        const uint32_t nbk = job_buckets(nb, cfg.kmers_per_bucket);
        const bool empty = j == cuts.back();
        if (!empty && (bases + nb > cfg.batch_bases || buckets + nbk > MAX_BUCKETS)) {
            cuts.push_back(j);
            bases = 0;
            buckets = 0;
        }
        bases += nb;
        buckets += nbk;
    }
    if (cuts.back() != jt.n) cuts.push_back(jt.n);
}

BatchClass class_for(const tg_params* p, const JobTab& jt, const std::vector<uint32_t>& cuts,
                     const tg_batch_config& cfg)
{
    BatchClass c;
    c.k = p->k;
    c.m = p->m;
    c.ci = p->ci;
    c.cx = p->cx;
    c.nw = p->k > 32 ? 2u : 1u;
    c.max_tiles = cfg.max_tiles;
    c.fixb = tg::FIX_BLOCKS;
    uint64_t bases = 0, ch = 0;
    uint32_t jobs = 0, buckets = 0;
    for (size_t i = 0; i + 1 < cuts.size(); ++i) {
        const uint32_t j0 = cuts[i], j1 = cuts[i + 1];
        bases = std::max(bases, jt.job_base[j1] - jt.job_base[j0]);
        ch = std::max(ch, jt.job_chunk[j1] - jt.job_chunk[j0]);
        jobs = std::max(jobs, j1 - j0);
        uint32_t bk = 0;
        for (uint32_t j = j0; j < j1; ++j)
            bk += job_buckets(jt.job_base[j + 1] - jt.job_base[j], cfg.kmers_per_bucket);
        buckets = std::max(buckets, bk);
    }
    c.cap_bases = pow2_ceil64(std::max<uint64_t>(bases, 1u << 16));
    c.cap_ch = pow2_ceil64(std::max<uint64_t>(ch, 1024));
    c.cap_strips = c.cap_ch + c.cap_bases / STRIP + 1;
    c.cap_jobs = (uint32_t)pow2_ceil64(std::max<uint32_t>(jobs, 64));
    c.nsp = (uint32_t)pow2_ceil64(std::max<uint32_t>(buckets, 2));
    c.xb = tg::HSEL_BITS - ilog2(c.nsp);
    c.sp_slots = pow2_ceil64(std::max<uint64_t>(c.cap_bases / 64, 1u << 16));
    c.rw = tg::fat_rw(c.k);
    // partition buffers: a typical bound (a superkmer every (W+1)/4 k-mers); a
    // batch that overflows it re-runs at the worst case (one per k-mer)
    c.rec_cap = std::min<uint64_t>(c.cap_bases, c.cap_bases * 4 / (c.k - c.m + 2) + 2 * c.cap_strips + 65536);
    return c;
}

}  // namespace

void destroy_batch(BatchState* s)
{
    if (!s) return;
    free_class(s);
    if (s->ev_start) hipEventDestroy(s->ev_start);
    if (s->ev_mid) hipEventDestroy(s->ev_mid);
    if (s->ev_end) hipEventDestroy(s->ev_end);
    if (s->h_tab) hipHostFree(s->h_tab);
    if (s->h_back) hipHostFree(s->h_back);
    free_store(s->many);
    free_store(s->one);
    delete s;
}

namespace {

BatchState* state(tg_ctx* c)
{
    if (!c->batch) {
        c->batch = new BatchState();
        tg_batch_config& g = c->batch->cfg;
        g.kmers_per_bucket = 16384;
        g.max_tiles = tg::ROUTE_MAX_TILES;
        g.batch_bases = 1ull << 28;
        g.load_factor = 0.75f;
        g.use_graph = 1;
        g.rho = 0.0f;
        g.slice_bases = 0;
        BatchState* s = c->batch;
        if (hipEventCreate(&s->ev_start) != hipSuccess || hipEventCreate(&s->ev_mid) != hipSuccess ||
            hipEventCreate(&s->ev_end) != hipSuccess)
            (void)hipGetLastError();
    }
    return c->batch;
}

// A single job's result that still lives in the batch buffers (view mode)
// moves into its own store, and the context's result pointers follow it.
// Called before anything may free or reuse those buffers.
int migrate_view(tg_ctx* ctx, BatchState* s)
{
    ResultStore& o = s->one;
    if (!o.view_keys) return TG_OK;
    hipStream_t st = ctx->stream;
    uint64_t* vk = o.view_keys;
    uint32_t* vc = o.view_cnt;
    o.view_keys = nullptr;
    o.view_cnt = nullptr;
    const uint64_t n = o.n;
    o.n = 0;
    int rc;
    if ((rc = grow_store(o, st, std::max<uint64_t>(n, 1), o.nw))) {
        if (ctx->d_keys == vk + o.off[0] * o.nw) {
            ctx->n_out = 0;
            ctx->d_keys = nullptr;
            ctx->d_cnt = nullptr;
        }
        return rc;
    }
    if (n) {
        HIPOK(hipMemcpyAsync(o.keys, vk, n * o.nw * 8ull, hipMemcpyDeviceToDevice, st));
        HIPOK(hipMemcpyAsync(o.cnt, vc, n * 4ull, hipMemcpyDeviceToDevice, st));
    }
    o.n = n;
    if (ctx->d_keys == vk + o.off[0] * o.nw) {
        ctx->d_keys = o.keys + o.off[0] * o.nw;
        ctx->d_cnt = o.cnt + o.off[0];
    }
    HIPOK(hipStreamSynchronize(st));
    return TG_OK;
}

int count_many_device(tg_ctx* ctx, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                      const uint64_t* d_off, uint64_t n_chunks, const JobTab& jt,
                      tg_job_stats* js, bool single = false)
{
    int rc = check_params(p, 0, 0);
    if (rc) return rc;
    if (p->k > 64)
        return fail(TG_EINVAL, "the batched engine counts k <= 64; count jobs of k = %u one at a time", p->k);
    if (!jt.job_chunk || !jt.job_base) return fail(TG_EINVAL, "null job table");
    if (jt.job_chunk[jt.n] > n_chunks || jt.job_base[jt.n] > n_bases)
        return fail(TG_EINVAL, "job table exceeds the input (%llu chunks, %llu bases)",
                    (unsigned long long)n_chunks, (unsigned long long)n_bases);
    for (uint32_t j = 0; j < jt.n; ++j) {
        if (jt.job_chunk[j + 1] < jt.job_chunk[j] || jt.job_base[j + 1] < jt.job_base[j])
            return fail(TG_EINVAL, "job table is not non-decreasing at job %u", j);
        if (jt.job_base[j + 1] - jt.job_base[j] >= (1ull << 31))
            return fail(TG_EINVAL, "job %u has %llu bases; a job is limited to 2^31 in a batch", j,
                        (unsigned long long)(jt.job_base[j + 1] - jt.job_base[j]));
    }
    HIPOK(hipSetDevice(ctx->device));
    BatchState* s = state(ctx);
    hipStream_t st = ctx->stream;
    if (single) {
        // the context's previous result is replaced: drop it now, so a call
        // that fails part-way leaves no pointer into buffers it reallocated
        ctx->n_out = 0;
        ctx->d_keys = nullptr;
        ctx->d_cnt = nullptr;
    } else if ((rc = migrate_view(ctx, s))) {
        // a many-job run may reallocate the batch buffers
        return rc;
    }
    const auto t_start = std::chrono::steady_clock::now();
    s->last = tg_many_stats{};
    s->ms_p1 = s->ms_p2 = 0;
    s->last_items = s->last_spilled = 0;
    const uint32_t nw = p->k > 32 ? 2u : 1u;
    s->count_blocks = tg::count_persistent_blocks(nw);
    tg::set_kernel_attributes();
    std::vector<uint32_t> cuts;
    plan_batches(jt, s->cfg, cuts);
    BatchClass want = class_for(p, jt, cuts, s->cfg);
    if (s->have_cls && s->cls.covers(want)) want = s->cls;
    ResultStore& out = single ? s->one : s->many;
    out.off.assign(jt.n, 0);
    out.len.assign(jt.n, 0);
    out.n = 0;
    out.view_keys = nullptr;
    out.view_cnt = nullptr;
    const bool view = single && cuts.size() == 2;
    if ((rc = grow_store(out, st, 1, nw))) return rc;
    uint64_t tot_kmers = 0, tot_uniq = 0;
    for (size_t bi = 0; bi + 1 < cuts.size(); ++bi) {
        const uint32_t j0 = cuts[bi], j1 = cuts[bi + 1];
        for (int attempt = 0;; ++attempt) {
            if (!s->have_cls || !(s->cls == want)) {
                if ((rc = alloc_class(s, want))) return rc;
            }
            const BatchClass& c = s->cls;
            // per-batch tables
            tg::BatchDesc* d = reinterpret_cast<tg::BatchDesc*>(s->h_tab);
            uint32_t* hjc = reinterpret_cast<uint32_t*>(s->h_tab + 256);
            uint32_t* hjr = hjc + c.cap_jobs + 1;
            const uint64_t c0 = jt.job_chunk[j0];
            d->bases = d_bases + jt.job_base[j0];
            d->off = d_off + c0;
            d->off_base = jt.job_base[j0];
            d->n_bases = jt.job_base[j1] - jt.job_base[j0];
            d->n_chunks = jt.job_chunk[j1] - c0;
            d->n_jobs = j1 - j0;
            const double target = (double)s->cfg.load_factor *
                                  (nw == 1 ? tg::LDS_CAP_NW1 : tg::LDS_CAP_NW2);
            d->scale = (float)((s->cfg.rho > 0 ? s->cfg.rho : s->rho) / target);
            uint32_t base = 0;
            for (uint32_t j = j0; j < j1; ++j) {
                const uint32_t nb =
                    job_buckets(jt.job_base[j + 1] - jt.job_base[j], s->cfg.kmers_per_bucket);
                hjc[j - j0] = (uint32_t)(jt.job_chunk[j] - c0);
                hjr[j - j0] = base | (ilog2(nb) << 16);
                base += nb;
            }
            hjc[j1 - j0] = (uint32_t)(jt.job_chunk[j1] - c0);
            const uint64_t tab_bytes = 256 + 4ull * (c.cap_jobs + 1) + 4ull * c.cap_jobs;
            HIPOK(hipMemcpyAsync(s->b.tab, s->h_tab, tab_bytes, hipMemcpyHostToDevice, st));
            HIPOK(hipEventRecord(s->ev_start, st));
            if (s->cfg.use_graph && !ctx->profiling) {
                if ((rc = ensure_graph(s, st))) return rc;
                HIPOK(hipGraphLaunch(s->exec[0], st));
                HIPOK(hipEventRecord(s->ev_mid, st));
                HIPOK(hipGraphLaunch(s->exec[1], st));
                s->last.n_graph_launches += 1;
            } else {
                // plain launches (also whenever per-kernel timing is on)
                enqueue_batch(s, st, ctx, 3);
                HIPOK(hipGetLastError());
            }
            HIPOK(hipEventRecord(s->ev_end, st));
            // readback: job offsets, superkmers, k-mers, distinct, flags, counters
            const uint32_t nj = j1 - j0;
            uint64_t* h_off = reinterpret_cast<uint64_t*>(s->h_back);
            uint64_t* h_nsk = h_off + c.cap_jobs + 1;
            uint64_t* h_nkm = h_nsk + c.cap_jobs;
            uint64_t* h_unq = h_nkm + c.cap_jobs;
            uint32_t* h_flags = reinterpret_cast<uint32_t*>(h_unq + c.cap_jobs);
            unsigned long long* h_ctr = reinterpret_cast<unsigned long long*>(h_flags + F_WORDS);
            HIPOK(hipMemcpyAsync(s->h_back, s->b.back, back_bytes(c.cap_jobs), hipMemcpyDeviceToHost, st));
            HIPOK(hipStreamSynchronize(st));
            {
                float a1 = 0, a2 = 0;
                if (hipEventElapsedTime(&a1, s->ev_start, s->ev_mid) == hipSuccess &&
                    hipEventElapsedTime(&a2, s->ev_mid, s->ev_end) == hipSuccess) {
                    s->ms_p1 += a1;
                    s->ms_p2 += a2;
                }
                (void)hipGetLastError();
            }
            if (h_flags[F_SKERR] & 4u)
                return fail(TG_EINVAL, "device chunk offsets disagree with the job table (batch %zu)",
                            bi);
            if (h_flags[F_SKERR] || h_flags[F_SPERR]) {
                if (attempt >= 4)
                    return fail(TG_ENOMEM, "batch overflow after %d retries (fix-up %u, spill %u)",
                                attempt, h_flags[F_SKERR], h_flags[F_SPERR]);
                if (h_flags[F_SKERR] & 1u)
                    want.fixb = (uint32_t)std::max<uint64_t>(
                        want.fixb * 2ull, (h_flags[F_FLAGS] + tg::SK_BLOCK - 1) / tg::SK_BLOCK);
                if (h_flags[F_SKERR] & 2u) want.rec_cap = want.cap_bases;
                if (h_flags[F_SPERR]) {
                    // items sized from a distinct/k-mer ratio learned on other
                    // input can hold several tables' worth of distinct k-mers:
                    // split them further as well as growing the spill table
                    want.sp_slots <<= 2;
                    s->rho = std::min(1.0, s->rho * 4.0);
                }
                s->last.n_retries += 1;
                continue;
            }
            const uint64_t n_out = h_off[nj];
            // a batch of one job is not gathered: its result is the items' output
            uint64_t* res_keys = nj == 1 ? s->b.out_keys : s->b.fin_keys;
            uint32_t* res_cnt = nj == 1 ? s->b.out_cnt : s->b.fin_cnt;
            if (view) {
                out.view_keys = res_keys;
                out.view_cnt = res_cnt;
                out.nw = nw;
            } else if ((rc = grow_store(out, st, out.n + n_out, nw))) {
                return rc;
            }
            if (n_out && !view) {
                HIPOK(hipMemcpyAsync(out.keys + out.n * nw, res_keys, n_out * nw * 8ull,
                                     hipMemcpyDeviceToDevice, st));
                HIPOK(hipMemcpyAsync(out.cnt + out.n, res_cnt, n_out * 4ull,
                                     hipMemcpyDeviceToDevice, st));
            }
            uint64_t bk = 0, bu = 0;
            for (uint32_t q = 0; q < nj; ++q) {
                const uint32_t j = j0 + q;
                out.off[j] = out.n + h_off[q];
                out.len[j] = h_off[q + 1] - h_off[q];
                if (js) {
                    js[j].n_bases = jt.job_base[j + 1] - jt.job_base[j];
                    js[j].n_kmers = h_nkm[q];
                    js[j].n_superkmers = h_nsk[q];
                    js[j].n_unique = h_unq[q];
                    js[j].n_out = out.len[j];
                }
                bk += h_nkm[q];
                bu += h_unq[q];
            }
            out.n += n_out;
            tot_kmers += bk;
            tot_uniq += bu;
            // distinct/k-mers of this batch sizes the splits of the next one
            if (bk) s->rho = std::min(1.0, std::max(0.02, 1.15 * (double)bu / (double)bk));
            s->last.n_batches += 1;
            s->last_items += h_flags[F_NITEMS];
            s->last_spilled += h_ctr[2];
            break;
        }
    }
    HIPOK(hipStreamSynchronize(st));
    s->last.n_jobs = jt.n;
    s->last.n_kmers = tot_kmers;
    s->last.n_unique = tot_uniq;
    s->last.n_out = out.n;
    s->last.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return TG_OK;
}

}  // namespace

int count_one_batched(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                      const uint64_t* d_off, uint64_t n_chunks, tg_stats* st)
{
    return count_range_batched(c, p, d_bases, n_bases, d_off, n_chunks, 0, n_chunks, 0, n_bases, st);
}

int count_range_batched(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                        const uint64_t* d_off, uint64_t n_chunks, uint64_t c0, uint64_t c1, uint64_t b0,
                        uint64_t b1, tg_stats* st)
{
    int rc = check_params(p, n_bases, n_chunks);
    if (rc) return rc;
    const uint64_t jc[2] = {c0, c1};
    const uint64_t jb[2] = {b0, b1};
    const JobTab jt{jc, jb, 1};
    tg_job_stats js{};
    if ((rc = count_many_device(c, p, d_bases, n_bases, d_off, n_chunks, jt, &js, true))) return rc;
    BatchState* s = c->batch;
    c->n_out = s->one.len[0];
    c->nw = s->one.nw;
    uint64_t* rk = s->one.view_keys ? s->one.view_keys : s->one.keys;
    uint32_t* rc_ = s->one.view_cnt ? s->one.view_cnt : s->one.cnt;
    c->d_keys = rk + s->one.off[0] * s->one.nw;
    c->d_cnt = rc_ + s->one.off[0];
    collect_prof(c);
    if (st) {
        tg_stats o{};
        o.n_bases = b1 - b0;
        o.n_chunks = c1 - c0;
        o.n_kmers = js.n_kmers;
        o.n_superkmers = js.n_superkmers;
        o.n_unique = js.n_unique;
        o.n_out = js.n_out;
        o.n_parts = round_pow2(p->n_partitions ? p->n_partitions : 16u);
        o.n_subparts = job_buckets(b1 - b0, s->cfg.kmers_per_bucket);
        o.n_lds_groups = (uint32_t)std::min<uint64_t>(s->last_items, 0xFFFFFFFFull);
        o.n_hbm_parts = (uint32_t)std::min<uint64_t>(s->last_spilled, 0xFFFFFFFFull);
        o.ms_phase1 = s->ms_p1;
        o.ms_phase2 = s->ms_p2;
        *st = o;
    }
    return TG_OK;
}

uint64_t slice_limit(tg_ctx* c)
{
    const uint64_t v = state(c)->cfg.slice_bases;
    return v ? v : DEFAULT_SLICE_BASES;
}

void drop_single_view(tg_ctx* c)
{
    if (!c->batch) return;
    ResultStore& o = c->batch->one;
    o.view_keys = nullptr;
    o.view_cnt = nullptr;
    o.n = 0;
}

}  // namespace tgi

using namespace tgi;

extern "C" {

int tg_get_batch_config(tg_ctx* c, tg_batch_config* cfg)
{
    if (!c || !cfg) return fail(TG_EINVAL, "null argument");
    *cfg = state(c)->cfg;
    return TG_OK;
}

int tg_set_batch_config(tg_ctx* c, const tg_batch_config* cfg)
{
    if (!c || !cfg) return fail(TG_EINVAL, "null argument");
    if (cfg->kmers_per_bucket < 256 || (cfg->kmers_per_bucket & (cfg->kmers_per_bucket - 1)))
        return fail(TG_EINVAL, "kmers_per_bucket must be a power of two >= 256");
    if (cfg->max_tiles < 1 || cfg->max_tiles > 4096)
        return fail(TG_EINVAL, "max_tiles must be in [1, 4096]");
    if (cfg->batch_bases < (1ull << 16) || cfg->batch_bases >= (1ull << 31))
        return fail(TG_EINVAL, "batch_bases must be in [2^16, 2^31)");
    if (!(cfg->load_factor > 0.05f && cfg->load_factor <= 1.0f))
        return fail(TG_EINVAL, "load_factor must be in (0.05, 1]");
    if (!(cfg->rho == 0.0f || (cfg->rho >= 0.001f && cfg->rho <= 1.0f)))
        return fail(TG_EINVAL, "rho must be 0 (learned) or in [0.001, 1]");
    if (cfg->slice_bases != 0 && (cfg->slice_bases < 1024 || cfg->slice_bases > DEFAULT_SLICE_BASES))
        return fail(TG_EINVAL, "slice_bases must be 0 (default) or in [1024, %llu]",
                    (unsigned long long)DEFAULT_SLICE_BASES);
    BatchState* s = state(c);
    HIPOK(hipSetDevice(c->device));
    HIPOK(hipStreamSynchronize(c->stream));
    if (cfg->max_tiles != s->cfg.max_tiles) {
        // the batch buffers are freed: a single job's result kept there moves first
        const int rc = migrate_view(c, s);
        if (rc) return rc;
        free_class(s);
    }
    s->cfg = *cfg;
    return TG_OK;
}

int tg_count_many_device(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                         const uint64_t* d_off, uint64_t n_chunks, const uint64_t* job_chunk,
                         const uint64_t* job_base, uint32_t n_jobs, tg_job_stats* st,
                         tg_many_stats* ms)
{
    if (!c) return fail(TG_EINVAL, "null context");
    const JobTab jt{job_chunk, job_base, n_jobs};
    const int rc = count_many_device(c, p, d_bases, n_bases, d_off, n_chunks, jt, st);
    if (rc == TG_OK && ms) *ms = c->batch->last;
    return rc;
}

int tg_count_many(tg_ctx* c, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                  const uint64_t* off, uint64_t n_chunks, const uint64_t* job_chunk,
                  uint32_t n_jobs, tg_job_stats* st, tg_many_stats* ms)
{
    if (!c) return fail(TG_EINVAL, "null context");
    if (!job_chunk || (!off && n_chunks) || (!bases && n_bases)) return fail(TG_EINVAL, "null argument");
    if (job_chunk[n_jobs] > n_chunks) return fail(TG_EINVAL, "job table exceeds the chunk count");
    HIPOK(hipSetDevice(c->device));
    std::vector<uint64_t> job_base(n_jobs + 1);
    for (uint32_t j = 0; j <= n_jobs; ++j) job_base[j] = n_chunks ? off[job_chunk[j]] : 0;
    uint8_t* d_b;
    uint64_t* d_o;
    int rc;
    if ((rc = ensure_t(c, "many_bases", n_bases + 64, &d_b)) ||
        (rc = ensure_t(c, "many_off", n_chunks + 1, &d_o)))
        return rc;
    if (n_bases) HIPOK(hipMemcpyAsync(d_b, bases, n_bases, hipMemcpyHostToDevice, c->stream));
    if (n_chunks) HIPOK(hipMemcpyAsync(d_o, off, (n_chunks + 1) * 8ull, hipMemcpyHostToDevice, c->stream));
    const JobTab jt{job_chunk, job_base.data(), n_jobs};
    rc = count_many_device(c, p, d_b, n_bases, d_o, n_chunks, jt, st);
    if (rc == TG_OK && ms) *ms = c->batch->last;
    return rc;
}

// Coordinate search over the batched path's tunables on the caller's jobs:
// every candidate runs twice (the first run sizes buffers and captures the
// graph), the second is timed; every run must reproduce the per-job totals of
// the starting configuration.  The fastest configuration is installed.
int tg_autotune_many(tg_ctx* c, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                     const uint64_t* d_off, uint64_t n_chunks, const uint64_t* job_chunk,
                     const uint64_t* job_base, uint32_t n_jobs, tg_batch_config* best_out,
                     double* best_ms)
{
    if (!c || !job_chunk || !job_base) return fail(TG_EINVAL, "null argument");
    const JobTab jt{job_chunk, job_base, n_jobs};
    BatchState* s = state(c);
    tg_batch_config best = s->cfg;
    best.rho = 0.0f;
    std::vector<tg_job_stats> ref(n_jobs), st(n_jobs);
    auto run = [&](const tg_batch_config& cfg, std::vector<tg_job_stats>& out, double* ms) -> int {
        int rc = tg_set_batch_config(c, &cfg);
        if (rc) return rc;
        // the first run builds the buffers and the graph; the best of the
        // next three is the candidate's time (one run is within noise of
        // its neighbours)
        double m = 0;
        for (int rep = 0; rep < 4; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = count_many_device(c, p, d_bases, n_bases, d_off, n_chunks, jt, out.data())))
                return rc;
            const double dt =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (rep == 1 || (rep > 1 && dt < m)) m = dt;
        }
        *ms = m;
        return TG_OK;
    };
    double best_t = 0;
    int rc = run(best, ref, &best_t);
    if (rc) return rc;
    const uint64_t total = n_jobs ? job_base[n_jobs] - job_base[0] : 0;
    const uint32_t kpb[] = {4096, 8192, 16384, 32768, 65536};
    const float lf[] = {0.5f, 0.6f, 0.75f, 0.9f};
    const uint32_t tiles[] = {256, 512, 1024};
    const uint64_t bb[] = {1ull << 26, 1ull << 27, 1ull << 28, 1ull << 29, 1ull << 30};
    auto same = [&](const std::vector<tg_job_stats>& a) {
        for (uint32_t j = 0; j < n_jobs; ++j)
            if (a[j].n_kmers != ref[j].n_kmers || a[j].n_superkmers != ref[j].n_superkmers ||
                a[j].n_unique != ref[j].n_unique || a[j].n_out != ref[j].n_out)
                return false;
        return true;
    };
    auto consider = [&](tg_batch_config cand) -> int {
        double t = 0;
        int r = run(cand, st, &t);
        if (r) return r;
        if (!same(st)) return fail(TG_EDEVICE, "autotune: a candidate configuration changed the results");
        if (t < best_t) {
            best_t = t;
            best = cand;
        }
        return TG_OK;
    };
    for (uint32_t v : kpb) {
        tg_batch_config cand = best;
        cand.kmers_per_bucket = v;
        if (v != best.kmers_per_bucket && (rc = consider(cand))) return rc;
    }
    for (uint64_t v : bb) {
        tg_batch_config cand = best;
        cand.batch_bases = v;
        if (v > 2 * std::max<uint64_t>(total, 1u << 16)) continue;  // beyond one batch: no change
        if (v != best.batch_bases && (rc = consider(cand))) return rc;
    }
    for (float v : lf) {
        tg_batch_config cand = best;
        cand.load_factor = v;
        if (v != best.load_factor && (rc = consider(cand))) return rc;
    }
    for (uint32_t v : tiles) {
        tg_batch_config cand = best;
        cand.max_tiles = v;
        if (v != best.max_tiles && (rc = consider(cand))) return rc;
    }
    if ((rc = tg_set_batch_config(c, &best))) return rc;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t;
    return TG_OK;
}

uint64_t tg_many_result_size(tg_ctx* c, uint32_t job)
{
    if (!c || !c->batch || job >= c->batch->many.len.size()) return 0;
    return c->batch->many.len[job];
}

int tg_many_result_copy(tg_ctx* c, uint32_t job, uint64_t* keys, uint32_t* counts, uint64_t cap)
{
    if (!c) return fail(TG_EINVAL, "null context");
    BatchState* s = c->batch;
    if (!s || job >= s->many.len.size()) return fail(TG_EINVAL, "no result for job %u", job);
    const uint64_t n = std::min<uint64_t>(cap, s->many.len[job]);
    if (!n) return TG_OK;
    HIPOK(hipSetDevice(c->device));
    const uint32_t nw = s->many.nw;
    if (keys)
        HIPOK(hipMemcpyAsync(keys, s->many.keys + s->many.off[job] * nw, n * nw * 8ull,
                             hipMemcpyDeviceToHost, c->stream));
    if (counts)
        HIPOK(hipMemcpyAsync(counts, s->many.cnt + s->many.off[job], n * 4ull, hipMemcpyDeviceToHost,
                             c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_many_result_device(tg_ctx* c, uint32_t job, const uint64_t** d_keys, const uint32_t** d_counts,
                          uint64_t* n)
{
    if (!c) return fail(TG_EINVAL, "null context");
    BatchState* s = c->batch;
    if (!s || job >= s->many.len.size()) return fail(TG_EINVAL, "no result for job %u", job);
    if (d_keys) *d_keys = s->many.keys + s->many.off[job] * s->many.nw;
    if (d_counts) *d_counts = s->many.cnt + s->many.off[job];
    if (n) *n = s->many.len[job];
    return TG_OK;
}

}  // extern "C"
