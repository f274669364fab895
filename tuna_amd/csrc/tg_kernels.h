// tg_kernels.h — launch-side declarations of the gfx950 kernels (host and
// device code share these argument structs).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

// Superkmer records (device partition buffers), RW 32-bit words each:
//   slim (RW = 2, for exports):  w0 = global index of the first base,
//                                w1 = meta
//   fat  (RW = 4 for k <= 32, 8 for k <= 64, for counting):
//                                w0 = meta, w1.. = the record's bases packed
//                                16 per word, first base in bits 31:30
//   meta = len (bits 0-7) | minpos (bits 8-13) | hsel (bits 14-31)
//   hsel = (bucket << SB) | split bits
// A fat record holds at most 16*(RW-1) bases, so a longer superkmer is cut into
// pieces of the same minimizer (same bucket); the first piece keeps the
// superkmer's minpos, the others set REC_CONT in the length byte (fat lengths
// are < 128).  Pieces change nothing counted: every k-mer is in one piece.
constexpr uint32_t REC_LEN_MASK = 0xFFu;
constexpr uint32_t REC_FAT_LEN_MASK = 0x7Fu;
constexpr uint32_t REC_CONT = 0x80u;
constexpr uint32_t REC_MINPOS_SHIFT = 8;
constexpr uint32_t REC_HSEL_SHIFT = 14;
constexpr uint32_t HSEL_BITS = 18;
constexpr uint32_t LOOKBACK_FULL = 0xFFFFFFFFu;
constexpr uint32_t SK_BLOCK = 256;      // lanes (strips) per superkmer workgroup
constexpr uint32_t SK_WAVE_BUF = 256;   // records staged per wave before a flush

// words per record of the counting path
__host__ __device__ constexpr uint32_t fat_rw(uint32_t k) { return k <= 32 ? 4u : 8u; }

// Phase-1 output is a list of "runs": run r holds run_cnt[r] records starting
// at record index run_start[r] of `recs`.  The main pass writes one compacted
// run per workgroup into the slot space of its strips (a strip of K k-mers has
// at most K records, so slot = first k-mer index: no global atomics); the
// fix-up pass writes one run per workgroup into a region of SK_BLOCK*S records
// per block starting at fix_base.
struct SkArgs {
    const uint32_t* packed;
    const uint64_t* off;        // chunk offsets (global base index), n_chunks+1
    const uint32_t* strip_chunk;
    const uint32_t* strip_tb;   // first owned k-mer of the strip (chunk-relative)
    const uint64_t* strip_slot; // main pass: record slot of the strip's first k-mer
    const uint32_t* list;       // fix-up pass: strip indices; nullptr in the main pass
    uint32_t n;                 // strips (or list entries) to process
    uint32_t k, m, W, S;
    uint32_t capk;              // a superkmer spans at most capk k-mers (256-k)
    uint32_t piece;             // fat records: k-mers per piece (16*(RW-1)-k+1); 0 for slim
    uint32_t lookback;          // k-mers of look-back before tb, or LOOKBACK_FULL
    uint32_t log2n, xb;         // routing field: job bucket bits on top, xb bits below
    // batched jobs: job of each chunk and its bucket range (base | log2(buckets) << 16);
    // nullptr for a single job, whose buckets are the reference partitions (log2n bits)
    const uint32_t* chunk_job;
    const uint32_t* job_route;
    // device-side sizes (nullptr: n is exact): main pass reads the strip total,
    // fix-up pass reads flag_count; n is then the capacity of the launch
    const uint64_t* n_dev;
    uint32_t fix_dev;
    uint32_t* err;              // bit 0: fix-up capacity exceeded
    uint32_t* recs;             // record storage (RW words per record)
    uint32_t* metas;            // the meta word of every record, same slots (routing histograms)
    uint32_t fix_base;          // fix-up pass: first record index of its regions
    uint32_t* run_start;        // run table, indexed from run_base + blockIdx.x
    uint32_t* run_cnt;
    uint32_t run_base;
    unsigned long long* sk_count;  // superkmers (first pieces) per job (one entry for a single job)
    uint32_t* flag_list;
    uint32_t* flag_count;
    uint32_t dq_cap;            // k > 64 main pass: deque entries per lane (1..WDQ_MAX)
    uint32_t* ovf_list;         // k > 64 main pass: strips whose deque overflowed
    uint32_t* ovf_count;
};

// k > 64 (keys of 4 or 8 words): the k-mers of every strip go into one
// open-addressed HBM table whose slots hold a key's position in the packed
// stream (positions below 2^40) and its count.
constexpr uint32_t WIDE_WMAX = 256;  // m-mers per window (k - m + 1) on that path
constexpr uint32_t WDQ_MAX = 16;     // deque entries per lane of its superkmer pass
struct WideArgs {
    const uint32_t* packed;
    const uint64_t* off;        // chunk offsets (global base index), n_chunks+1
    const uint32_t* strip_chunk;
    const uint32_t* strip_tb;
    uint32_t n_strips;          // capacity of the launch
    const uint64_t* n_dev;      // live strip total
    uint32_t S, k;
    uint64_t* tab;              // two words per slot: tag (0 empty), count
    uint64_t mask;              // slots - 1 (a power of two, >= 2^16)
    uint32_t* err;              // set when a key found no slot
};

struct RouteArgs {
    const uint32_t* recs;      // record storage of the runs (RW words per record)
    const uint32_t* metas;     // meta word per record slot (histograms read only these)
    const uint32_t* run_start;
    const uint32_t* run_cnt;
    const uint64_t* run_off;   // exclusive scan of run_cnt (n_runs+1)
    uint32_t n_runs;
    uint64_t n_rec;
    uint32_t tile;             // records per tile
    uint32_t nsp, sb, k;
    uint32_t* hsk;             // [ntiles][nsp] superkmers, becomes tile-exclusive offsets
    uint32_t* hkm;             // [ntiles][nsp] k-mers
    const uint64_t* part_off;  // nsp+1 (scatter only)
    uint32_t* dst;             // partition-ordered records (scatter only)
    // device-side record total (nullptr: n_rec/tile/grid are exact); the
    // kernels then derive tile and tile count themselves, up to max_tiles.  A
    // total above rec_cap (the partition buffers' capacity) routes nothing and
    // sets bit 1 of *err.
    const uint64_t* n_rec_dev;
    uint32_t max_tiles;
    uint64_t rec_cap;
    uint32_t* err;
    // scatter: the meta word of every coarse-ordered record (read by the fine
    // pass to size its groups before moving records); the fine pass orders a
    // sub-partition's records by the top split_bits of their split field and,
    // when split_off is set, writes the start of every (sub-partition, split
    // group): nsp << split_bits + 1 entries (see route_split_bits)
    uint32_t* tmp_meta;
    uint64_t* split_off;
    uint32_t split_bits;
    // which coarse buckets are split-ordered: those holding a bucket whose
    // k-mers x *split_scale exceed 1 (the work items' split rule); nullptr: none
    const uint64_t* split_km;
    const float* split_scale;
    uint32_t* split_flags;  // per coarse bucket (256 words): 1 = split-ordered
    // group counts of the FINE_RANGES record ranges of every split-ordered
    // coarse bucket, FINE_RANGES << HSEL_BITS words (nullptr: each bucket is
    // scattered by one workgroup, which sizes its groups itself)
    uint32_t* fine_hist;
};

// Record ranges a split-ordered coarse bucket is cut into for the fine
// scatter, one workgroup each, so that a large bucket does not hold up the
// pass (the buckets of a skewed input differ in size by 2x and more).
constexpr uint32_t FINE_RANGES = 4;

// Split-field bits the fine scatter orders records by (at most 512 groups
// per coarse bucket).
uint32_t route_split_bits(uint32_t nsp);

// Phase-2 work item: records [x, y) of consecutive sub-partitions; records with
// (split bits & z) == w are counted by this item.
struct CountArgs {
    const uint32_t* rec;       // partition-ordered fat records
    const uint4* items;
    uint32_t n_items;
    uint32_t k;
    uint32_t ci;
    uint64_t cx;
    uint32_t probe_max;
    // k <= 32: records of at least per_len bases (longer than a superkmer of
    // one minimizer occurrence, k + W; 0 = never) are tested for a short period
    // and then counted with weights (a tandem repeat's k-mers recur every P)
    uint32_t per_len;
    // spill table (global, shared by all items; keys are disjoint by partition)
    uint64_t* sp_keys;         // NW words per slot
    uint32_t* sp_cnt;
    uint64_t sp_mask;
    uint32_t* sp_err;
    // output
    uint64_t* out_keys;        // NW words per k-mer
    uint32_t* out_cnt;
    unsigned long long* counters;  // [0] n_out, [1] n_unique, [2] spilled k-mers
    uint64_t out_cap;
    uint32_t* item_uniq;       // optional: distinct keys per item (calibration)
    uint32_t* item_sp;         // optional: k-mers spilled per item (debug statistics)
    // batched jobs (persistent launch): items are taken from a device queue;
    // spill keys carry the job as a tag; each item reports its output segment
    const uint32_t* n_items_dev;
    uint32_t* item_q;
    const uint32_t* item_job;
    uint64_t* seg_pos;
    uint64_t* seg_n;
    unsigned long long* job_uniq;
};

// Batch descriptor (device memory, rewritten before every batch): where the
// batch's bases and chunk offsets are and how many there are.
struct BatchDesc {
    const uint8_t* bases;      // first base of the batch
    const uint64_t* off;       // offsets of the batch's chunks (n_chunks+1, absolute)
    uint64_t off_base;         // off[0]: subtracted to get batch-relative positions
    uint64_t n_bases;
    uint64_t n_chunks;
    uint32_t n_jobs;
    float scale;               // work-item splits: estimated distinct = k-mers x scale / table
};

// Batch job tables (device): job j owns chunks [job_chunk[j], job_chunk[j+1])
// (batch-relative) and buckets [base, base + 2^L), job_route[j] = base | L << 16.
struct BatchJobs {
    const uint32_t* job_chunk;
    const uint32_t* job_route;
};

// launchers (defined in kernels.hip)
void launch_pack(hipStream_t st, const uint8_t* bases, uint64_t n, uint32_t* out, uint64_t n_words);
// long_list: fill_list_words(strip capacity) words, cleared by the chunk
// counts and used by the strip fill that follows them on the same stream
uint64_t fill_list_words(uint64_t cap_strips);
void launch_chunk_counts(hipStream_t st, const uint64_t* off, uint64_t n_chunks, uint32_t k,
                         uint32_t S, uint64_t* nstr, uint64_t* nkm, uint64_t* long_list);
void launch_fill_strips(hipStream_t st, uint64_t n_chunks, const uint64_t* nstr_excl,
                        const uint64_t* nkm_excl, const uint64_t* nkm, uint32_t S,
                        uint32_t* strip_chunk, uint32_t* strip_tb, uint64_t* strip_slot,
                        uint64_t* long_list, uint64_t list_words);
// exclusive scan of n u64 values into out[0..n] (out[n] = total); tmp >= scan_tmp_words(n)
uint64_t scan_tmp_words(uint64_t n);
void launch_scan_u64(hipStream_t st, const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* tmp);
void launch_u32_to_u64(hipStream_t st, const uint32_t* in, uint64_t n, uint64_t* out);
// rw: words per record (2 = slim, 4 or 8 = fat)
void launch_superkmers(hipStream_t st, const SkArgs& a, uint32_t rw);
void launch_superkmers_wide(hipStream_t st, const SkArgs& a);
void launch_wide_insert(hipStream_t st, const WideArgs& w, uint32_t nw);
// drain blocks of a table: per-block totals (written, distinct), scanned by the
// caller, then the written keys at the scanned block offsets
uint32_t wide_drain_blocks(uint64_t slots);
void launch_wide_drain_count(hipStream_t st, const WideArgs& w, uint32_t ci, uint64_t cx, uint64_t* blk_keep,
                             uint64_t* blk_occ);
void launch_wide_drain_write(hipStream_t st, const WideArgs& w, uint32_t nw, uint32_t ci, uint64_t cx, uint32_t onw,
                             const uint64_t* blk_off, uint64_t* out_keys, uint32_t* out_cnt);
void launch_route_hist(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint32_t rw);
void launch_route_colscan(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint64_t* part_sk,
                          uint64_t* part_km);
void launch_route_scatter(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint32_t* tmp,
                          uint32_t rw);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint64_t n);
// Zeroes up to ZERO_MAX regions in one launch (elements of 4 or 8 bytes).
constexpr int ZERO_MAX = 6;
struct ZeroRegions {
    void* p[ZERO_MAX];
    uint64_t n[ZERO_MAX];
    uint32_t wide[ZERO_MAX];  // 1: 8-byte elements
    uint32_t count;
};
void launch_zero_regions(hipStream_t st, const ZeroRegions& z);
void launch_fill_u64(hipStream_t st, uint64_t* p, uint64_t v, uint64_t n);
void launch_count(hipStream_t st, const CountArgs& a, uint32_t nw);
void launch_spill_drain(hipStream_t st, const CountArgs& a, uint32_t nw, uint64_t n_slots);
uint32_t count_persistent_blocks(uint32_t nw);

// batched path
void launch_pack_desc(hipStream_t st, const BatchDesc* d, uint32_t* out, uint64_t cap_words);
void launch_batch_setup(hipStream_t st, const BatchDesc* d, BatchJobs j, uint32_t cap_jobs,
                        uint64_t* boff, uint32_t* chunk_job, uint64_t cap_ch, uint32_t* bucket_job,
                        uint32_t nsp, uint32_t* err);
void set_kernel_attributes();
// nitem: nsp per-bucket item counts; scan_tmp >= scan_tmp_words(nsp)
// split_off / split_bits: the fine scatter's group starts; a bucket split into
// at most 2^split_bits items gets contiguous record ranges (no filter).
void launch_build_items(hipStream_t st, const uint64_t* part_km, const uint64_t* part_off,
                        uint32_t nsp, uint32_t xb, const uint32_t* bucket_job, const float* scale,
                        uint4* items, uint32_t* item_job, uint64_t* bucket_item, uint32_t* n_items,
                        uint64_t* nitem, uint64_t* scan_tmp, const uint64_t* split_off,
                        uint32_t split_bits);
void launch_count_batch(hipStream_t st, const CountArgs& a, uint32_t nw, uint32_t blocks);
// spill entries of a batch: per-job distinct and reported counts, then placement
void launch_spill_jobs(hipStream_t st, const CountArgs& a, uint32_t nw, uint64_t n_slots,
                       unsigned long long* job_sp);
// The batch readback in one buffer, laid out as [job offsets: cap_jobs + 1]
// [superkmers: cap_jobs][k-mers: cap_jobs][distinct: cap_jobs] (u64), then
// n_flags u32 flags and 8 u64 counters, so the host needs one copy.
void launch_pack_back(hipStream_t st, const BatchDesc* d, uint32_t cap_jobs, const uint64_t* job_off,
                      const uint64_t* job_nsk, const uint64_t* job_nkm,
                      const unsigned long long* job_uniq, const uint32_t* flags, uint32_t n_flags,
                      const unsigned long long* counters, uint8_t* back);
void launch_job_totals(hipStream_t st, const BatchDesc* d, BatchJobs j, uint32_t cap_jobs,
                       const uint64_t* bucket_item, const uint64_t* seg_x,
                       const unsigned long long* job_sk, const unsigned long long* job_sp,
                       const uint64_t* nkm_x, uint64_t* job_tot, uint64_t* job_item_tot,
                       uint64_t* job_nsk, uint64_t* job_nkm);
// A batch of one job (d->n_jobs == 1) is not gathered: its result is a.out_keys
// (the items' reservations tile it from 0) with the spill entries placed after
// them.
void launch_gather_items(hipStream_t st, const CountArgs& a, uint32_t nw, BatchJobs j, const BatchDesc* d,
                         const uint64_t* bucket_item, const uint64_t* seg_x,
                         const uint64_t* job_off, uint64_t* dst_keys, uint32_t* dst_cnt,
                         uint32_t blocks);
void launch_spill_place(hipStream_t st, const CountArgs& a, const BatchDesc* d, uint32_t nw, uint64_t n_slots,
                        const uint64_t* job_off, const uint64_t* job_item_tot,
                        unsigned long long* job_cur, uint64_t* dst_keys, uint32_t* dst_cnt);

// Merge table of a job counted in slices (merge.hip): slot i holds nw key
// words at keys[nw*i] and a count at cnt[i]; mask + 1 slots (a power of two);
// claimed counts the slots taken; err is set if a probe found the table full.
struct MergeTab {
    uint64_t* keys;
    uint32_t* cnt;
    uint64_t mask;
    unsigned long long* claimed;
    uint32_t* err;
};
void launch_merge_clear(hipStream_t st, const MergeTab& t, uint32_t nw);
void launch_merge_insert(hipStream_t st, const MergeTab& t, uint32_t nw, const uint64_t* keys,
                         const uint32_t* cnt, uint64_t n);
void launch_merge_rehash(hipStream_t st, const MergeTab& dst, const MergeTab& src, uint32_t nw);
void launch_merge_drain(hipStream_t st, const MergeTab& t, uint32_t nw, uint32_t ci, uint64_t cx,
                        uint64_t* out_keys, uint32_t* out_cnt, unsigned long long* ctr);

// table geometry of the LDS counting kernels
constexpr uint32_t LDS_CAP_NW1 = 8192;  // 8192 x (8+4) B = 96 KiB
constexpr uint32_t LDS_CAP_NW2 = 2048;  // 2048 x (16+4) B = 40 KiB
// count workgroup: k <= 32, 16 waves around a 96 KiB table (one per CU);
// k > 32, 8 waves around a 40 KiB table (two per CU)
constexpr uint32_t count_block(int nw) { return nw == 1 ? 1024u : 512u; }
constexpr uint32_t ROUTE_MAX_TILES = 512;
constexpr uint32_t FIX_BLOCKS = 8;      // fix-up capacity of a batch: 8 x SK_BLOCK strips

}  // namespace tg
