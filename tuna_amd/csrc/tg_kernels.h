// tg_kernels.h — launch-side declarations of the gfx950 kernels (host and
// device code share these argument structs).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

// Superkmer record, 8 bytes (device partition buffers):
//   x = global base index of the first base (packed stream)
//   y = len (bits 0-7) | minpos (bits 8-13) | hsel (bits 14-31)
//   hsel = (sub-partition id << SB) | split bits, SB = 18 - log2(NSP)
constexpr uint32_t REC_LEN_MASK = 0xFFu;
constexpr uint32_t REC_MINPOS_SHIFT = 8;
constexpr uint32_t REC_HSEL_SHIFT = 14;
constexpr uint32_t HSEL_BITS = 18;
constexpr uint32_t LOOKBACK_FULL = 0xFFFFFFFFu;
constexpr uint32_t SK_BLOCK = 256;      // lanes (strips) per superkmer workgroup
constexpr uint32_t SK_WAVE_BUF = 256;   // records staged per wave before a flush

// Phase-1 output is a list of "runs": run i holds run_cnt[i] records at
// run_ptr[i].  The main pass writes one compacted run per workgroup; the
// fix-up pass (strips redone from their chunk start) writes further runs.
struct SkArgs {
    const uint32_t* packed;
    const uint64_t* off;        // chunk offsets (global base index), n_chunks+1
    const uint32_t* strip_chunk;
    const uint32_t* strip_tb;   // first owned k-mer of the strip (chunk-relative)
    const uint64_t* strip_slot; // record slot base of the strip (main pass)
    const uint32_t* list;       // fix-up pass: strip indices; nullptr in the main pass
    uint32_t n;                 // strips (or list entries) to process
    uint32_t k, m, W, S;
    uint32_t capk;              // a superkmer spans at most capk k-mers (256-k)
    uint32_t lookback;          // k-mers of look-back before tb, or LOOKBACK_FULL
    uint32_t log2n, log2s, sb;  // routing: n partitions, s sub-partitions each, SB split bits
    uint2* rec;                 // main pass: slot space; fix-up: S*SK_BLOCK records per block
    uint2** run_ptr;            // run table (indexed from run_base + blockIdx.x)
    uint32_t* run_cnt;
    uint32_t run_base;
    uint32_t* flag_list;
    uint32_t* flag_count;
};

struct RouteArgs {
    uint2* const* run_ptr;
    const uint32_t* run_cnt;
    const uint64_t* run_off;   // exclusive scan of run_cnt (n_runs+1)
    uint32_t n_runs;
    uint64_t n_rec;
    uint32_t tile;             // records per tile
    uint32_t nsp, sb, k;
    uint32_t* hsk;             // [ntiles][nsp] superkmers, becomes tile-exclusive offsets
    uint32_t* hkm;             // [ntiles][nsp] k-mers
    const uint64_t* part_off;  // nsp+1 (scatter only)
    uint2* dst;                // partition-ordered records (scatter only)
};

// Phase-2 work item: records [x, y) of consecutive sub-partitions; records with
// (split bits & z) == w are counted by this item.
struct CountArgs {
    const uint32_t* packed;
    const uint2* rec;
    const uint4* items;
    uint32_t n_items;
    uint32_t k;
    uint32_t ci;
    uint64_t cx;
    uint32_t probe_max;
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
};

// launchers (defined in kernels.hip)
void launch_pack(hipStream_t st, const uint8_t* bases, uint64_t n, uint32_t* out, uint64_t n_words);
void launch_chunk_counts(hipStream_t st, const uint64_t* off, uint64_t n_chunks, uint32_t k,
                         uint32_t S, uint64_t* nstr, uint64_t* nkm);
void launch_fill_strips(hipStream_t st, uint64_t n_chunks, const uint64_t* nstr_excl,
                        const uint64_t* nkm_excl, const uint64_t* nkm, uint32_t S,
                        uint32_t* strip_chunk, uint32_t* strip_tb, uint64_t* strip_slot);
// exclusive scan of n u64 values into out[0..n] (out[n] = total); tmp >= scan_tmp_words(n)
uint64_t scan_tmp_words(uint64_t n);
void launch_scan_u64(hipStream_t st, const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* tmp);
void launch_u32_to_u64(hipStream_t st, const uint32_t* in, uint64_t n, uint64_t* out);
void launch_superkmers(hipStream_t st, const SkArgs& a);
void launch_route_hist(hipStream_t st, const RouteArgs& a, uint32_t ntiles);
void launch_route_colscan(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint64_t* part_sk,
                          uint64_t* part_km);
void launch_route_scatter(hipStream_t st, const RouteArgs& a, uint32_t ntiles, uint2* tmp);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint64_t n);
void launch_fill_u64(hipStream_t st, uint64_t* p, uint64_t v, uint64_t n);
void launch_count(hipStream_t st, const CountArgs& a, uint32_t nw);
void launch_spill_drain(hipStream_t st, const CountArgs& a, uint32_t nw, uint64_t n_slots);

// table geometry of the LDS counting kernels
constexpr uint32_t LDS_CAP_NW1 = 4096;  // 4096 x (8+4) B = 48 KiB
constexpr uint32_t LDS_CAP_NW2 = 2048;  // 2048 x (16+4) B = 40 KiB
constexpr uint32_t COUNT_BLOCK = 512;

}  // namespace tg
