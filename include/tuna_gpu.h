/*
 * tuna_gpu.h — C ABI of the MI355X (gfx950) k-mer counting engine.
 *
 * The reference (vicLeva/tuna v1.4.0) has no FFI or C ABI: it is a
 * header-only C++ library plus a CLI.  The entry points below replace its two
 * in-memory phase seams and the public API built on them:
 *
 *   partition_kmers_mem<k,m>(cfg, bufs)            src/partition_hash.hpp:556-565
 *   count_and_write_mem<k,m>(cfg, K, bufs, out, kff) src/count.hpp:875-970
 *   count_and_callback_mem<k,m>(cfg, K, bufs, cb)  src/count.hpp:311-388
 *   tuna::count / count_to / count_to_raw          include/tuna/tuna.hpp:91-195
 *
 * Conventions: every call returns 0 on success or a negative TG_E* status;
 * tg_last_error() gives a thread-local message for the last failure.  All
 * pointers are plain host or device pointers with explicit sizes; no C++ or
 * torch types cross this boundary.  A context owns one HIP stream, the device
 * buffers of its jobs (kept resident and reused across calls) and the result
 * of its last job.  Calls on one context are serialised by the caller; distinct
 * contexts may be used from distinct threads.
 *
 * K-mer keys use the layout of the reference's Kmer<k> (Kmer.hpp:44-48):
 * ceil(k/32) little-endian 64-bit words, kache 2-bit codes (A0 C1 G2 T3), the
 * last base in bits 1:0 of word 0, the first base most significant; keys are
 * canonical (the smaller of the k-mer and its reverse complement).
 */
#ifndef TUNA_GPU_H
#define TUNA_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK 0
#define TG_EINVAL (-1)   /* bad argument (k, m, sizes, pointers) */
#define TG_EDEVICE (-2)  /* HIP runtime failure */
#define TG_ENOMEM (-3)   /* device memory admission failed */
#define TG_ECAP (-4)     /* caller buffer too small */
#define TG_EINPUT (-5)   /* input file / parse error */
#define TG_ETIMEOUT (-6) /* request not finished within the wait */

typedef struct tg_ctx tg_ctx;

/* Counting parameters; names follow the reference Config (src/Config.hpp:10-26). */
typedef struct {
    uint32_t k;            /* k-mer length, 2..64 on this engine */
    uint32_t m;            /* minimizer length, 1..min(32, k-1) */
    uint32_t n_partitions; /* reference -n: rounded up to a power of 2; 0 -> 16 */
    uint32_t ci;           /* report counts >= ci (reference -ci) */
    uint64_t cx;           /* report counts <= cx (reference -cx) */
} tg_params;

/* Totals of one job; mirror the reference's structured stderr lines
 * (src/pipeline.hpp:174-178) plus engine-side figures. */
typedef struct {
    uint64_t n_bases;      /* ACGT bases handed in */
    uint64_t n_chunks;     /* ACGT runs */
    uint64_t n_kmers;      /* k-mers inserted (reference total_inserted) */
    uint64_t n_superkmers; /* reference "superkmers:" */
    uint64_t n_unique;     /* distinct canonical k-mers before ci/cx */
    uint64_t n_out;        /* reference "unique_kmers:" (after ci/cx) */
    uint32_t n_parts;      /* reference "n_parts:" */
    uint32_t n_subparts;   /* device sub-partitions (n_parts * s) */
    uint32_t n_lds_groups; /* LDS-table work groups */
    uint32_t n_hbm_parts;  /* partitions counted in HBM tables */
    double ms_phase1;      /* device time: pack + superkmers + routing */
    double ms_phase2;      /* device time: counting + drain */
} tg_stats;

/* ---- context -------------------------------------------------------- */
int tg_ctx_create(int device, tg_ctx** out);
void tg_ctx_destroy(tg_ctx* ctx);
const char* tg_last_error(void);
/* Per-kernel HIP-event timing on the context stream (0 = off). */
int tg_set_profiling(tg_ctx* ctx, int on);
/* Number of profiled kernels; fills names (NUL-separated, cap bytes), total ms
 * and launch counts for up to n entries. */
int tg_kernel_times(tg_ctx* ctx, char* names, uint64_t cap, double* ms,
                    uint64_t* launches, int n);
void tg_reset_kernel_times(tg_ctx* ctx);

/* ---- device memory helpers (bench and tests keep inputs resident) ----- */
int tg_dev_alloc(tg_ctx* ctx, uint64_t bytes, void** dptr);
int tg_dev_free(tg_ctx* ctx, void* dptr);
int tg_h2d(tg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int tg_d2h(tg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int tg_sync(tg_ctx* ctx);

/* ---- whole job ------------------------------------------------------ */
/* Count the canonical k-mers of ACGT chunks: bases[off[i]..off[i+1]) is chunk
 * i (only [ACGTacgt] bytes), off has n_chunks+1 entries.  Host pointers; the
 * data is copied to the device.  Replaces partition_kmers_mem +
 * count_and_write_mem (src/pipeline.hpp:111-178). */
int tg_count_host(tg_ctx* ctx, const tg_params* p, const uint8_t* bases,
                  uint64_t n_bases, const uint64_t* off, uint64_t n_chunks,
                  tg_stats* st);
/* Count the canonical k-mers of FASTA/FASTQ files (gzip when a name ends in
 * ".gz"; BGZF files are inflated on `threads` threads, other gzip files on one
 * thread while the calling thread parses what has arrived).  Each file's bases
 * go to the device as they are parsed, so reading and the upload overlap; the
 * result equals tg_count_host on the files' chunks concatenated in order.
 * Replaces the reference's run over its input files (src/pipeline.hpp:64-178,
 * reading through src/seq_source.hpp:190-205).  Input errors carry the
 * reference's messages ("Cannot open file: <path>", "Empty file: <path>",
 * "Cannot open: <path>", "Cannot read: <path>"). */
int tg_count_files(tg_ctx* ctx, const tg_params* p, const char* const* paths, uint32_t n_files,
                   uint32_t threads, tg_stats* st);
/* Same with bases/off already resident in device memory (d_bases, d_off). */
int tg_count_device(tg_ctx* ctx, const tg_params* p, const uint8_t* d_bases,
                    uint64_t n_bases, const uint64_t* d_off, uint64_t n_chunks,
                    tg_stats* st);
/* How a context runs single jobs.  Default (0): through the batched engine as
 * a batch of one job (one captured launch sequence, no host round trip before
 * the result).  1: the plain launch sequence, driven from the host between
 * its steps (calibration and sizing read back in between), as
 * partition_kmers_mem then count_and_write_mem run in the reference. */
int tg_set_plain_sequence(tg_ctx* ctx, int on);
/* Debug statistics per reference partition (after src/count.hpp:42-60 and
 * emit_debug_stats :478-565, which the reference collects under -dbg).  With
 * on = 1, jobs counted with the plain sequence record, per partition: k-mers
 * inserted (with multiplicity), distinct keys held in the LDS tables, k-mers
 * sent to the HBM spill table (the reference's n_overflow) and the table slots
 * used (work items x LDS table size; tables are fixed, never resized).  Work
 * items then never span two partitions; results are unchanged. */
int tg_set_debug_stats(tg_ctx* ctx, int on);
/* Entries of the per-lane deque that the superkmer pass of k > 64 keeps for
 * the window minimum (1..16, default 16).  A strip whose deque would grow past
 * it is counted again by the pass that keeps the whole window, so the totals
 * do not depend on it; small values send most strips there (a test of that
 * pass).  No counterpart in the reference, whose superkmer pass keeps the
 * window (src/superkmer_helpers.hpp). */
int tg_set_wide_deque(tg_ctx* ctx, uint32_t entries);
/* Bytes of input that tg_count_files keeps on the device (0, the default: half
 * the free device memory at the call).  Larger inputs of k <= 64 are read into
 * host memory and counted slice by slice (tg_batch_config.slice_bases), each
 * slice uploaded, counted and added into the device merge table, so device
 * memory bounds the distinct k-mers, not the input.  Replaces the reference's
 * choice of its disk pipeline for inputs larger than memory
 * (src/pipeline.hpp:77-98,186-263). */
int tg_set_input_budget(tg_ctx* ctx, uint64_t bytes);
/* Copies the statistics of the last such job; returns the number of
 * partitions, TG_ECAP when n is smaller, TG_EINVAL when none were recorded.
 * Null arrays are skipped. */
int64_t tg_debug_partitions(tg_ctx* ctx, uint64_t* inserted, uint64_t* lds_unique, uint64_t* spilled,
                            uint64_t* table_cap, uint64_t n);
/* Minimizer coverage histogram (src/count.hpp:84-171, 665-705 under -dbg): for
 * every distinct minimizer (canonical ntHash of the m-mer at a superkmer's
 * minimizer position) the number of k-mers of its superkmers; returns the
 * number of distinct coverage values, with coverage[i] ascending and
 * n_minimizers[i] minimizers having it (TG_ECAP when cap is smaller).  Runs
 * Phase 1 again and groups on the host. */
int64_t tg_debug_min_coverage(tg_ctx* ctx, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                              const uint64_t* off, uint64_t n_chunks, uint32_t* coverage,
                              uint64_t* n_minimizers, uint64_t cap);
/* Result of the last job: n_out (key, count) pairs in unspecified order. */
uint64_t tg_result_size(tg_ctx* ctx);
/* keys: n*ceil(k/32) words; counts: n.  Copies min(n, cap) pairs. */
int tg_result_copy(tg_ctx* ctx, uint64_t* keys, uint32_t* counts, uint64_t cap);

/* ---- phase-level entry points (parity tests, tools) ------------------ */
/* 2-bit pack of ASCII ACGT: word j holds bases 16j..16j+15, first base in
 * bits 31:30, kache codes (src/partition_hash.hpp:67-70). */
int tg_pack_host(tg_ctx* ctx, const uint8_t* bases, uint64_t n_bases, uint32_t* out_words);
/* Phase 1 only: superkmers of every chunk (src/partition_hash.hpp:48-101).
 * For superkmer i: start = global base index of its first base, len = bases,
 * minpos = minimizer offset from start, pid = minimizer hash & (n_subparts-1).
 * Order is unspecified.  *n_sk receives the count (even if > cap). */
int tg_superkmers_host(tg_ctx* ctx, const tg_params* p, const uint8_t* bases,
                       uint64_t n_bases, const uint64_t* off, uint64_t n_chunks,
                       uint64_t* start, uint32_t* len, uint32_t* minpos,
                       uint32_t* pid, uint64_t cap, uint64_t* n_sk,
                       uint32_t* n_subparts);
/* Phase 1 + routing: superkmers grouped by reference partition
 * (hash & (n_partitions-1)); part_off has n_partitions+1 entries.  Each record
 * is serialised in the reference byte format [len][min_pos][packed]
 * (src/superkmer_io.hpp:154-172) into rec_bytes (rec_off per record,
 * n_sk+1 entries).  Replaces partition_kmers_mem's per-partition buffers. */
int tg_partitions_host(tg_ctx* ctx, const tg_params* p, const uint8_t* bases,
                       uint64_t n_bases, const uint64_t* off, uint64_t n_chunks,
                       uint64_t* part_off, uint8_t* rec_bytes, uint64_t bytes_cap,
                       uint64_t* rec_off, uint64_t rec_cap, uint64_t* n_sk,
                       uint64_t* n_bytes);

/* ---- many independent jobs per launch sequence ------------------------ */
/* The reference counts one sample per run (src/main.cpp:77 ->
 * src/pipeline.hpp:111-178); counting many samples is one pipeline each.  These
 * calls count many jobs together: job j is chunks [job_chunk[j], job_chunk[j+1])
 * of the input, and every job gets its own result, identical to counting it
 * alone with tg_count_host.  Jobs share k, m, ci, cx (n_partitions is not used:
 * the result does not depend on it). */
typedef struct {
    uint64_t n_bases;      /* ACGT bases of the job */
    uint64_t n_kmers;      /* reference total_inserted */
    uint64_t n_superkmers; /* reference "superkmers:" */
    uint64_t n_unique;     /* distinct canonical k-mers before ci/cx */
    uint64_t n_out;        /* reference "unique_kmers:" */
} tg_job_stats;

typedef struct {
    uint32_t n_jobs, n_batches;
    uint32_t n_graph_launches; /* batches run by replaying a captured hipGraph */
    uint32_t n_captures;       /* graphs captured (one per capacity class) */
    uint32_t n_retries;        /* batches re-run with larger overflow buffers */
    uint64_t n_kmers, n_unique, n_out;
    double ms_wall;            /* host wall time of the call */
} tg_many_stats;

/* Tunables of the batched path (tg_autotune_many picks them). */
typedef struct {
    uint32_t kmers_per_bucket; /* a job of B bases gets pow2(B / this) routing buckets */
    uint32_t max_tiles;        /* routing tiles of a batch */
    uint64_t batch_bases;      /* bases per batch (a larger job runs alone) */
    float load_factor;         /* LDS table fill target when splitting work items */
    uint32_t use_graph;        /* 1: replay one captured hipGraph per batch; 0: plain launches */
    float rho;                 /* distinct/k-mers estimate for splits; 0 = learned per batch */
    uint64_t slice_bases;      /* a single job (tg_count_host / tg_count_device) of more bases
                                  is counted in slices of whole chunks of at most this many
                                  bases, merged exactly on the device; 0 = 2^31 - 2^20 */
} tg_batch_config;

int tg_get_batch_config(tg_ctx* ctx, tg_batch_config* cfg);
int tg_set_batch_config(tg_ctx* ctx, const tg_batch_config* cfg);

/* Autotuner: times candidate settings (bucket size, batch size, load factor,
 * routing tiles) on the given device-resident jobs by coordinate search, checks
 * every candidate reproduces the same per-job totals, installs the fastest in
 * ctx and returns it (best_ms: wall time of one call with it). */
int tg_autotune_many(tg_ctx* ctx, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                     const uint64_t* d_off, uint64_t n_chunks, const uint64_t* job_chunk,
                     const uint64_t* job_base, uint32_t n_jobs, tg_batch_config* best,
                     double* best_ms);

/* Host inputs: bases/off as for tg_count_host plus the job table (n_jobs+1
 * entries).  st (n_jobs entries) and ms are optional. */
int tg_count_many(tg_ctx* ctx, const tg_params* p, const uint8_t* bases, uint64_t n_bases,
                  const uint64_t* off, uint64_t n_chunks, const uint64_t* job_chunk,
                  uint32_t n_jobs, tg_job_stats* st, tg_many_stats* ms);
/* Device-resident inputs.  job_chunk and job_base (n_jobs+1 entries each) are
 * host arrays: job_base[j] = off[job_chunk[j]], so batches are planned without
 * reading device memory. */
int tg_count_many_device(tg_ctx* ctx, const tg_params* p, const uint8_t* d_bases, uint64_t n_bases,
                         const uint64_t* d_off, uint64_t n_chunks, const uint64_t* job_chunk,
                         const uint64_t* job_base, uint32_t n_jobs, tg_job_stats* st,
                         tg_many_stats* ms);
/* Results of the last tg_count_many* call, per job (unspecified order). */
uint64_t tg_many_result_size(tg_ctx* ctx, uint32_t job);
int tg_many_result_copy(tg_ctx* ctx, uint32_t job, uint64_t* keys, uint32_t* counts, uint64_t cap);
int tg_many_result_device(tg_ctx* ctx, uint32_t job, const uint64_t** d_keys,
                          const uint32_t** d_counts, uint64_t* n);

/* ---- serving: a request loop with continuous batching ----------------- */
/* The reference has no serving mode (one process run per count).  A server
 * keeps `lanes` engine contexts resident (stream, batch buffers, captured
 * graph, device input area, worker thread).  A free lane takes every queued
 * request that fits its share of the HBM budget, in arrival order, counts them
 * together (one job each: results identical to counting a request alone) and
 * copies each result to pinned host memory.  Requests arriving while lanes are
 * busy wait for the next free lane; with two or more lanes, batches overlap on
 * the device.  All requests of a server share k, m, ci, cx. */
typedef struct tg_server tg_server;

typedef struct {
    uint32_t lanes;           /* resident contexts / worker threads (default 2) */
    uint32_t max_batch_jobs;  /* requests per lane batch (default 64) */
    uint64_t max_batch_bases; /* bases per lane batch (default 2^27) */
    uint64_t hbm_budget;      /* device bytes for all lanes (0 = half the free HBM) */
} tg_server_config;

typedef struct {
    double submit_to_start_s; /* queueing */
    double submit_to_done_s;  /* latency: result in host memory */
    uint32_t lane, batch_jobs;
} tg_request_times;

typedef struct {
    uint32_t lanes;
    uint64_t lane_bases;      /* admission limit of one lane batch */
    uint64_t hbm_budget, hbm_free;
    uint64_t queued, completed, failed, rejected, batches;
    uint64_t pinned_pool_bytes;
} tg_server_info;

int tg_server_create(int device, const tg_params* p, const tg_server_config* cfg, tg_server** out);
/* Finishes queued requests, then stops the lanes. */
void tg_server_destroy(tg_server* s);
/* Queue a request: bases/off as for tg_count_host (off[0] = 0, off[n_chunks] =
 * n_bases).  The caller keeps both arrays alive until the request is done (pinned
 * memory from tg_host_alloc makes the upload faster).  TG_ENOMEM when the
 * request exceeds a lane's admission limit. */
int tg_server_submit(tg_server* s, const uint8_t* bases, uint64_t n_bases, const uint64_t* off,
                     uint64_t n_chunks, uint64_t* id);
/* 1 done, 0 queued or running, < 0 the request's error. */
int tg_server_poll(tg_server* s, uint64_t id);
/* Block until done (timeout_s < 0: no limit); TG_ETIMEOUT otherwise. */
int tg_server_wait(tg_server* s, uint64_t id, double timeout_s);
/* Result of a finished request, lent until tg_server_release: n (key, count)
 * pairs (keys n*ceil(k/32) words), stats and timings (each optional). */
int tg_server_result(tg_server* s, uint64_t id, const uint64_t** keys, const uint32_t** counts,
                     uint64_t* n, tg_job_stats* st, tg_request_times* t);
int tg_server_release(tg_server* s, uint64_t id);
int tg_server_stats(tg_server* s, tg_server_info* st);
/* Pinned host memory for request inputs. */
int tg_host_alloc(uint64_t bytes, void** p);
int tg_host_free(void* p);

/* ---- host input path (A1) --------------------------------------------- */
/* Parse FASTA/FASTQ bytes (format by first byte, '@' = FASTQ) into ACGT
 * chunks with the reference parser's rules (src/seq_source.hpp:33-39,169-185).
 * out_bases needs len bytes; out_off n_chunks+1 entries.  Returns chunk count
 * or a negative status (TG_ECAP when off_cap is short, TG_EINPUT on empty). */
int64_t tg_parse_buffer(const uint8_t* data, uint64_t len, uint8_t* out_bases,
                        uint64_t* out_off, uint64_t off_cap);
/* The same parse on `threads` host threads, with the same output (bytes,
 * offsets and status) as tg_parse_buffer: the input is cut into sections at
 * FASTQ record starts (every fourth line) or anywhere in FASTA (a chunk that
 * crosses a cut is joined); inputs whose FASTQ records do not follow the
 * lines are parsed sequentially.  Replaces the reference's parallel readers
 * (src/seq_source.hpp:48-205). */
int64_t tg_parse_buffer_mt(const uint8_t* data, uint64_t len, uint8_t* out_bases, uint64_t* out_off,
                           uint64_t off_cap, uint32_t threads);

/* Read a whole input file into a malloc'd buffer (gzip-compressed when the
 * name ends in ".gz", as src/seq_source.hpp:196-204).  Errors follow the
 * reference messages: "Cannot open file: <path>", "Empty file: <path>". */
int tg_read_file(const char* path, uint8_t** data, uint64_t* len);
void tg_free(void* p);
/* Parse one input file the way tg_count_files reads it: gzip inflation
 * overlapped with a parse in sections of `section_bytes` of text (0: 32 MiB),
 * BGZF members inflated on `threads` threads, plain files mapped and parsed
 * on `threads` threads.  *bases (n_bases bytes) and *off (the returned chunk
 * count + 1 entries) are malloc'd; release both with tg_free.  The output
 * equals tg_parse_buffer on the file's whole text.  Returns the chunk count
 * or a negative status. */
int64_t tg_parse_file(const char* path, uint32_t threads, uint64_t section_bytes, uint8_t** bases,
                      uint64_t* n_bases, uint64_t** off);
/* A plain (not gzip) input file as a read-only mapping, parsed in place with
 * no copy; same error messages as tg_read_file.  Release with tg_unmap_file. */
int tg_map_file(const char* path, const uint8_t** data, uint64_t* len);
void tg_unmap_file(const uint8_t* data, uint64_t len);

/* Reference auto-tuned partition count (src/setup.hpp:82-196): depends on the
 * host's L2/L3 sizes, RLIMIT_NOFILE, the input sizes, k and threads.  Not yet
 * rounded to a power of two (src/main.cpp:47-52 does that). */
uint32_t tg_auto_partitions(const char* const* files, uint32_t n_files, uint32_t k,
                            uint32_t threads);
/* The same formula with every host-dependent input given: the input files'
 * sizes in bytes and gzip flags (a gzip file counts 6x its size), the L2 cache
 * of one core and the L3 cache in bytes (the reference reads them from sysfs
 * cpu0/cache, src/setup.hpp:14-51) and the open-file limit (RLIMIT_NOFILE soft
 * limit; the reference uses 1024 when it is infinite or unknown). */
uint32_t tg_auto_partitions_with(const uint64_t* sizes, const uint8_t* gz, uint32_t n_files, uint32_t k,
                                 uint32_t threads, uint64_t l2_bytes, uint64_t l3_bytes,
                                 uint64_t fd_limit);

/* ---- output formats (A10) ---------------------------------------------- */
/* TSV lines "<k-mer>\t<count>\n" (src/count.hpp:180-216) for n pairs; returns
 * bytes written or TG_ECAP.  cap >= n*(k+12) always suffices. */
int64_t tg_format_tsv(uint32_t k, const uint64_t* keys, const uint32_t* counts, uint64_t n,
                      char* out, uint64_t cap);
/* KFF file with one raw section (src/kff_output.hpp:20-57): encoding A0 C1 G2 T3,
 * unique, canonical, k / max=1 / data_size=4, each k-mer as ceil(k/4) bytes
 * MSB-first followed by its count as 4 big-endian bytes. */
int tg_write_kff(const char* path, uint32_t k, const uint64_t* keys, const uint32_t* counts,
                 uint64_t n);

/* The same formats built on the device from the context's last single-job
 * result (src/count.hpp:180-265 write_counts / write_counts_kff, whose
 * per-k-mer formatting they replace); the host only copies bytes.
 *   tg_result_tsv: TSV lines of result entries [first, first + n) into out
 *     (cap bytes); returns the byte count, TG_ECAP if it exceeds cap, or with
 *     out == NULL just the byte count.
 *   tg_result_kff: the KFF raw-section records of those entries, ceil(k/4)
 *     key bytes + 4-byte big-endian count each (no file header).
 *   tg_result_write: the whole result as a TSV file (kff = 0) or a KFF file
 *     (kff != 0, the bytes of tg_write_kff), formatted on the device in chunks
 *     whose copies overlap the file writes.
 * k must match the result's key width; output order is the result order. */
int64_t tg_result_tsv(tg_ctx* ctx, uint32_t k, uint64_t first, uint64_t n, char* out, uint64_t cap);
int64_t tg_result_kff(tg_ctx* ctx, uint32_t k, uint64_t first, uint64_t n, uint8_t* out,
                      uint64_t cap);
int tg_result_write(tg_ctx* ctx, uint32_t k, const char* path, int kff);
/* tg_result_tsv for callers that cannot allocate device memory (serving loops
 * that stream a result in chunks, code that must not call hipMalloc): the same
 * arguments, bytes and errors, for at most 262144 lines per call (TG_EINVAL
 * beyond).  Its device scratch is one fixed array of about 20 MB per GPU, held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU race.  No counterpart in the reference, which has no
 * device memory. */
int64_t tg_result_tsv_noalloc(tg_ctx* ctx, uint32_t k, uint64_t first, uint64_t n, char* out, uint64_t cap);

/* ---- synthetic inputs (seeded, deterministic) ------------------------- */
/* Generate the FASTA/FASTQ text of a named workload into a library-owned
 * buffer: "genome_fa" (uniform genome), "reads_fq" (short reads from a uniform
 * genome with substitutions and N), "repeat_reads" (repetitive genome reads).
 * Parameters are read from the key=value string spec.  Returns the size in
 * bytes; tg_synth_copy copies it out. */
int64_t tg_synth(const char* spec);
int tg_synth_copy(uint8_t* dst, uint64_t cap);
void tg_synth_free(void);
/* Thread-safe form: the text goes to a malloc'd buffer (release with tg_free). */
int64_t tg_synth_alloc(const char* spec, uint8_t** data);

#ifdef __cplusplus
}
#endif
#endif
