/*
 * oracle.h — CPU restatement of tuna's two-phase k-mer counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * library; it is the checker that tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py compare the HIP path against.
 *
 * Every function names the reference file:line it restates.  Status of each
 * function's pinning (what independent check holds it fixed) is listed in the
 * header comment of oracle.c.
 */
#ifndef TUNA_ORACLE_H
#define TUNA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A1: FASTA/FASTQ bytes -> maximal ACGT runs ("chunks").  Chunk bytes are
 * appended to out_bases (capacity >= len); out_off receives n_chunks+1
 * offsets (out_off[0] = 0).  Returns n_chunks, -1 if off_cap is too small,
 * -2 for an empty buffer. */
int64_t orc_parse(const uint8_t* data, uint64_t len, uint8_t* out_bases,
                  uint64_t* out_off, uint64_t off_cap);

/* A2 encoders on one ASCII byte. */
uint32_t orc_kache_code(uint8_t c);
uint32_t orc_nt_code(uint8_t c);

/* A3: canonical ntHash of one m-mer, closed form (no rolling). */
uint64_t orc_nthash_direct(const uint8_t* ascii, int m);
/* A3: canonical ntHash of every m-mer of an ACGT string, rolling form. */
void orc_nthash_roll(const uint8_t* ascii, uint64_t len, int m, uint64_t* out);

/* A4: window minimum over W = k-m+1 m-mer hashes for every k-mer of a chunk
 * with M m-mers (M - W + 1 k-mers).  H[t] = minimizer hash, pos[t] =
 * chunk-relative m-mer position of the minimizer (tie rule of the reference).
 *   orc_window_stream  — restates the streaming two-stack object;
 *   orc_window_brute   — applies the normative rule directly per window. */
void orc_window_stream(const uint64_t* h, uint64_t M, int W, uint64_t* H, uint32_t* pos);
void orc_window_brute(const uint64_t* h, uint64_t M, int W, uint64_t* H, uint32_t* pos);

/* A5: superkmers of every chunk.  For superkmer i: sk_start = global base
 * index of its first base (chunk offset + t0), sk_len = bases,
 * sk_minpos = minimizer offset from t0, sk_hash = minimizer hash.
 * Returns the number of superkmers (-1 if cap too small). */
int64_t orc_superkmers(const uint8_t* bases, const uint64_t* off, uint64_t n_chunks,
                       int k, int m, uint64_t* sk_start, uint32_t* sk_len,
                       uint32_t* sk_minpos, uint64_t* sk_hash, uint64_t cap,
                       uint64_t* kmers_out);

/* A6: tuna's partition record [len][min_pos][ceil(len/4) bytes MSB-first]
 * with hdr_t = u8 when 2k-m <= 255, else u16 (little-endian).  Returns bytes
 * written. */
uint64_t orc_pack_record(const uint8_t* ascii, uint32_t len, uint32_t min_pos,
                         int k, int m, uint8_t* out);

/* A8: canonical k-mer words (kache 2-bit, first base most significant, word 0
 * = low 64 bits) of the k-mer starting at ascii.  nw = ceil(k/32) <= 2. */
void orc_canonical(const uint8_t* ascii, int k, uint64_t* words);

/* A9: brute-force counting of canonical k-mers over every chunk of length
 * >= k.  Writes the sorted unique keys (nw words each, word 0 first) and
 * counts, filtered by ci <= count <= cx.  Returns the number written, or -1
 * if cap is too small.  *total_kmers receives the number of k-mers seen and
 * *n_unique the unfiltered distinct count. */
int64_t orc_count(const uint8_t* bases, const uint64_t* off, uint64_t n_chunks,
                  int k, uint32_t ci, uint64_t cx, uint64_t* keys,
                  uint32_t* counts, uint64_t cap, uint64_t* total_kmers,
                  uint64_t* n_unique);

/* A5-A11 together on `threads` threads: superkmers routed to n_parts
 * partitions (hash & (n_parts-1)), then one counting table per partition
 * (cpu_count.c).  Returns the number of k-mers with ci <= count <= cx, writing
 * them (unordered) when keys != NULL, -1 on failure or a short cap, -2 for bad
 * parameters.  stats[0..2] = k-mers, superkmers, distinct; phase_s[0..1] =
 * wall seconds of the two phases. */
int64_t orc_count_parallel(const uint8_t* bases, const uint64_t* off, uint64_t n_chunks, int k, int m,
                           uint32_t n_parts, uint32_t ci, uint64_t cx, int threads, uint64_t* keys,
                           uint32_t* counts, uint64_t cap, uint64_t* stats, double* phase_s);

/* A10: TSV line "<label>\t<count>\n" for one key. Returns bytes written. */
uint64_t orc_tsv_line(const uint64_t* words, int k, uint32_t count, char* out);

/* A12: round up to a power of two (reference round_pow2). */
uint32_t orc_round_pow2(uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
