// Host-side input reading shared by the C ABI and the engine's file entry
// point.
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

namespace tgi {

// Receives bases as they become final: `n` bases that belong at offset `at`
// of the file's bases.  `pinned` says the memory is page-locked; it stays
// valid only during the call.  A nonzero return stops the read with that code.
using SectionFn = std::function<int(const uint8_t* bases, uint64_t n, uint64_t at, bool pinned)>;

// Reads one FASTA/FASTQ file (gzip when the name ends in ".gz") and hands its
// bases to `ready` in order; `off` receives the chunk ends (off[0] = 0) and
// *n_bases the total.  The output equals tg_parse_buffer on the file's text.
// `section` is the text size a streamed (gzip) parse waits for between
// sections (0: 32 MiB).
int parse_file_sections(const char* path, uint32_t threads, uint64_t section, const SectionFn& ready,
                        std::vector<uint64_t>& off, uint64_t* n_bases);

}  // namespace tgi
