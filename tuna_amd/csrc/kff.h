// kff.h — the fixed parts of a KFF file with one raw section, as written by
// src/kff_output.hpp:20-57: header, global variables and section header
// before the records, and the footer after them.
#pragma once
#include <cstdint>
#include <string>

namespace tgi {

// magic, version 1.0, encoding A0 C1 G2 T3, unique, canonical, no metadata,
// variables k / max = 1 / data_size = 4, then a raw section of n blocks
std::string kff_header(uint32_t k, uint64_t n);
// variables section "first_index" / "footer_size", then the magic
std::string kff_footer();

}  // namespace tgi
