// pgd_reader.hpp -- the host half of the graph reader (not public ABI).
#pragma once
#include <cstdint>
#include <vector>

#include "pgd_hip.h"

namespace pgd {

// A file's edge lines after the reference reader's text rules (parse,
// carried values, start-index fix), before the self-loop removal and, for
// edge lists, the first-appearance renumbering.
// Usually (every line sets both of its fields) v, u hold one entry per edge
// line, self loops included, still to be renumbered (edge lists) or checked
// (.mtx).  When some line leaves a field unset, its value comes from the
// previous line's variables after their in-place decrement and renumbering
// (graphlet_core.cpp:413-424, 505-506); those files are replayed line by
// line on the host and final_ids is set: v, u are the final ids without self
// loops, with n and loops.
struct RawLines {
    bool mtx = false;          // .mtx rules: no renumbering, no duplicate removal
    int64_t maxv = 0;          // edge lists: the reference's lookup table holds raw ids [0, maxv]
    std::vector<int32_t> v, u; // see above
    bool final_ids = false;
    int64_t n = 0, loops = 0;  // with final_ids
};

// Parse `path` with `threads` host threads (0: the hardware's count).
// PGD_ERR_INVALID for a missing file.
pgd_status read_raw_lines(const char* path, int threads, RawLines& out);

}  // namespace pgd
