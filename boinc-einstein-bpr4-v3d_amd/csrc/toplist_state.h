// toplist_state.h -- the host toplist behind the opaque brp_toplist handle
// (include/brp/workload.hpp:16-23: five levels of `rows` rows, power-descending per
// level), shared by the host science (science.cpp) and the device bank loop (engine.hip),
// which uploads and downloads its rows.
#pragma once
#include <cstdint>
#include <vector>

#include "brp_hip.h"

struct brp_toplist {
    uint32_t rows = 100;
    std::vector<brp_bank_row> table;  // 5 levels x rows, power-descending per level
};
