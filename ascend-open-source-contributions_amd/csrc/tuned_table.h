// tuned_table.h -- plan table written by cgemm3m/autotune.py from MI355X measurements.
// Key: (floor(log2(m*n)), floor(log2(k)), grouped); value: output tile edge and K partitions.
// Bins not listed fall back to the planner's occupancy rule.
#pragma once

namespace c3m {

struct TunedEntry {
    signed char mnLog2, kLog2, grouped, tile, parts;
};

static const TunedEntry kTunedTable[] = {
    {0, 0, 0, 0, 0},  // sentinel (no measured entries yet)
};

}  // namespace c3m
