// tuned_table.h -- plan table written by cgemm3m/autotune.py from MI355X measurements.
// Key: (floor(log2(m*n)), floor(log2(k)), grouped); value: output tile edge and K partitions
// (tile 32: the one-launch small-problem kernel).
// Bins not listed fall back to the planner's occupancy rule.
#pragma once

namespace c3m {

struct TunedEntry {
    short mnLog2, kLog2, grouped, tile, parts;
};

static const TunedEntry kTunedTable[] = {
    {10, 6, 1, 64, 1},
    {10, 8, 1, 64, 1},
    {10, 10, 1, 64, 2},
    {10, 11, 1, 64, 4},
    {12, 6, 0, 32, 1},
    {12, 6, 1, 64, 1},
    {12, 8, 1, 64, 1},
    {12, 10, 0, 32, 1},
    {12, 10, 1, 64, 4},
    {12, 11, 1, 64, 4},
    {14, 6, 1, 64, 1},
    {14, 7, 0, 32, 1},
    {14, 8, 1, 64, 1},
    {14, 9, 0, 32, 1},
    {14, 10, 1, 64, 1},
    {14, 11, 1, 64, 1},
    {14, 12, 0, 64, 16},
    {16, 6, 1, 128, 1},
    {16, 8, 0, 32, 1},
    {16, 8, 1, 128, 1},
    {16, 10, 1, 128, 1},
    {16, 11, 1, 128, 1},
    {16, 13, 0, 64, 16},
    {16, 16, 0, 64, 16},
    {17, 8, 0, 32, 1},
    {18, 6, 1, 128, 1},
    {18, 7, 0, 32, 1},
    {18, 8, 1, 128, 1},
    {18, 9, 0, 64, 2},
    {18, 10, 1, 128, 1},
    {18, 11, 1, 128, 1},
    {18, 14, 0, 128, 16},
    {19, 9, 0, 64, 2},
    {20, 8, 0, 64, 1},
    {20, 9, 0, 64, 1},
    {20, 10, 0, 128, 4},
    {20, 11, 0, 64, 2},
    {22, 8, 0, 128, 1},
    {22, 11, 0, 128, 1},
    {24, 9, 0, 128, 1},
    {24, 12, 0, 128, 1},
    {26, 13, 0, 128, 1},
    {0, 0, 0, 0, 0},  // sentinel
};

}  // namespace c3m
