// packed_presets.h -- lanes per column / row of the packed iteration kernels
// by LP shape, as measured by the project's autotuner (mipx_amd/autotune.py,
// which rewrote this file from its measurements on AMD Radeon Graphics).
// Buckets: b_c = round(log2(mean entries per column)) in 0..5,
// b_r = round(log2(mean entries per row)) in 0..7; -1 = not tuned (the
// nearest-shape heuristic decides).  Presets 0..7 are (lanes per column,
// lanes per row) = (1,1), (2,2), (2,4), (4,4), (4,8), (8,8), (8,16), (16,16).
// Measured (primal + dual kernel microseconds per iteration, every LP of the
// training batch active, presets 0..7):
//   (2,4): 81 59 47 53 49 51 57 69 -> preset 2 (2, 4)
//   (3,4): 79 54 43 42 38 43 46 50 -> preset 4 (4, 8)
//   (3,3): 54 39 33 30 29 30 35 40 -> preset 4 (4, 8)
//   (2,3): 54 40 34 34 32 37 42 48 -> preset 4 (4, 8)
//   (3,5): 266 164 115 114 94 90 88 96 -> preset 6 (8, 16)
#pragma once

#include <cmath>

namespace mpx {

constexpr int kTuneColBuckets = 6, kTuneRowBuckets = 8;

constexpr signed char kTunedPreset[kTuneColBuckets][kTuneRowBuckets] = {
    {-1, -1, -1, -1, -1, -1, -1, -1},
    {-1, -1, -1, -1, -1, -1, -1, -1},
    {-1, -1, -1,  4,  2, -1, -1, -1},
    {-1, -1, -1,  4,  4,  6, -1, -1},
    {-1, -1, -1, -1, -1, -1, -1, -1},
    {-1, -1, -1, -1, -1, -1, -1, -1},
};

inline int tuneBucket(double mean, int buckets) {
    if (!(mean > 0.0)) return 0;
    const long b = std::lround(std::log2(mean));
    return static_cast<int>(b < 0 ? 0 : (b >= buckets ? buckets - 1 : b));
}

// The tuned preset for a matrix with nnz entries, m rows, n columns (-1: none).
inline int tunedPreset(long long nnz, int m, int n) {
    if (m <= 0 || n <= 0) return -1;
    const int bc = tuneBucket(static_cast<double>(nnz) / n, kTuneColBuckets);
    const int br = tuneBucket(static_cast<double>(nnz) / m, kTuneRowBuckets);
    return kTunedPreset[bc][br];
}

}  // namespace mpx
