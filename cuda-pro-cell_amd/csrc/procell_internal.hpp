// procell_internal.hpp -- host-side problem set-up and the device launch
// interface shared by the C-ABI (procell_api.hip) and the kernels
// (procell_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "procell.h"

namespace procell {

// Limits of the packed work-queue entry (cell 32 bits, row 24, type 8).
constexpr uint64_t kMaxCells = 0xFFFFFFFFull;
constexpr uint64_t kMaxRows = 1ull << 24;
constexpr uint32_t kMaxTypes = 256;
constexpr int kMaxGeneration = 62;  // heap index of a node must fit 64 bits

// Counters block in device memory (uint64 each).
enum Counter : int {
    kPoolCount = 0,   // dividing initial cells appended by the root kernel
    kPoolNext = 1,    // next unclaimed queue entry
    kDivisions = 2,
    kRecorded = 3,
    kDropped = 4,
    kLost = 5,
    kRejected = 6,
    kCells = 7,
    kNumCounters = 8,
};

// One dividing initial cell, produced by the root kernel and consumed by the
// tree kernel: s0 = completion time of its first division.
struct alignas(16) PoolEntry {
    double s0;
    uint32_t cell;
    uint32_t row_type;  // nonzero row index | sorted type index << 24
};

// Problem after validation: nonzero rows in file order, types sorted by
// descending proportion (src/io/parser.cu:184), output bin set.
struct Setup {
    std::vector<double> v;          // value per nonzero row
    std::vector<uint64_t> start;    // first cell id per nonzero row, plus total at the end
    std::vector<uint32_t> row_info; // generation cap | (value >= phi) << 8
    uint64_t n_cells = 0;
    uint32_t n_types = 0;
    std::vector<uint32_t> type_file;  // sorted position -> file index
    std::vector<double> type_cum, type_mean, type_sd;  // sorted order
    double phi = 0.0;
    double t_max = 0.0;
    int kmax = 0;
    std::vector<double> bins;       // output bin set, ascending, distinct
};

int build_setup(const procell_params *p, Setup &s, bool want_bins);

// Arguments of one run, passed by value to the kernels.
struct RunArgs {
    uint64_t seed;
    double t_max;
    uint64_t n_cells;
    uint32_t n_rows;
    uint32_t n_types;
    uint32_t kdim;
    uint32_t pad;
    const uint64_t *row_start;   // [n_rows + 1]
    const uint32_t *row_info;    // [n_rows]
    const double *type_cum;      // [n_types] sorted order
    const double *type_mean;
    const double *type_sd;
    const uint32_t *type_file;
    unsigned long long *tally;   // [n_rows][kdim][n_types] file type order
    PoolEntry *pool;             // [n_cells]
    unsigned long long *counters;  // [kNumCounters]
};

struct LaunchShape {
    int roots_blocks;
    int roots_block_size;
    uint32_t roots_window_rows;
    int tree_blocks;
    int tree_block_size;
    uint32_t chunk;
    size_t tree_lds_bytes;
};

hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t st);
hipError_t launch_roots(const RunArgs &a, const LaunchShape &ls, hipStream_t st);
hipError_t launch_tree(const RunArgs &a, const LaunchShape &ls, hipStream_t st);
hipError_t launch_debug_draws(uint64_t seed, const uint64_t *cell, const uint64_t *node,
                              const uint32_t *attempt, uint64_t n, uint32_t *words, double *z,
                              hipStream_t st);
hipError_t tree_occupancy(int block_size, size_t lds_bytes, int *blocks_per_cu);
hipError_t roots_occupancy(int block_size, size_t lds_bytes, int *blocks_per_cu);

}  // namespace procell
