// procell_internal.hpp -- the device launch interface shared by the C-ABI
// (procell_api.hip) and the kernels (procell_kernels.hip); the host-only
// set-up it builds on is in procell_setup.hpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "procell.h"
#include "procell_setup.hpp"

namespace procell {

// Per-parameter-set values of a plan: one set for a plain run, many for a
// batch.  Rows, initial cells, type-table entries, tally entries and
// segments of all sets are concatenated in set order.
struct SetInfo {
    double t_max;
    uint64_t cell_base;   // first initial cell of the set (cell ids of the stream are relative to it)
    uint64_t tally_base;  // tally index of the set's first nonzero row
    uint32_t kdim;        // generations + 1
    uint32_t n_types;
    uint32_t type_base;   // first entry of the set in the type tables
    uint32_t row_base;    // first nonzero row of the set
    uint32_t pad[4];
};

// A dividing initial cell, buffered in LDS between a root step and the
// hand-out to an idle lane, with every field the lane's state needs (three
// 16-byte LDS reads, no dependent lookups).
struct QueueEntry {
    double s0;          // completion time of its first division
    double mean, sd;    // timer parameters of its type
    uint32_t cell;      // initial cell id, relative to the set
    uint32_t tb_lds;    // tally window index of (row, generation 0, type), or none
    uint32_t rowtf;     // row relative to the set's first row | file type << 24
    uint32_t kcap;      // generation cap of the row
    uint32_t pad[2];
};
static_assert(sizeof(QueueEntry) == 48, "three 16-byte LDS words");

// Copies of the counters block: workgroup b adds into copy b % kCounterCopies,
// so the device atomics on any one counter stay few; the host sums the copies.
constexpr uint32_t kCounterCopies = 64;

// Arguments of one run, passed by value to the kernel.
struct RunArgs {
    uint64_t seed;          // seed of a single-set run (n_sets == 1)
    uint64_t n_cells;
    uint32_t n_sets;
    uint32_t n_rows;
    uint32_t n_types;       // most types of any set
    uint32_t kdim;          // most generations + 1 of any set
    uint32_t stack_levels;  // per-lane stack depth (levels 1 .. kdim-1)
    uint32_t win_rows;      // rows of the LDS tally window (0: none)
    uint32_t idle_limit;    // idle lanes tolerated before a wavefront refills
    uint32_t rows_in_lds;   // nonzero: the row tables are copied to LDS
    const SetInfo *sets;         // [n_sets]
    const uint64_t *set_seed;    // [n_sets] seeds of a batch run (n_sets > 1)
    const uint32_t *seg_set;     // [segments] set of each segment
    const uint32_t *row_start;   // [n_rows + 1] first cell of each nonzero row (cells < 2^32)
    const uint32_t *row_info;    // [n_rows]: generation cap | (value >= phi) << 8
    const double *type_cum;      // [n_types] sorted order
    const double *type_mean;
    const double *type_sd;
    const uint32_t *type_file;   // sorted position -> file index
    const uint64_t *seg_begin;   // [segments] first initial cell of each workgroup's segment
    const uint64_t *seg_end;     // [segments] one past its last initial cell
    const uint32_t *seg_row;     // [segments] row of each segment's first cell (window base)
    unsigned long long *tally;   // per set [rows][kdim][n_types] of the set, file type order
    unsigned long long *counters;  // [n_sets][kCounterCopies][kNumCounters]
    double *gstack;              // [segments][stack_levels][block] lane stacks in device memory, or null: in LDS
};

struct LaunchShape {
    int blocks;             // = segments
    int block_size;
    size_t lds_bytes;
};

size_t sim_lds_bytes(uint32_t win_rows, uint32_t kdim, uint32_t n_types, uint32_t stack_levels, uint32_t block_size,
                     uint32_t lds_rows);
hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t st);
hipError_t launch_sim(const RunArgs &a, const LaunchShape &ls, hipStream_t st);
hipError_t launch_debug_draws(uint64_t seed, const uint64_t *cell, const uint64_t *node,
                              const uint32_t *attempt, uint64_t n, uint32_t *words, double *z,
                              hipStream_t st);
hipError_t sim_occupancy(int block_size, size_t lds_bytes, bool rows_in_lds, uint32_t kdim, bool gstack,
                         int *blocks_per_cu);

}  // namespace procell
