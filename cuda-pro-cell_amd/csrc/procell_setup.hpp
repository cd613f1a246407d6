// procell_setup.hpp -- host-only part of libprocell: validation and set-up
// of a problem, the cost model that splits initial cells into segments, and
// the merge of a raw tally into the output histogram.  No HIP: this code is
// compiled by the host compiler (procell_setup.cpp), so that it can also be
// built and checked on its own.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "procell.h"

namespace procell {

// Limits of the packed root entry (cell 32 bits, row 24, type 8).
constexpr uint64_t kMaxCells = 0xFFFFFFFFull;
constexpr uint64_t kMaxRows = 1ull << 24;
constexpr uint32_t kMaxTypes = 256;
constexpr int kMaxGeneration = 62;  // heap index of a node must fit 64 bits

// Event counters of a run (uint64 each; the device keeps several copies).
enum Counter : int {
    kPoolCount = 0,   // initial cells that divided at least once
    kDivisions = 1,
    kRecorded = 2,    // unused on the device: derived by the host
    kDropped = 3,     // unused on the device: derived by the host
    kLost = 4,
    kRejected = 5,
    kCells = 6,
    kNumCounters = 8,
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

int generation_cap(double v, double phi);
double resolve_phi(const procell_params *p);
int build_setup(const procell_params *p, Setup &s, bool want_bins);

// Expected lane-steps of one initial cell of row `row` and of a whole set.
double row_cost(const Setup &s, uint32_t row);
double set_cost(const Setup &s);
// Contiguous segments of initial cells whose cumulative expected costs end
// at the fractions cum[0] < ... < cum[g-1] = 1: begin[g+1], row_of[g].
void split_segments(const Setup &s, const std::vector<double> &cum, std::vector<uint64_t> &begin,
                    std::vector<uint32_t> &row_of);

// Event statistics from the summed counters (kNumCounters) and the raw
// tally; the launch fields of `st` are left zero.
void fill_stats(const unsigned long long *cnt, const int64_t *tally, size_t tally_len, procell_stats &st);
// Merge a raw tally [rows][kmax+1][max(T,1)] (file type order) into r's
// output histogram; PROCELL_E_CAPACITY when r->capacity is too small.
int merge_tally(const Setup &s, const int64_t *tally, procell_result *r);

}  // namespace procell
