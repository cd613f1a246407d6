// procell_formats.cpp -- the reference's text formats.
//
// Histogram file: whitespace-separated "value frequency" pairs read with
// operator>> into (double, uint64) until a read fails (src/io/parser.cu:99-103).
// Types file: "proportion mean sd" triples (src/io/parser.cu:167).
// Output: one line per value with a nonzero count, ascending,
// "value<TAB>count[<TAB>count_type_j ...]" with the stream precision set to 10
// significant digits (src/io/parser.cu:187-217).
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <iostream>
#include <vector>

#include "procell.h"

extern "C" {

int procell_read_histogram(const char *path, procell_row **rows, uint64_t *n_rows)
{
    if (!path || !rows || !n_rows)
        return PROCELL_E_INVALID;
    std::vector<procell_row> v;
    std::ifstream in(path);
    double value = 0.0;
    uint64_t frequency = 0;
    while (in >> value >> frequency)
        v.push_back(procell_row{value, frequency});
    procell_row *out = (procell_row *)malloc((v.size() ? v.size() : 1) * sizeof(procell_row));
    if (!out)
        return PROCELL_E_OUT_OF_MEMORY;
    if (!v.empty())
        memcpy(out, v.data(), v.size() * sizeof(procell_row));
    *rows = out;
    *n_rows = v.size();
    return PROCELL_OK;
}

int procell_read_types(const char *path, procell_type **types, uint32_t *n_types)
{
    if (!path || !types || !n_types)
        return PROCELL_E_INVALID;
    std::vector<procell_type> v;
    std::ifstream in(path);
    double proportion = 0.0, mean = 0.0, sd = 0.0;
    while (in >> proportion >> mean >> sd)
        v.push_back(procell_type{proportion, mean, sd});
    procell_type *out = (procell_type *)malloc((v.size() ? v.size() : 1) * sizeof(procell_type));
    if (!out)
        return PROCELL_E_OUT_OF_MEMORY;
    if (!v.empty())
        memcpy(out, v.data(), v.size() * sizeof(procell_type));
    *types = out;
    *n_types = (uint32_t)v.size();
    return PROCELL_OK;
}

static void write_rows(std::ostream &os, const procell_result *r, uint32_t n_types, int32_t track_ratio)
{
    os.precision(10);
    for (uint64_t i = 0; i < r->n_out; ++i) {
        if (r->count[i] <= 0)
            continue;
        os << r->value[i] << "\t" << r->count[i];
        if (track_ratio && r->type_count)
            for (uint32_t j = 0; j < n_types; ++j)
                os << "\t" << r->type_count[i * n_types + j];
        os << "\n";
    }
    os.flush();
}

int procell_write_histogram(const char *path, const procell_result *r, uint32_t n_types, int32_t track_ratio)
{
    if (!r || (r->n_out && (!r->value || !r->count)))
        return PROCELL_E_INVALID;
    if (!path) {
        write_rows(std::cout, r, n_types, track_ratio);
        return std::cout.good() ? PROCELL_OK : PROCELL_E_IO;
    }
    std::ofstream os(path);
    if (!os)
        return PROCELL_E_IO;
    write_rows(os, r, n_types, track_ratio);
    return os.good() ? PROCELL_OK : PROCELL_E_IO;
}

void procell_free(void *ptr) { free(ptr); }

}  // extern "C"
