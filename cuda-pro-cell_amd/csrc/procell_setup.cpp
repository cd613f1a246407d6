// procell_setup.cpp -- host-only part of libprocell: validation and set-up
// (the reference's parse-time rules, src/io/parser.cu:46-151), the cost
// model behind the segment split, the merge of a raw tally into the output
// histogram (src/simulation/proliferation.cu:353-380), and the sizing entry
// points of the C ABI.  Compiled by the host compiler.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "procell_setup.hpp"

namespace procell {

// ---------------------------------------------------------------------------
// Validation and set-up (mirrors the reference's parse-time rules).
// ---------------------------------------------------------------------------
int generation_cap(double v, double phi)
{
    // proliferation.cu:323: divide only while f/2 > phi; f halves exactly.
    int k = 0;
    double f = v;
    while (f / 2 > phi) {
        f = f / 2;
        if (++k > kMaxGeneration)
            return k;
    }
    return k;
}

double resolve_phi(const procell_params *p)
{
    if (p->phi > 0.0)
        return p->phi;
    // parser.cu:80-96: smallest value among rows with a nonzero count.
    double best = 0.0;
    bool have = false;
    for (uint64_t i = 0; i < p->n_rows; ++i) {
        const procell_row &r = p->rows[i];
        if (r.count > 0 && (!have || r.value < best)) {
            best = r.value;
            have = true;
        }
    }
    return best;
}

int build_setup(const procell_params *p, Setup &s, bool want_bins)
{
    if (!p || (p->n_rows && !p->rows) || (p->n_types && !p->types))
        return PROCELL_E_INVALID;
    if (isnan(p->t_max) || p->t_max < 0.0)
        return PROCELL_E_INVALID;
    s = Setup();
    s.t_max = p->t_max;
    uint64_t total = 0, n_nz = 0;
    for (uint64_t i = 0; i < p->n_rows; ++i) {
        if (!isfinite(p->rows[i].value))
            return PROCELL_E_NONFINITE;
        const uint64_t c = p->rows[i].count;
        if (c == 0)
            continue;
        ++n_nz;
        if (total + c < total || total + c > kMaxCells)
            return PROCELL_E_TOO_MANY_CELLS;
        total += c;
    }
    if (n_nz > kMaxRows)
        return PROCELL_E_TOO_MANY_CELLS;
    if (p->n_types == 0)
        return total ? PROCELL_E_PROPORTIONS : PROCELL_OK;
    if (p->n_types > kMaxTypes)
        return PROCELL_E_TOO_MANY_CELLS;
    // parser.cu:46-66: proportions (left fold, file order) must sum to 1.
    double sum = 0.0;
    for (uint32_t j = 0; j < p->n_types; ++j) {
        const procell_type &t = p->types[j];
        if (!isfinite(t.proportion) || !isfinite(t.mean) || !isfinite(t.sd))
            return PROCELL_E_NONFINITE;
        sum = sum + t.proportion;
    }
    if (fabs(1.0 - sum) > 1.0 / pow(10.0, 8.0))
        return PROCELL_E_PROPORTIONS;
    for (uint32_t j = 0; j < p->n_types; ++j)
        if (p->types[j].mean == 0.0)
            return PROCELL_E_ZERO_MEAN;  // cell.cu:110 would leave the timer unset
    const double phi = resolve_phi(p);
    if (total > 0 && !(phi > 0.0))
        return PROCELL_E_PHI;
    s.phi = phi;
    s.n_cells = total;
    s.n_types = p->n_types;
    s.v.reserve(n_nz);
    s.start.reserve(n_nz + 1);
    s.row_info.reserve(n_nz);
    uint64_t acc = 0;
    for (uint64_t i = 0; i < p->n_rows; ++i) {
        const procell_row &r = p->rows[i];
        if (r.count == 0)
            continue;
        const int kc = generation_cap(r.value, phi);
        if (kc > kMaxGeneration)
            return PROCELL_E_DEPTH;
        s.kmax = std::max(s.kmax, kc);
        s.v.push_back(r.value);
        s.start.push_back(acc);
        s.row_info.push_back((uint32_t)kc | ((r.value >= phi ? 1u : 0u) << 8));
        acc += r.count;
    }
    s.start.push_back(acc);
    // parser.cu:184: descending proportion; ties keep file order.
    s.type_file.resize(s.n_types);
    for (uint32_t j = 0; j < s.n_types; ++j)
        s.type_file[j] = j;
    std::stable_sort(s.type_file.begin(), s.type_file.end(), [&](uint32_t x, uint32_t y) {
        return p->types[x].proportion > p->types[y].proportion;
    });
    // cell.cu:86-101: running sum in sorted order.
    double c = 0.0;
    for (uint32_t j = 0; j < s.n_types; ++j) {
        const procell_type &t = p->types[s.type_file[j]];
        c = c + t.proportion;
        s.type_cum.push_back(c);
        s.type_mean.push_back(t.mean);
        s.type_sd.push_back(t.sd);
    }
    if (want_bins) {
        // parser.cu:126-151: every v / 2^k >= phi, exact halving, distinct.
        for (double v : s.v) {
            double f = v;
            while (f >= phi) {
                s.bins.push_back(f);
                f = f / 2;
            }
        }
        std::sort(s.bins.begin(), s.bins.end());
        s.bins.erase(std::unique(s.bins.begin(), s.bins.end()), s.bins.end());
    }
    return PROCELL_OK;
}

// Lane-steps of one initial cell of a row: its root step (one lane of a
// full-width step, about what a division costs) plus the expected number of
// divisions, with division k+1 completing at D = tau0 (1 + U) + tau_1 + ... +
// tau_k approximated as normal (truncation ignored).  Only used to balance
// the workgroups' segments; claims and stealing even out its errors.
constexpr double kRootCost = 1.0;
double row_cost(const Setup &s, uint32_t row)
{
    const int kcap = (int)(s.row_info[row] & 0xFFu);
    double w = kRootCost;
    double p_prev = 0.0;
    for (uint32_t j = 0; j < s.n_types; ++j) {
        const double m = s.type_mean[j];
        const double sd = fabs(s.type_sd[j]);
        const double p = s.type_cum[j] - p_prev;
        p_prev = s.type_cum[j];
        if (m < 0.0)
            continue;
        double e = 0.0;
        for (int k = 0; k < kcap; ++k) {
            const double mu = 1.5 * m + k * m;
            const double var = (7.0 / 3.0) * sd * sd + m * m / 12.0 + k * sd * sd;
            const double z = (s.t_max - mu) / sqrt(var > 0.0 ? var : 1e-300);
            const double prob = 0.5 * erfc(-z / sqrt(2.0));
            const double add = ldexp(prob, k);
            e += add;
            if (prob < 1e-12)
                break;
        }
        w += p * e;
    }
    return w;
}

// Split the initial cells into contiguous segments whose cumulative costs end
// at the fractions cum[0] < cum[1] < ... < cum[g-1] = 1 of the total.
void split_segments(const Setup &s, const std::vector<double> &cum, std::vector<uint64_t> &begin,
                    std::vector<uint32_t> &row_of)
{
    const uint32_t g = (uint32_t)cum.size();
    const uint32_t n_rows = (uint32_t)s.v.size();
    begin.assign(g + 1, s.n_cells);
    row_of.assign(g, n_rows ? n_rows - 1 : 0);
    begin[0] = 0;
    std::vector<double> w(n_rows);
    double total = 0.0;
    for (uint32_t r = 0; r < n_rows; ++r) {
        w[r] = row_cost(s, r);
        total += w[r] * (double)(s.start[r + 1] - s.start[r]);
    }
    uint32_t next = 1;
    double acc = 0.0;
    for (uint32_t r = 0; r < n_rows && next < g; ++r) {
        const uint64_t n = s.start[r + 1] - s.start[r];
        const double row_total = w[r] * (double)n;
        while (next < g && acc + row_total >= total * cum[next - 1]) {
            const double want = total * cum[next - 1] - acc;
            uint64_t cut = s.start[r] + (uint64_t)(want / w[r]);
            if (cut > s.start[r + 1])
                cut = s.start[r + 1];
            if (cut < begin[next - 1])
                cut = begin[next - 1];
            begin[next++] = cut;
        }
        acc += row_total;
    }
    for (uint32_t i = 0; i < g; ++i) {
        const uint64_t c = begin[i] < s.n_cells ? begin[i] : (s.n_cells ? s.n_cells - 1 : 0);
        row_of[i] = (uint32_t)(std::upper_bound(s.start.begin(), s.start.end() - 1, c) - s.start.begin()) - 1;
    }
}

// Expected work of a set (sum of row costs over its cells).
double set_cost(const Setup &s)
{
    double w = 0.0;
    for (uint32_t r = 0; r < (uint32_t)s.v.size(); ++r)
        w += row_cost(s, r) * (double)(s.start[r + 1] - s.start[r]);
    return w;
}

void fill_stats(const unsigned long long *cnt, const int64_t *tally, size_t tally_len, procell_stats &st)
{
    memset(&st, 0, sizeof st);
    st.divisions = cnt[kDivisions];
    // recorded = tally sum; dropped from the event balance (every division
    // adds one cell; every cell ends recorded, dropped or lost)
    uint64_t recorded = 0;
    for (size_t i = 0; i < tally_len; ++i)
        recorded += (uint64_t)tally[i];
    st.recorded = recorded;
    st.dropped = cnt[kCells] + cnt[kDivisions] - recorded - cnt[kLost];
    st.lost = cnt[kLost];
    st.rejected = cnt[kRejected];
    st.cells = cnt[kCells];
    st.pool = cnt[kPoolCount];
}

int merge_tally(const Setup &s, const int64_t *tally, procell_result *r)
{
    const uint32_t T = s.n_types ? s.n_types : 1;
    const uint32_t kdim = (uint32_t)s.kmax + 1;
    const uint64_t n_rows = s.v.size();
    const uint64_t n_bins = s.bins.size();
    if (!r || (r->capacity && !(r->value && r->count)))
        return PROCELL_E_INVALID;
    // (row, generation) -> output value row_value * 2^-generation, exact
    // equality merge into the bin set (proliferation.cu:353-380 searches the
    // same set by exact equality).
    std::vector<int64_t> bc(n_bins, 0), bt((size_t)n_bins * T, 0);
    for (uint64_t row = 0; row < n_rows; ++row) {
        double f = s.v[row];
        for (uint32_t k = 0; k < kdim; ++k, f = f / 2) {
            const int64_t *cell = &tally[((size_t)row * kdim + k) * T];
            int64_t tot = 0;
            for (uint32_t j = 0; j < T; ++j)
                tot += cell[j];
            if (!tot)
                continue;
            auto it = std::lower_bound(s.bins.begin(), s.bins.end(), f);
            if (it == s.bins.end() || *it != f)
                continue;  // cannot happen: recorded values are >= phi
            const size_t b = (size_t)(it - s.bins.begin());
            bc[b] += tot;
            for (uint32_t j = 0; j < T; ++j)
                bt[b * T + j] += cell[j];
        }
    }
    uint64_t n_out = 0;
    for (uint64_t b = 0; b < n_bins; ++b)
        n_out += bc[b] > 0;
    r->n_out = n_out;
    if (n_out > r->capacity)
        return PROCELL_E_CAPACITY;
    uint64_t o = 0;
    for (uint64_t b = 0; b < n_bins; ++b) {
        if (bc[b] <= 0)
            continue;
        r->value[o] = s.bins[b];
        r->count[o] = bc[b];
        if (r->type_count)
            for (uint32_t j = 0; j < s.n_types; ++j)
                r->type_count[o * s.n_types + j] = bt[b * T + j];
        ++o;
    }
    return PROCELL_OK;
}

}  // namespace procell

using namespace procell;

extern "C" {

int procell_output_capacity(const procell_params *p, uint64_t *capacity)
{
    if (!capacity)
        return PROCELL_E_INVALID;
    Setup s;
    const int rc = build_setup(p, s, true);
    if (rc != PROCELL_OK)
        return rc;
    *capacity = s.bins.size();
    return PROCELL_OK;
}

int procell_tally_shape(const procell_params *p, uint64_t *nonzero_rows, uint32_t *kdim)
{
    if (!nonzero_rows || !kdim)
        return PROCELL_E_INVALID;
    Setup s;
    const int rc = build_setup(p, s, false);
    if (rc != PROCELL_OK)
        return rc;
    *nonzero_rows = s.v.size();
    *kdim = (uint32_t)s.kmax + 1;
    return PROCELL_OK;
}

int procell_resolve_phi(const procell_params *p, double *phi)
{
    if (!p || !phi || (p->n_rows && !p->rows))
        return PROCELL_E_INVALID;
    *phi = resolve_phi(p);
    return PROCELL_OK;
}

}  // extern "C"
