// host_checks.cpp -- CPU checks of libprocell's host-only code, built by
// tests/test_sanitizers.py with -fsanitize=address,undefined together with
// cuda-pro-cell_amd/csrc/procell_setup.cpp, procell_formats.cpp and the C
// oracle (oracle/procell_oracle.c, the checker).
//
// usage: host_checks <golden dir> <scratch dir>
//
//   * set-up (build_setup), the sizing entry points and the event balance
//     (fill_stats) on the golden inputs, on generated problems and on every
//     documented error case;
//   * the merge (merge_tally) and the writer against the oracle's own
//     file-level run: same files, seed, t_max and phi -> the same text;
//   * the segment split (split_segments): boundaries ordered, every segment's
//     row holds its first cell;
//   * the readers on malformed, binary and missing files (no crash, no
//     out-of-bounds access, sane counts).
// Exit status 0 when every check held.
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "procell.h"
#include "procell_oracle.h"
#include "procell_setup.hpp"

using namespace procell;

static int g_failures = 0;
static int g_checks = 0;

#define CHECK(c)                                                                    \
    do {                                                                            \
        ++g_checks;                                                                 \
        if (!(c)) {                                                                 \
            fprintf(stderr, "CHECK failed at %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++g_failures;                                                           \
        }                                                                           \
    } while (0)

static std::string read_text(const std::string &path)
{
    std::string s;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f)
        return s;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0)
        s.append(buf, n);
    fclose(f);
    return s;
}

static void write_text(const std::string &path, const std::string &text)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fwrite(text.data(), 1, text.size(), f);
    fclose(f);
}

struct Files {
    std::string hist, types;
};

// A generated problem written as the two input files (%.17g round-trips).
static Files write_problem(const std::string &dir, int id, std::mt19937_64 &rng, uint32_t n_rows, uint32_t n_types,
                           uint64_t max_count)
{
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::string h, t;
    char line[128];
    for (uint32_t i = 0; i < n_rows; ++i) {
        const double v = pow(10.0, 4.0 * u01(rng));
        const uint64_t c = (rng() % 4 == 0) ? 0 : rng() % (max_count + 1);  // zero-count rows too
        snprintf(line, sizeof line, "%.17g %" PRIu64 "\n", v, c);
        h += line;
    }
    std::vector<double> p(n_types);
    double sum = 0.0;
    for (uint32_t j = 0; j < n_types; ++j)
        sum += p[j] = 0.05 + u01(rng);
    double acc = 0.0;
    for (uint32_t j = 0; j < n_types; ++j) {
        // proportions written so that their left fold is 1 within 1e-8
        p[j] = j + 1 == n_types ? 1.0 - acc : p[j] / sum;
        acc += p[j];
        const bool quiescent = j + 1 == n_types && n_types > 1 && rng() % 2;
        const double m = quiescent ? -1.0 : 20.0 + 80.0 * u01(rng);
        const double sd = quiescent ? -1.0 : m * (0.1 + 0.4 * u01(rng));
        snprintf(line, sizeof line, "%.17g %.17g %.17g\n", p[j], m, sd);
        t += line;
    }
    Files f{dir + "/gen" + std::to_string(id) + "_hist.txt", dir + "/gen" + std::to_string(id) + "_types.txt"};
    write_text(f.hist, h);
    write_text(f.types, t);
    return f;
}

struct Loaded {
    procell_row *rows = nullptr;
    procell_type *types = nullptr;
    uint64_t n_rows = 0;
    uint32_t n_types = 0;
    ~Loaded()
    {
        procell_free(rows);
        procell_free(types);
    }
};

static bool load(const Files &f, Loaded &L)
{
    return procell_read_histogram(f.hist.c_str(), &L.rows, &L.n_rows) == PROCELL_OK &&
           procell_read_types(f.types.c_str(), &L.types, &L.n_types) == PROCELL_OK;
}

static procell_params params_of(const Loaded &L, double t_max, double phi, uint64_t seed, int track)
{
    procell_params p;
    memset(&p, 0, sizeof p);
    p.rows = L.rows;
    p.n_rows = L.n_rows;
    p.types = L.types;
    p.n_types = L.n_types;
    p.t_max = t_max;
    p.phi = phi;
    p.seed = seed;
    p.track_ratio = track;
    return p;
}

// Merge + writer against the oracle's file-level run; set-up against the
// oracle's phi and generation cap; the event balance against its stats.
static void check_run(const Files &f, const std::string &dir, double t_max, double phi, uint64_t seed, int track)
{
    Loaded L;
    CHECK(load(f, L));
    const procell_params p = params_of(L, t_max, phi, seed, track);
    Setup s;
    const int rc = build_setup(&p, s, true);
    CHECK(rc == PROCELL_OK);
    if (rc != PROCELL_OK)
        return;
    std::vector<double> vals(L.n_rows), prop(L.n_types), mean(L.n_types), sd(L.n_types);
    std::vector<uint64_t> cnts(L.n_rows);
    for (uint64_t i = 0; i < L.n_rows; ++i) {
        vals[i] = L.rows[i].value;
        cnts[i] = L.rows[i].count;
    }
    for (uint32_t j = 0; j < L.n_types; ++j) {
        prop[j] = L.types[j].proportion;
        mean[j] = L.types[j].mean;
        sd[j] = L.types[j].sd;
    }
    oracle_problem op;
    op.values = vals.data();
    op.counts = cnts.data();
    op.n_rows = L.n_rows;
    op.proportion = prop.data();
    op.mean = mean.data();
    op.sd = sd.data();
    op.n_types = L.n_types;
    op.t_max = t_max;
    op.phi = phi;
    op.seed = seed;
    CHECK(oracle_resolve_phi(&op) == s.phi);
    CHECK(oracle_max_generation(&op) == s.kmax);

    const uint32_t T = s.n_types ? s.n_types : 1;
    const size_t tally_len = s.v.size() * (size_t)(s.kmax + 1) * T;
    std::vector<int64_t> tally(tally_len + 1, 0);
    uint64_t st[ORACLE_N_STATS] = {};
    CHECK(oracle_simulate(&op, tally.data(), nullptr, st, 1) == 0);

    // event balance: the host derives recorded and dropped from the counters
    unsigned long long cnt[kNumCounters] = {};
    cnt[kDivisions] = st[ORACLE_STAT_DIVISIONS];
    cnt[kLost] = st[ORACLE_STAT_LOST];
    cnt[kRejected] = st[ORACLE_STAT_REJECTED];
    cnt[kCells] = st[ORACLE_STAT_CELLS];
    procell_stats ps;
    fill_stats(cnt, tally.data(), tally_len, ps);
    CHECK(ps.recorded == st[ORACLE_STAT_RECORDED]);
    CHECK(ps.dropped == st[ORACLE_STAT_DROPPED]);
    CHECK(ps.cells == s.n_cells);

    uint64_t cap = 0;
    CHECK(procell_output_capacity(&p, &cap) == PROCELL_OK && cap == s.bins.size());
    uint64_t nz = 0;
    uint32_t kdim = 0;
    CHECK(procell_tally_shape(&p, &nz, &kdim) == PROCELL_OK && nz == s.v.size() && kdim == (uint32_t)s.kmax + 1);
    std::vector<double> value(cap + 1);
    std::vector<int64_t> count(cap + 1), type_count((cap + 1) * T);
    procell_result r;
    memset(&r, 0, sizeof r);
    r.capacity = cap;
    r.value = value.data();
    r.count = count.data();
    r.type_count = type_count.data();
    CHECK(merge_tally(s, tally.data(), &r) == PROCELL_OK);
    CHECK(r.n_out <= cap);
    int64_t total = 0;
    for (uint64_t i = 0; i < r.n_out; ++i) {
        total += r.count[i];
        CHECK(i == 0 || r.value[i - 1] < r.value[i]);
    }
    CHECK((uint64_t)total == st[ORACLE_STAT_RECORDED]);
    if (r.n_out) {  // one slot short of what is needed
        procell_result small = r;
        small.capacity = r.n_out - 1;
        CHECK(merge_tally(s, tally.data(), &small) == PROCELL_E_CAPACITY);
    }

    const std::string mine = dir + "/mine.txt", theirs = dir + "/oracle.txt";
    CHECK(procell_write_histogram(mine.c_str(), &r, L.n_types, track) == PROCELL_OK);
    CHECK(oracle_run_files(f.hist.c_str(), f.types.c_str(), t_max, phi, track, seed, theirs.c_str(), 1) == 0);
    const std::string a = read_text(mine), b = read_text(theirs);
    CHECK(a == b);
    if (a != b)
        fprintf(stderr, "  texts differ for %s (t_max %g, phi %g, seed %" PRIu64 ")\n", f.hist.c_str(), t_max, phi,
                seed);

    // segment split
    for (uint32_t g : {1u, 2u, 7u, 64u, 1000u}) {
        std::vector<double> cum(g);
        for (uint32_t k = 0; k < g; ++k)
            cum[k] = k + 1 == g ? 1.0 : (double)(k + 1) / g;
        std::vector<uint64_t> begin;
        std::vector<uint32_t> row_of;
        split_segments(s, cum, begin, row_of);
        CHECK(begin.size() == g + 1 && row_of.size() == g);
        CHECK(begin[0] == 0 && begin[g] == s.n_cells);
        for (uint32_t k = 0; k < g; ++k) {
            CHECK(begin[k] <= begin[k + 1]);
            if (begin[k] < s.n_cells) {
                CHECK(row_of[k] < s.v.size());
                CHECK(s.start[row_of[k]] <= begin[k] && begin[k] < s.start[row_of[k] + 1]);
            }
        }
    }
    CHECK(set_cost(s) >= (double)s.n_cells);
}

static procell_params one_row(const procell_row *row, uint64_t n_rows, const procell_type *t, uint32_t n_types)
{
    procell_params p;
    memset(&p, 0, sizeof p);
    p.rows = row;
    p.n_rows = n_rows;
    p.types = t;
    p.n_types = n_types;
    p.t_max = 100.0;
    return p;
}

static void check_errors()
{
    Setup s;
    const procell_type t1[] = {{1.0, 40.0, 10.0}};
    const procell_row r1[] = {{100.0, 10}};
    CHECK(build_setup(nullptr, s, true) == PROCELL_E_INVALID);
    procell_params p = one_row(r1, 1, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_OK && s.n_cells == 10 && s.phi == 100.0);
    p.t_max = -1.0;
    CHECK(build_setup(&p, s, true) == PROCELL_E_INVALID);
    p.t_max = NAN;
    CHECK(build_setup(&p, s, true) == PROCELL_E_INVALID);
    p.t_max = INFINITY;  // allowed
    CHECK(build_setup(&p, s, true) == PROCELL_OK);

    const procell_row rnan[] = {{NAN, 3}};
    p = one_row(rnan, 1, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_NONFINITE);
    const procell_type tnan[] = {{1.0, NAN, 1.0}};
    p = one_row(r1, 1, tnan, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_NONFINITE);
    const procell_type tsum[] = {{0.5, 40.0, 10.0}, {0.4, 20.0, 5.0}};
    p = one_row(r1, 1, tsum, 2);
    CHECK(build_setup(&p, s, true) == PROCELL_E_PROPORTIONS);
    const procell_type tzero[] = {{1.0, 0.0, 1.0}};
    p = one_row(r1, 1, tzero, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_ZERO_MEAN);
    p = one_row(r1, 1, nullptr, 0);
    CHECK(build_setup(&p, s, true) == PROCELL_E_PROPORTIONS);  // cells but no types
    p = one_row(nullptr, 0, nullptr, 0);
    CHECK(build_setup(&p, s, true) == PROCELL_OK && s.n_cells == 0);
    p = one_row(nullptr, 3, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_INVALID);

    const procell_row rover[] = {{1.0, 0xFFFFFFFFull}, {2.0, 1}};
    p = one_row(rover, 2, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_TOO_MANY_CELLS);
    const procell_row rwrap[] = {{1.0, ~0ull}, {2.0, 2}};
    p = one_row(rwrap, 2, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_TOO_MANY_CELLS);
    std::vector<procell_type> many(257, procell_type{1.0 / 257, 40.0, 10.0});
    p = one_row(r1, 1, many.data(), 257);
    CHECK(build_setup(&p, s, true) == PROCELL_E_TOO_MANY_CELLS);

    const procell_row rdeep[] = {{1e300, 1}};
    p = one_row(rdeep, 1, t1, 1);
    p.phi = 1e-300;
    CHECK(build_setup(&p, s, true) == PROCELL_E_DEPTH);
    p.phi = 1e300 / ldexp(1.0, 62);  // exactly 62 halvings allowed (f/2 > phi stops at 62)
    CHECK(build_setup(&p, s, true) == PROCELL_OK && s.kmax <= kMaxGeneration);

    const procell_row rzero[] = {{5.0, 0}, {0.0, 0}};
    p = one_row(rzero, 2, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_OK && s.n_cells == 0 && s.v.empty());
    const procell_row rphi0[] = {{0.0, 4}};
    p = one_row(rphi0, 1, t1, 1);
    CHECK(build_setup(&p, s, true) == PROCELL_E_PHI);  // auto phi = 0

    double phi = 0.0;
    const procell_row rmix[] = {{8.0, 0}, {3.0, 2}, {5.0, 1}};
    p = one_row(rmix, 3, t1, 1);
    CHECK(procell_resolve_phi(&p, &phi) == PROCELL_OK && phi == 3.0);
    CHECK(procell_resolve_phi(nullptr, &phi) == PROCELL_E_INVALID);
    uint64_t cap = 0;
    CHECK(procell_output_capacity(&p, nullptr) == PROCELL_E_INVALID);
    CHECK(procell_output_capacity(&p, &cap) == PROCELL_OK && cap == 2);  // 3 and 5 (5/2 < 3)
    CHECK(procell_write_histogram("/nonexistent-dir/x/y.txt", nullptr, 1, 0) != PROCELL_OK);
}

static void check_readers(const std::string &dir)
{
    struct Case {
        const char *name;
        std::string text;
    };
    std::string binary;
    std::mt19937_64 rng(7);
    for (int i = 0; i < 4096; ++i)
        binary.push_back((char)(rng() & 0xFF));
    const std::vector<Case> cases = {
        {"empty", ""},
        {"spaces", "   \n\n\t\n"},
        {"truncated", "1.5 10\n2.5"},
        {"garbage", "abc def\n1 2\n"},
        {"negative_count", "1.5 -3\n"},
        {"huge_line", std::string(100000, '9') + " 1\n"},
        {"binary", binary},
        {"types_short", "0.5 10\n"},
        {"many_numbers", "1 2 3 4 5 6 7 8 9 10 11 12 13\n"},
    };
    for (const Case &c : cases) {
        const std::string path = dir + "/bad_" + c.name + ".txt";
        write_text(path, c.text);
        procell_row *rows = nullptr;
        uint64_t n_rows = 0;
        CHECK(procell_read_histogram(path.c_str(), &rows, &n_rows) == PROCELL_OK);
        CHECK(rows != nullptr);
        for (uint64_t i = 0; i < n_rows; ++i)
            (void)rows[i].value;  // every reported row is addressable
        procell_free(rows);
        procell_type *types = nullptr;
        uint32_t n_types = 0;
        CHECK(procell_read_types(path.c_str(), &types, &n_types) == PROCELL_OK);
        CHECK(types != nullptr);
        for (uint32_t j = 0; j < n_types; ++j)
            (void)types[j].sd;
        procell_free(types);
    }
    procell_row *rows = nullptr;
    uint64_t n_rows = 7;
    CHECK(procell_read_histogram((dir + "/missing.txt").c_str(), &rows, &n_rows) == PROCELL_OK && n_rows == 0);
    procell_free(rows);
    CHECK(procell_read_histogram(nullptr, &rows, &n_rows) == PROCELL_E_INVALID);
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s <golden dir> <scratch dir>\n", argv[0]);
        return 2;
    }
    const std::string golden = argv[1], dir = argv[2];
    const Files readme{golden + "/readme_histogram.txt", golden + "/readme_types.txt"};
    int runs = 0;
    for (uint64_t seed : {1ull, 2ull, 12345ull}) {
        check_run(readme, dir, 168.0, 0.0, seed, 1);
        check_run(readme, dir, 1000.0, 0.0, seed, 0);
        runs += 2;
    }
    std::mt19937_64 rng(2024);
    for (int id = 0; id < 12; ++id) {
        const uint32_t n_rows = 1 + (uint32_t)(rng() % 40);
        const uint32_t n_types = 1 + (uint32_t)(rng() % 5);
        const Files f = write_problem(dir, id, rng, n_rows, n_types, 60);
        const double t_max = (id % 4 == 0) ? 0.0 : 50.0 + (double)(rng() % 200);
        const double phi = (id % 3 == 0) ? 0.0 : 0.5 + (double)(rng() % 20);
        check_run(f, dir, t_max, phi, 1000 + id, id % 2);
        ++runs;
    }
    check_errors();
    check_readers(dir);
    printf("host checks: %d checks, %d failed, %d runs against the oracle\n", g_checks, g_failures, runs);
    return g_failures ? 1 : 0;
}
