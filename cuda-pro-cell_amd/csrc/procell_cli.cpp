// procell_cli.cpp -- the `procell` command (src/main.cu:18-34,
// src/io/cmdargs.cpp:11-375).
//
// Options, messages (on stdout) and exit codes follow the reference's
// argument parser, with these decisions:
//   * -p/--phi-min is optional: without it phi is the smallest value with a
//     nonzero count, as README.md:98-102 documents (the reference's parser
//     made it mandatory, cmdargs.cpp:48-49).
//   * the README's spellings --output and --phi are accepted as aliases of
//     --output-histogram and --phi-min.
//   * --seed <u64> (default 0) makes runs reproducible; the reference seeds
//     from time(NULL).  --device <n> selects the HIP device.
//   * -d/--tree-depth is validated as in the reference and otherwise ignored:
//     the depth-first kernels have no level batching to bound.
//   * --help prints the reference's line followed by a usage summary.
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>

#include <iostream>
#include <string>
#include <vector>

#include "procell.h"

namespace {

constexpr unsigned kMaxTreeDepth = 23;

struct Args {
    bool h0_given = false, types_given = false, out_given = false, t_given = false;
    bool phi_given = false, depth_given = false, track_ratio = false, seed_given = false;
    bool device_given = false;
    std::string h0, types, out;
    double t_max = 0.0, phi = 0.0;
    uint64_t seed = 0;
    int device = 0;
};

[[noreturn]] void fail(const std::string &msg)
{
    std::cout << msg << std::endl;
    exit(EXIT_FAILURE);
}

bool file_option(int argc, char **argv, int &i, const char *lng, const char *shrt, const char *alias,
                 const char *label, bool &given, std::string &dst)
{
    const std::string s(argv[i]);
    if (s != lng && s != shrt && !(alias && s == alias))
        return false;
    if (given)
        fail(std::string("Option ") + label + " already given");
    if (i >= argc - 1)
        fail(std::string("Option ") + label + " requires a filename");
    dst = argv[++i];
    given = true;
    return true;
}

bool parse_t_max(int argc, char **argv, int &i, Args &a)
{
    const std::string s(argv[i]);
    if (s != "--time-max" && s != "-t")
        return false;
    if (a.t_given)
        fail("Option --time-max (-t) already given");
    if (i >= argc - 1)
        fail("Option --time-max (-t) requires an integer value >= 0");
    const double v = atof(argv[++i]);
    if (!(v >= 0))
        fail("Option --time-max (-t) requires an integer value >= 0");
    a.t_max = v;
    a.t_given = true;
    return true;
}

bool parse_phi(int argc, char **argv, int &i, Args &a)
{
    const std::string s(argv[i]);
    if (s != "--phi-min" && s != "-p" && s != "--phi")
        return false;
    if (a.phi_given)
        fail("Option --phi-min (-p) already given");
    if (i >= argc - 1)
        fail("Option --phi-min (-p) requires a double value > 0");
    const double v = atof(argv[++i]);
    if (!(v > 0))
        fail("Option --phi-min (-p) requires a double value > 0");
    a.phi = v;
    a.phi_given = true;
    return true;
}

bool parse_depth(int argc, char **argv, int &i, Args &a)
{
    const std::string s(argv[i]);
    if (s != "--tree-depth" && s != "-d")
        return false;
    if (a.depth_given)
        fail("Option --tree-depth (-d) already given");
    if (i >= argc - 1)
        fail("Option --tree-depth (-d) requires an integer value >= 1");
    int v = 0;
    if (sscanf(argv[++i], "%d", &v) != 1)
        v = 0;
    if (v < 1 || v > (int)kMaxTreeDepth)
        fail("Option --tree-depth (-d) requires an integer value >= 1 && <= " + std::to_string(kMaxTreeDepth));
    a.depth_given = true;
    return true;
}

bool parse_track_ratio(int, char **argv, int &i, Args &a)
{
    const std::string s(argv[i]);
    if (s != "--track-ratio" && s != "-r")
        return false;
    if (a.track_ratio)
        fail("Option --track-ratio (-r) already given");
    a.track_ratio = true;
    return true;
}

bool parse_seed(int argc, char **argv, int &i, Args &a)
{
    const std::string s(argv[i]);
    if (s != "--seed" && s != "-s")
        return false;
    if (a.seed_given)
        fail("Option --seed (-s) already given");
    if (i >= argc - 1)
        fail("Option --seed (-s) requires an unsigned integer value");
    char *end = nullptr;
    const char *txt = argv[++i];
    const unsigned long long v = strtoull(txt, &end, 0);
    if (!end || *end != '\0' || end == txt)
        fail("Option --seed (-s) requires an unsigned integer value");
    a.seed = v;
    a.seed_given = true;
    return true;
}

bool parse_device(int argc, char **argv, int &i, Args &a)
{
    const std::string s(argv[i]);
    if (s != "--device")
        return false;
    if (a.device_given)
        fail("Option --device already given");
    if (i >= argc - 1)
        fail("Option --device requires an integer value >= 0");
    const int v = atoi(argv[++i]);
    if (v < 0)
        fail("Option --device requires an integer value >= 0");
    a.device = v;
    a.device_given = true;
    return true;
}

void usage()
{
    std::cout << "Questo è l'help" << std::endl
              << "usage: procell -h <histogram> -c <cell-types> -t <time-max> [-o <output>] [-p <phi>]\n"
                 "               [-r] [--seed <u64>] [--device <n>] [-d <1..23>]" << std::endl;
}

Args parse(int argc, char **argv)
{
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string s(argv[i]);
        bool legal = file_option(argc, argv, i, "--histogram", "-h", nullptr, "--histogram (-h)", a.h0_given, a.h0);
        legal = legal || file_option(argc, argv, i, "--cell-types", "-c", nullptr, "--cell-types (-c)",
                                     a.types_given, a.types);
        legal = legal || file_option(argc, argv, i, "--output-histogram", "-o", "--output",
                                     "--output-histogram (-o)", a.out_given, a.out);
        legal = legal || parse_t_max(argc, argv, i, a);
        legal = legal || parse_phi(argc, argv, i, a);
        legal = legal || parse_depth(argc, argv, i, a);
        legal = legal || parse_track_ratio(argc, argv, i, a);
        legal = legal || parse_seed(argc, argv, i, a);
        legal = legal || parse_device(argc, argv, i, a);
        if (s == "--help") {
            usage();
            exit(EXIT_SUCCESS);
        }
        if (!legal)
            fail("Invalid option " + s);
    }
    if (!a.h0_given || !a.types_given || !a.t_given) {
        std::cout << "The following missing arguments are required:" << std::endl;
        if (!a.h0_given)
            std::cout << "--histogram (-h)" << std::endl;
        if (!a.types_given)
            std::cout << "--cell-types (-c)" << std::endl;
        if (!a.t_given)
            std::cout << "--t-max (-t)" << std::endl;
        exit(EXIT_FAILURE);
    }
    return a;
}

}  // namespace

int main(int argc, char **argv)
{
    const Args a = parse(argc, argv);
    procell_row *rows = nullptr;
    procell_type *types = nullptr;
    uint64_t n_rows = 0;
    uint32_t n_types = 0;
    if (procell_read_histogram(a.h0.c_str(), &rows, &n_rows) != PROCELL_OK ||
        procell_read_types(a.types.c_str(), &types, &n_types) != PROCELL_OK)
        fail("ERROR: cannot read the input files, aborting.");

    procell_params p = {};
    p.rows = rows;
    p.n_rows = n_rows;
    p.types = types;
    p.n_types = n_types;
    p.t_max = a.t_max;
    p.phi = a.phi_given ? a.phi : 0.0;
    p.seed = a.seed;
    p.track_ratio = a.track_ratio ? 1 : 0;
    p.device = a.device;

    uint64_t cap = 0;
    int rc = procell_output_capacity(&p, &cap);
    if (rc == PROCELL_E_PROPORTIONS)
        fail("ERROR: proportion distribution of cell types does not sum to 1, aborting.");
    if (rc != PROCELL_OK)
        fail(std::string("ERROR: ") + procell_strerror(rc) + ", aborting.");
    std::vector<double> value(cap ? cap : 1);
    std::vector<int64_t> count(cap ? cap : 1);
    std::vector<int64_t> type_count((cap ? cap : 1) * (n_types ? n_types : 1));
    procell_result r = {};
    r.capacity = cap;
    r.value = value.data();
    r.count = count.data();
    r.type_count = type_count.data();
    rc = procell_simulate(&p, &r);
    if (rc != PROCELL_OK)
        fail(std::string("ERROR: ") + procell_strerror(rc) + ", aborting.");
    rc = procell_write_histogram(a.out_given ? a.out.c_str() : nullptr, &r, n_types, p.track_ratio);
    procell_free(rows);
    procell_free(types);
    if (rc != PROCELL_OK) {
        std::cerr << "ERROR: " << procell_strerror(rc) << std::endl;
        return EXIT_FAILURE;
    }
    return 0;
}
