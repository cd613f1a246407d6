// cmdargs_harness.cpp -- test infrastructure only: a main() around the
// reference's own argument parser (src/io/cmdargs.cpp:11-375 of
// cuda-pro-cell, compiled from where it lies by oracle/build.py into
// oracle/_ref/cmdargs_ref).
//
// The parser prints its messages and calls exit() itself on --help and on
// every error (cmdargs.cpp:33-38, 40-44, 47-74).  When it returns, this main
// prints one line with the parsed fields so that a test can compare accepted
// argument lines too, then exits 0.  Fields the reference leaves unset when
// an optional option is absent are printed as "-".
#include <stdio.h>

#include "io/cmdargs.h"

int main(int argc, char **argv)
{
    procell::io::CmdArgs a(argc, argv);
    printf("parsed h0=%s cell_types=%s output=%s t_max=%.17g phi_min=%.17g tree_depth=%u track_ratio=%d\n",
           a.h0.c_str(), a.cell_types.c_str(), a.output_histogram_given ? a.output_histogram.c_str() : "-",
           (double)a.t_max, (double)a.phi_min, (unsigned)a.tree_depth, a.track_ratio ? 1 : 0);
    return 0;
}
