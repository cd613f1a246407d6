/*
 * procell_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement (plain C11 + OpenMP) of ProCell's plain proliferation
 * method, used as the correctness checker for the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load or call it, and only as the checker / the timed CPU baseline; the
 * product library never links it.
 *
 * Reference behaviour restated (cuda-pro-cell):
 *   src/io/parser.cu:68-154   histogram rows, prefix bounds, auto-phi, output bin set
 *   src/io/parser.cu:156-185  cell types, proportion check, descending sort
 *   src/io/parser.cu:187-217  output text format
 *   src/simulation/cell.cu:25-143          initial cell: type, truncated timer, phase
 *   src/simulation/proliferation.cu:309-410 divide / record / drop rule
 * The random stream is this project's own (Philox4x32-10 keyed on seed, cell
 * id and tree path, FP64 Box-Muller), not cuRAND's: see procell_oracle.c.
 */
#ifndef PROCELL_ORACLE_H
#define PROCELL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const double *values;       /* histogram rows in file order (zero counts allowed) */
    const uint64_t *counts;
    uint64_t n_rows;
    const double *proportion;   /* cell types in file order */
    const double *mean;         /* mean < 0: quiescent */
    const double *sd;
    uint32_t n_types;
    double t_max;
    double phi;                 /* <= 0: smallest value with a nonzero count */
    uint64_t seed;
} oracle_problem;

enum {
    ORACLE_STAT_DIVISIONS = 0,
    ORACLE_STAT_RECORDED = 1,
    ORACLE_STAT_DROPPED = 2,
    ORACLE_STAT_LOST = 3,
    ORACLE_STAT_REJECTED = 4,     /* normal draws rejected because the timer was <= 0 */
    ORACLE_STAT_CELLS = 5,
    ORACLE_N_STATS = 6
};

/* Primitive pieces, exported so tests can pin each one on its own. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void oracle_block(uint64_t seed, uint64_t cell, uint64_t node, uint32_t attempt, uint32_t out[4]);
double oracle_u01_open_closed(uint32_t hi, uint32_t lo);
double oracle_u01_closed_open(uint32_t hi, uint32_t lo);
double oracle_log(double x);
double oracle_sqrt(double x);
void oracle_sincos_turns(uint32_t hi, uint32_t lo, double *s, double *c);
void oracle_normal_pair(const uint32_t w[4], double *z0, double *z1);

/* Resolved phi (auto rule applied) and generation cap per nonzero row. */
double oracle_resolve_phi(const oracle_problem *p);
int oracle_max_generation(const oracle_problem *p);

/*
 * Full simulation.  tally: int64 [n_nonzero_rows][kdim][n_types] (types in file
 * order), kdim = oracle_max_generation(p) + 1; drops: int64 [n_nonzero_rows][kdim]
 * (may be NULL); stats: uint64 [ORACLE_N_STATS].  n_threads <= 0: OpenMP default.
 * Returns 0, or a negative code for invalid input.
 */
int oracle_simulate(const oracle_problem *p, int64_t *tally, int64_t *drops,
                    uint64_t *stats, int n_threads);

/*
 * Relaxed estimator (this project's differentiable relaxation, see the
 * source): h = expected tally [n_nonzero_rows][kdim][n_types] (file type
 * order, may be NULL); with g = dL/dh of the same shape (may be NULL),
 * grad = dL/d(mean, sd, proportion) as [3][n_types] (file order).
 * beta > 0 is the sigmoid width.  Returns 0 or a negative code.
 */
int oracle_relaxed(const oracle_problem *p, double beta, const double *g, double *h, double *grad,
                   int n_threads);

/* File-level run: reads the two input files, simulates, writes the output
 * text (value<TAB>count[<TAB>per-type...] per line).  Returns 0 or negative. */
int oracle_run_files(const char *histogram_path, const char *types_path,
                     double t_max, double phi, int track_ratio, uint64_t seed,
                     const char *output_path, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
