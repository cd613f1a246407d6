/*
 * procell.h -- C ABI of the MI355X (gfx950) ProCell proliferation library,
 * libprocell.so.  Plain C types only; no HIP or torch types in signatures
 * (streams travel as void*).
 *
 * The reference (cuda-pro-cell) has no C ABI; its seam between the driver
 * (src/simulation/simulator.cu:10-56) and the simulation is a pair of C++
 * calls plus the text-format helpers.  Each entry point below names the
 * reference interface it replaces.
 *
 * Semantics (what one simulation computes) follow the reference's plain
 * method: every initial cell of a fluorescence histogram draws a type by
 * proportion; a proliferating cell (mean >= 0) whose next division completes
 * by t_max and whose halved fluorescence stays above phi divides into two
 * daughters with fresh Normal(mean, sd) timers truncated to > 0; a cell out
 * of time is tallied at its fluorescence; a cell in time whose halved
 * fluorescence would not exceed phi is dropped.  Random draws come from
 * Philox4x32-10 keyed on (seed, cell id, tree path), so results are a pure
 * function of the inputs and the seed and do not depend on the device, the
 * tuning knobs or the launch shape.
 */
#ifndef PROCELL_H
#define PROCELL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PROCELL_ABI_VERSION 2

/* ---- status codes (never exit(); every entry point returns one) ---- */
#define PROCELL_OK 0
#define PROCELL_E_INVALID (-1)        /* null pointer, bad size or bad knob */
#define PROCELL_E_PROPORTIONS (-2)    /* proportions do not sum to 1 within 1e-8 */
#define PROCELL_E_ZERO_MEAN (-3)      /* a type with mean == 0 (timer undefined) */
#define PROCELL_E_NONFINITE (-4)      /* NaN / inf in values, proportions, means, sds */
#define PROCELL_E_PHI (-5)            /* resolved phi is not > 0 */
#define PROCELL_E_DEPTH (-6)          /* a row would need more than 62 generations */
#define PROCELL_E_TOO_MANY_CELLS (-7) /* more than 2^32-1 cells, 2^24 rows or 256 types */
#define PROCELL_E_CAPACITY (-8)       /* result buffers too small */
#define PROCELL_E_HIP (-9)            /* a HIP runtime call failed */
#define PROCELL_E_NO_DEVICE (-10)     /* no usable gfx950 device */
#define PROCELL_E_IO (-11)            /* file could not be written */
#define PROCELL_E_OUT_OF_MEMORY (-12) /* host or device allocation failed */
#define PROCELL_E_SCRATCH (-13)       /* needs more than PROCELL_NOALLOC_SCRATCH_BYTES (procell_simulate_noalloc) */

/* One histogram row: fluorescence value and number of cells.
 * Replaces struct fluorescence (src/simulation/data_types.h:29-33). */
typedef struct {
    double value;
    uint64_t count;
} procell_row;

/* One cell type, in file order (its index is the reference's `name`).
 * Replaces struct cell_type (src/simulation/data_types.h:21-27); mean < 0
 * marks a quiescent type (README.md:84-90). */
typedef struct {
    double proportion;
    double mean;
    double sd;
} procell_type;

/* Launch knobs.  0 means "choose automatically".  Results never depend on
 * them; only speed does. */
typedef struct {
    uint32_t block_size;    /* threads per workgroup: 64..256, multiple of 64 */
    uint32_t blocks_per_cu; /* resident workgroups per CU */
    uint32_t segments;      /* segments of initial cells per resident workgroup (1..64); the
                               workgroups of later segments start as earlier ones finish */
    uint32_t idle_lanes;    /* idle lanes a wavefront tolerates before refilling them (0..63) */
    uint32_t flags;         /* PROCELL_FLAG_* */
} procell_tuning;

#define PROCELL_FLAG_TIMING 1u   /* record per-kernel device times into procell_stats */

typedef struct {
    const procell_row *rows;     /* histogram rows in file order; zero counts are skipped */
    uint64_t n_rows;
    const procell_type *types;   /* cell types in file order */
    uint32_t n_types;
    double t_max;                /* >= 0 (may be +inf) */
    double phi;                  /* > 0, or <= 0 for "smallest value with a nonzero count" */
    uint64_t seed;
    int32_t track_ratio;         /* nonzero: fill per-type columns */
    int32_t device;              /* HIP device ordinal */
    procell_tuning tuning;
    void *stream;                /* hipStream_t for the plan API; NULL = library-owned stream */
} procell_params;

typedef struct {
    uint64_t divisions;   /* cells that divided into two daughters */
    uint64_t recorded;    /* cells tallied into the output histogram */
    uint64_t dropped;     /* cells in time whose halved fluorescence would not exceed phi */
    uint64_t lost;        /* initial cells out of time with value < phi (not in the output bins) */
    uint64_t rejected;    /* normal draws rejected because the timer was <= 0 */
    uint64_t cells;       /* initial cells */
    uint64_t pool;        /* initial cells that divided at least once */
    double ms_tree;       /* device time of the last run's kernels: clearing, roots, trees (PROCELL_FLAG_TIMING) */
    double ms_run;        /* device time of the whole run incl. clearing (PROCELL_FLAG_TIMING; = ms_tree) */
    uint32_t workgroups;  /* launch shape chosen for the run */
    uint32_t block_size;
    uint32_t lds_bytes;
    uint32_t window_rows; /* rows of the per-workgroup LDS tally window */
} procell_stats;

typedef struct {
    uint64_t capacity;     /* rows available in value / count (/ type_count) */
    double *value;         /* [capacity] output fluorescence values, ascending */
    int64_t *count;        /* [capacity] cells per value */
    int64_t *type_count;   /* [capacity * n_types] per-type counts, file order; may be NULL */
    uint64_t n_out;        /* out: rows written (only values with count > 0) */
    int64_t *raw_tally;    /* optional [nonzero_rows][kdim][n_types] counts by (row, generation, type) */
    uint64_t raw_tally_len;/* elements available in raw_tally */
    procell_stats stats;   /* out */
} procell_result;

/* Run one simulation end to end (upload, kernels, download, merge).
 * Replaces create_cells_population (src/simulation/cells_population.h:12-18)
 * followed by proliferate (src/simulation/proliferation.h:12-19).  Unlike the
 * reference, results are reproducible for a given seed, counts are 64-bit and
 * errors are returned instead of ignored. */
int procell_simulate(const procell_params *p, procell_result *r);

/* procell_simulate without device allocation: the same inputs, outputs,
 * results (bit for bit) and errors, but every device buffer of the run (its
 * tables, segments, counters and tally) is carved from one static device
 * array of PROCELL_NOALLOC_SCRATCH_BYTES per GPU, which the library holds
 * until the process ends.  It calls no hipMalloc/hipFree, creates no stream
 * or event, and runs on p->stream (NULL: the default stream); host-side
 * set-up uses ordinary heap memory as procell_simulate does.  An input whose
 * device buffers exceed the array returns PROCELL_E_SCRATCH before any
 * kernel runs.  Calls on one GPU are serialised by a lock inside the library
 * (they would otherwise race on the shared array); the array keeps the last
 * call's contents and is cleared where a run needs zeros.  PROCELL_FLAG_TIMING
 * is ignored (no events).
 * No reference counterpart: this is for callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc). */
#define PROCELL_NOALLOC_SCRATCH_BYTES (16u * 1024u * 1024u)
int procell_simulate_noalloc(const procell_params *p, procell_result *r);

/* Number of distinct output values (the output bin set: every v / 2^k >= phi
 * of every nonzero row), an upper bound for r->capacity.
 * Replaces the bin-set part of load_fluorescences (src/io/parser.h:15-21). */
int procell_output_capacity(const procell_params *p, uint64_t *capacity);

/* Shape of raw_tally: number of nonzero rows and generations + 1.
 * No reference counterpart (the reference tallies by value only). */
int procell_tally_shape(const procell_params *p, uint64_t *nonzero_rows, uint32_t *kdim);

/* Resolved phi (the README rule for an absent -p, src/io/parser.cu:80-96). */
int procell_resolve_phi(const procell_params *p, double *phi);

const char *procell_strerror(int code);
int procell_abi_version(void);

/* ---- plan API: inputs uploaded once, many runs on one stream ----------
 * No reference counterpart; it splits procell_simulate so that callers
 * (benchmarks, batch drivers) can keep inputs resident in device memory. */
typedef struct procell_plan procell_plan;
int procell_plan_create(const procell_params *p, procell_plan **plan);
int procell_plan_run(procell_plan *plan, uint64_t seed);            /* enqueue only */
int procell_plan_result(procell_plan *plan, procell_result *r);     /* wait, download, merge */
int procell_plan_destroy(procell_plan *plan);
void *procell_plan_stream(procell_plan *plan);                       /* hipStream_t */
/* With PROCELL_FLAG_TIMING: device time (ms) of the kernels of each of the
 * last min(runs, 256, max_runs) runs since the previous call (the clearing
 * kernel, the root kernel and the tree kernel, back to back), oldest first,
 * from HIP events recorded on the plan's stream around the three kernels.
 * Waits for the stream; resets the run count. */
int procell_plan_kernel_times(procell_plan *plan, double *ms, uint32_t max_runs, uint32_t *n_runs);

/* ---- batched plain method: many parameter sets per launch --------------
 * No reference counterpart (the reference simulates one parameter set per
 * process).  The sets' initial cells, rows, types and tallies are
 * concatenated and run by one launch of each kernel; every tree workgroup
 * works on one set.  Set i's result equals procell_simulate of sets[i] with
 * seed seeds[i], bit for bit.  Runs replay one hipGraph of the three
 * kernels, captured on the first run.  All sets must name the same device;
 * the launch knobs, flags and stream of sets[0] apply to the batch.  On a
 * batch plan, procell_plan_run gives every set the same seed and
 * procell_plan_result returns set 0. */
int procell_batch_create(const procell_params *sets, uint32_t n_sets, procell_plan **plan);
int procell_batch_run(procell_plan *plan, const uint64_t *seeds);              /* seeds[n_sets]; enqueue only */
int procell_batch_result(procell_plan *plan, uint32_t set, procell_result *r); /* wait, download, merge */
/* Autotuner of a batch plan: procell_plan_autotune (below). */
int procell_batch_autotune(procell_plan *plan, const uint64_t *seeds, uint32_t reps, procell_tuning *chosen);
/* Autotuner of any plan (one set or a batch): times `reps` runs
 * (seeds[n_sets], one per set) of every candidate layout -- resident
 * workgroups per CU, which sets the width of the LDS tally window, times
 * segments per resident workgroup, i.e. the granularity of the work handed
 * to the workgroup dispatcher -- and, for each layout, of every idle-lane
 * threshold (how many lanes of a wavefront may idle before it refills
 * them), on the plan's own inputs; the plan keeps the fastest for later
 * runs and *chosen returns it.  One initial cell per lane at a time (the
 * lane walks its tree depth first).  Results do not depend on the knobs.
 * No reference counterpart (the reference's one knob, -d, only batched
 * levels: src/io/cmdargs.cpp:295-351). */
int procell_plan_autotune(procell_plan *plan, const uint64_t *seeds, uint32_t reps, procell_tuning *chosen);

/* ---- relaxed (differentiable) estimator --------------------------------
 * No reference counterpart (the reference only samples).  For calibrating
 * (mean, sd, proportion) of each type to an observed histogram.  The plan
 * fixes the histogram, t_max, phi, seed, the type count and which types are
 * quiescent (mean < 0 in *p); beta > 0 is the temperature of
 * sigmoid((t_max - s) / beta), which replaces the plain method's "division
 * completes by t_max" test.  Every initial cell takes every type j with
 * weight p_j; timers are mean_j + sd_j z with z the first normal candidates
 * of the plain method's Philox stream, untruncated; a node of weight W
 * records W (1 - d) at (row, generation, type) and its daughters carry W d;
 * fluorescence and the generation cap are the plain method's.  Subtrees
 * whose weight falls below 1e-15 are not expanded.
 *   theta : device [3][n_types] = mean, sd, proportion (file order); the
 *           entries of quiescent types' mean and sd are ignored.  Keep
 *           mean > 0 and sd < 0.85 mean for proliferating types: beyond
 *           that, untruncated timers let the expected tree grow until the
 *           generation cap (2^cap nodes per walk).
 *   h     : device [n_rows][kdim][n_types] relaxed expected tally (rows =
 *           nonzero rows in file order, as raw_tally).
 *   g     : device, shape of h: the cotangent dL/dh.
 *   grad  : device [3][n_types] = dL/dmean, dL/dsd, dL/dproportion.
 * The backward pass recomputes every subtree from the stream; nothing is
 * kept between calls.  Calls are enqueued on `stream` (a hipStream_t; NULL is
 * the default stream), in order with the caller's other work on it, and
 * return without waiting.  Calls on one plan must not overlap (they share
 * its per-lane stacks). */
typedef struct procell_relaxed procell_relaxed;
int procell_relaxed_create(const procell_params *p, double beta, procell_relaxed **plan);
int procell_relaxed_shape(const procell_relaxed *plan, uint64_t *n_rows, uint32_t *kdim, uint32_t *n_types);
int procell_relaxed_forward(procell_relaxed *plan, const double *theta, double *h, void *stream);
int procell_relaxed_backward(procell_relaxed *plan, const double *theta, const double *g, double *grad,
                             void *stream);
/* Cell stride of the plan's later passes: only the initial cells c (ids
 * within the set) with (c + phase) a multiple of `stride` (a power of two,
 * at most 65536) are walked, so the passes of phases 0 .. stride-1 sum to
 * the full pass; 1 (the default) walks every cell. */
int procell_relaxed_set_stride(procell_relaxed *plan, uint32_t stride, uint32_t phase);
/* h and its Jacobian in one pass.  The entries of type j depend on type j's
 * (mean, sd, proportion) only, so the Jacobian is three numbers per entry:
 *   jac : device [2][n_rows][kdim][n_types] = dh/dmean_j, dh/dsd_j of each
 *         entry's own type j (0 for quiescent types); dh/dproportion_j is
 *         h / proportion_j.
 * Enqueued like forward.  No counterpart in the reference: it serves
 * second-order (Gauss-Newton) calibration. */
int procell_relaxed_jacobian(procell_relaxed *plan, const double *theta, double *h, double *jac, void *stream);
/* A plan over n_sets parameter sets walked by the same launches (sets may
 * differ in histogram, types, t_max, phi and seed; one device).  theta, h,
 * jac, g and grad are the sets' arrays concatenated in set order: set i's
 * theta starts at 3 x (types of sets before i), its h at h_base (below).
 * procell_relaxed_shape describes set 0. */
int procell_relaxed_create_batch(const procell_params *sets, uint32_t n_sets, double beta, procell_relaxed **plan);
int procell_relaxed_set_shape(const procell_relaxed *plan, uint32_t set, uint64_t *h_base, uint64_t *n_rows,
                              uint32_t *kdim, uint32_t *n_types);
/* Expansions of the plan's last call (roots drawn plus nodes whose daughters
 * were drawn); waits for the device.  Returns PROCELL_E_DEPTH (with *n set)
 * when a walk needed more than 2^22 expansions and was cut short: the call's
 * h or gradient is then incomplete (parameters outside the estimator's
 * domain, see theta above). */
int procell_relaxed_expansions(procell_relaxed *plan, uint64_t *n);
/* Walks cut short (over the 2^22-expansion budget) in the plan's last call,
 * per set: enqueues on `stream` a copy of the plan's counters into the
 * caller's device array cuts[n_sets] and returns without waiting, so a
 * caller can reject a set's cut pass without a host round trip.  No
 * reference counterpart. */
int procell_relaxed_set_cuts(procell_relaxed *plan, uint64_t *cuts, void *stream);
int procell_relaxed_destroy(procell_relaxed *plan);

/* ---- batched calibration ----------------------------------------------
 * No reference counterpart.  Fits each parameter set's types to a target
 * histogram with the relaxed estimator above: per set, the loss is the
 * squared L2 distance between the normalised relaxed histogram (h summed
 * over types, each (row, generation) entry merged into its output value
 * v / 2^k) and the normalised target; one step runs, for all sets at once,
 * the relaxed Jacobian pass (h and dh/dtheta in one walk), the loss and its
 * gradient (the Jacobian contracted with dL/dh; with a cell stride k > 1:
 * the forward pass, the loss and its cotangent, the adjoint pass) and
 * an Adam update of (log mean, log sd) of every proliferating type and the
 * proportion logits of every type (proportions = softmax of the logits), all
 * on the device, replayed from one hipGraph captured at the first step.
 *   sets    : [n_sets] histograms, t_max, phi, seeds of the relaxed streams
 *             and the starting parameters (proliferating types need mean,
 *             sd and proportion > 0); the device, and the stream when not
 *             NULL, of sets[0] apply to the whole calibration.
 *   targets : [n_sets] target histograms in the plain method's output form
 *             (value[n_out], count[n_out]); every value must be one of the
 *             set's output values (e.g. a procell_simulate result).
 *   adam    : lr, beta1, beta2, eps (the update of Kingma & Ba, 2015);
 *             each update keeps sd <= 0.8 mean (see theta above).
 *   cell_stride : 1 walks every initial cell each step; a power of two
 *             k > 1 makes the forward pass walk the cells of phase
 *             (step mod k) and the adjoint those of phase (step + 1 mod k)
 *             only (see procell_relaxed_set_stride): k times cheaper steps,
 *             and a residual and Jacobian from independent cells (double
 *             sampling), so the gradient of the squared loss is unbiased. */
typedef struct {
    double lr;
    double beta1;
    double beta2;
    double eps;
} procell_adam;
typedef struct procell_calib procell_calib;
int procell_calib_create(const procell_params *sets, uint32_t n_sets, double beta, const procell_result *targets,
                         const procell_adam *adam, uint32_t cell_stride, procell_calib **calib);
/* Enqueues n_steps steps (graph replays) on the calibration's stream. */
int procell_calib_step(procell_calib *calib, uint32_t n_steps);
/* Waits; current parameters: per set [3][n_types] (mean, sd, proportion;
 * -1 for a quiescent type's mean and sd), sets concatenated. */
int procell_calib_theta(procell_calib *calib, double *theta, uint64_t capacity);
/* Waits; loss of each set at the last step's parameters, the steps taken
 * and the expansions of the last step (its walks: the Jacobian pass, or
 * with a cell stride the forward and adjoint passes); steps and
 * expansions may be NULL.  Returns PROCELL_E_DEPTH when a walk of the last
 * step exceeded 2^22 expansions and was cut short. */
int procell_calib_loss(procell_calib *calib, double *loss, uint32_t capacity, uint64_t *steps,
                       uint64_t *expansions);
void *procell_calib_stream(procell_calib *calib);                   /* hipStream_t */
int procell_calib_destroy(procell_calib *calib);

/* ---- text formats ------------------------------------------------------
 * Replace load_fluorescences' file reading (src/io/parser.cu:68-124),
 * load_cell_types (src/io/parser.cu:156-185) and save_fluorescences
 * (src/io/parser.cu:187-217).  Arrays are malloc'ed; release with
 * procell_free.  A file that cannot be opened reads as empty, as in the
 * reference. */
int procell_read_histogram(const char *path, procell_row **rows, uint64_t *n_rows);
int procell_read_types(const char *path, procell_type **types, uint32_t *n_types);
/* path == NULL writes to stdout.  Format: "%.10g\t%lld[\t%lld...]\n" per row. */
int procell_write_histogram(const char *path, const procell_result *r, uint32_t n_types,
                            int32_t track_ratio);
void procell_free(void *ptr);

/* ---- diagnostics: the random stream evaluated on the device ------------
 * For each i: words[4i..4i+3] = Philox block of (seed, cell[i], node[i],
 * attempt[i]); z[2i], z[2i+1] = the Box-Muller pair of that block.  Host
 * pointers; runs one small kernel.  No reference counterpart. */
int procell_debug_draws(int32_t device, uint64_t seed, const uint64_t *cell, const uint64_t *node,
                        const uint32_t *attempt, uint64_t n, uint32_t *words, double *z);

#ifdef __cplusplus
}
#endif

#endif /* PROCELL_H */
