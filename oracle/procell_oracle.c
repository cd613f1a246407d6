/*
 * procell_oracle.c -- TEST INFRASTRUCTURE ONLY (see procell_oracle.h).
 *
 * Plain C restatement of ProCell's plain proliferation method with this
 * project's random stream.  It is the checker for the HIP implementation and
 * the timed CPU baseline; it is never linked into the product.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -mfma -fopenmp (oracle/build.py).
 * -ffp-contract=off keeps every floating-point operation exactly as written,
 * so that the GPU implementation, which is written to the same operation
 * sequence and also built without contraction, produces identical bits.
 *
 * Random stream (this project's own design; the reference uses cuRAND XORWOW
 * re-initialised per draw, src/utils/util.cu:143-169, which is not reproduced):
 *   Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 *   easy as 1, 2, 3", SC'11), key = (seed lo, seed hi),
 *   counter = (node lo, node hi, cell lo, cell hi[15:0] | attempt << 16).
 *   node is the heap index of a cell in its division tree (root = 1,
 *   daughters of n are 2n and 2n+1); node 0 holds the root's setup draws.
 *     block(cell, 0, 0)     -> the root's draws, one word each: type uniform
 *                              in [0,1) (word 0), phase U in (0,1] (word 1),
 *                              and the Box-Muller pair of its first timer's
 *                              two candidates (words 2, 3)
 *     block(cell, n, 0)     -> Box-Muller pair: timers of daughters 2n and 2n+1
 *     block(cell, n, a>=1)  -> truncated-normal redraw stream of node n (the
 *                              root's first timer continues in node 1's
 *                              stream when both of its candidates are <= 0)
 */
#define _POSIX_C_SOURCE 200809L
#include "procell_oracle.h"

#include <inttypes.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Error codes; numerically identical to the product's include/procell.h. */
#define OR_OK 0
#define OR_E_INVALID (-1)
#define OR_E_PROPORTIONS (-2)
#define OR_E_ZERO_MEAN (-3)
#define OR_E_NONFINITE (-4)
#define OR_E_PHI (-5)
#define OR_E_DEPTH (-6)
#define OR_E_TOO_MANY_CELLS (-7)
#define OR_E_IO (-11)

#define OR_MAX_GENERATION 62
#define OR_MAX_CELLS (((uint64_t)1) << 48)

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Random123 constants, SC'11 paper, Table 2).           */
/* ------------------------------------------------------------------ */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)x2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        x0 = n0; x1 = n1; x2 = n2; x3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

void oracle_block(uint64_t seed, uint64_t cell, uint64_t node, uint32_t attempt, uint32_t out[4])
{
    uint32_t ctr[4];
    uint32_t key[2];
    ctr[0] = (uint32_t)node;
    ctr[1] = (uint32_t)(node >> 32);
    ctr[2] = (uint32_t)cell;
    ctr[3] = ((uint32_t)(cell >> 32) & 0xFFFFu) | (attempt << 16);
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    oracle_philox4x32_10(ctr, key, out);
}

/* 53-bit uniforms: m = hi * 2^21 + (lo >> 11) is formed exactly. */
double oracle_u01_closed_open(uint32_t hi, uint32_t lo)
{
    double m = fma((double)hi, 0x1p21, (double)(lo >> 11));
    return m * 0x1p-53;
}

double oracle_u01_open_closed(uint32_t hi, uint32_t lo)
{
    double m = fma((double)hi, 0x1p21, (double)(lo >> 11));
    return (m + 1.0) * 0x1p-53;
}

#define PROCELL_TABLE_QUAL static
#include "procell_math_tables.inc"

/* ------------------------------------------------------------------ */
/* FP64 log, table driven (the method of Tang, "Table-driven           */
/* implementation of the logarithm function in IEEE floating-point      */
/* arithmetic", ACM TOMS 16(4), 1990, in the fma form used by modern     */
/* libms).  x = 2^e m, m in [1, 2); j = top 7 fraction bits; for j >= 64 */
/* m is halved and e incremented, so the two table intervals next to 1   */
/* have c = 1 (exact results near x = 1).  With {invc, logc} from the    */
/* table (procell_math_tables.inc): r = fma(m, invc, -1), |r| < 2^-7,     */
/* log x = (e ln2_hi + logc) + (r + (r^2 P(r) + e ln2_lo)) with           */
/* P(r) = -1/2 + r/3 - r^2/4 + r^3/5 - r^4/6 + r^5/7 (Taylor series of    */
/* (log(1+r) - r)/r^2).  Valid for positive normal x.                    */
/* ------------------------------------------------------------------ */
static const double LN2_HI = 0x1.62e42fee00000p-1;
static const double LN2_LO = 0x1.a39ef35793c76p-33;
static const double LP1 = -0.5;
static const double LP2 = 0x1.5555555555555p-2;   /*  1/3 */
static const double LP3 = -0.25;
static const double LP4 = 0x1.999999999999ap-3;   /*  1/5 */
static const double LP5 = -0x1.5555555555555p-3;  /* -1/6 */
static const double LP6 = 0x1.2492492492492p-3;   /*  1/7 */

double oracle_log(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    int e = (int)(bits >> 52) - 1023;
    uint32_t j = (uint32_t)(bits >> 45) & 127u;
    uint64_t mbits = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &mbits, sizeof m);
    if (j >= 64u) {
        m = m * 0.5;
        e = e + 1;
    }
    double invc = procell_log_table[j][0];
    double logc = procell_log_table[j][1];
    double r = fma(m, invc, -1.0);
    double r2 = r * r;
    double p = LP6;
    p = fma(p, r, LP5);
    p = fma(p, r, LP4);
    p = fma(p, r, LP3);
    p = fma(p, r, LP2);
    p = fma(p, r, LP1);
    double de = (double)e;
    double hi = fma(de, LN2_HI, logc);
    double lo = de * LN2_LO;
    double t = fma(r2, p, lo);
    return hi + (r + t);
}

/* ------------------------------------------------------------------ */
/* FP64 sqrt, division-free: reciprocal square root seed from the        */
/* integer-shift estimate with Lomont's 64-bit constant 0x5FE6EB50C7B537A9*/
/* (C. Lomont, "Fast inverse square root", 2003; max rel. error 3.42%),  */
/* three Newton-Raphson steps y <- y + y (1/2 - a/2 y^2), then one       */
/* residual correction s <- s + (y/2)(a - s^2) on s = a y.               */
/* Valid for finite a >= 0 (a == 0 returns 0).                           */
/* ------------------------------------------------------------------ */
double oracle_sqrt(double a)
{
    if (!(a > 0.0))
        return 0.0;
    uint64_t ab;
    memcpy(&ab, &a, sizeof ab);
    uint64_t yb = 0x5FE6EB50C7B537A9ull - (ab >> 1);
    double y;
    memcpy(&y, &yb, sizeof y);
    double ha = 0.5 * a;
    double t, e;
    t = y * y; e = fma(-ha, t, 0.5); y = fma(y, e, y);
    t = y * y; e = fma(-ha, t, 0.5); y = fma(y, e, y);
    t = y * y; e = fma(-ha, t, 0.5); y = fma(y, e, y);
    double s = a * y;
    double res = fma(-s, s, a);
    return fma(0.5 * y, res, s);
}

/* ------------------------------------------------------------------ */
/* sin and cos of 2*pi*u, u = u01_closed_open(hi, lo), table driven.    */
/* j = nearest multiple of 1/256 turn (from the top bits of hi, 0..256), */
/* d = u - j/256 exactly (|d| <= 1/512), t = 2 pi d with 2 pi split in    */
/* hi and lo parts; sin t = t + t^3 (S1 + t^2 (S2 + t^2 S3)) and          */
/* cos t - 1 = t^2 (C1 + t^2 (C2 + t^2 C3)) (Taylor series); then the     */
/* addition theorems with {sin, cos}(2 pi j/256) from the table           */
/* (procell_math_tables.inc).                                             */
/* ------------------------------------------------------------------ */
static const double TWO_PI_HI = 0x1.921fb54442d18p+2;
static const double TWO_PI_LO = 0x1.1a62633145c07p-52;
static const double SC_S1 = -0x1.5555555555555p-3;  /* -1/6    */
static const double SC_S2 = 0x1.1111111111111p-7;   /*  1/120  */
static const double SC_S3 = -0x1.a01a01a01a01ap-13; /* -1/5040 */
static const double SC_C1 = -0.5;
static const double SC_C2 = 0x1.5555555555555p-5;   /*  1/24   */
static const double SC_C3 = -0x1.6c16c16c16c17p-10; /* -1/720  */

void oracle_sincos_turns(uint32_t hi, uint32_t lo, double *s_out, double *c_out)
{
    double u = oracle_u01_closed_open(hi, lo);
    uint32_t j = ((hi >> 23) + 1u) >> 1;          /* 0..256 */
    double d = u - (double)j * 0x1p-8;            /* exact */
    double t = fma(d, TWO_PI_HI, d * TWO_PI_LO);
    double t2 = t * t;
    double ps = fma(t2, SC_S3, SC_S2);
    ps = fma(t2, ps, SC_S1);
    double st = fma(t * t2, ps, t);
    double cm = t2 * fma(t2, fma(t2, SC_C3, SC_C2), SC_C1);
    double sa = procell_sincos_table[j & 255u][0];
    double ca = procell_sincos_table[j & 255u][1];
    *s_out = sa + fma(sa, cm, ca * st);
    *c_out = ca + fma(ca, cm, -(sa * st));
}

/* Box-Muller (Box & Muller 1958): z0 = rho cos(2 pi u2), z1 = rho sin(2 pi u2),
 * rho = sqrt(-2 log u1), u1 in (0,1] from words 0-1, u2 in [0,1) from 2-3. */
void oracle_normal_pair(const uint32_t w[4], double *z0, double *z1)
{
    double u1 = oracle_u01_open_closed(w[0], w[1]);
    double a = -2.0 * oracle_log(u1);
    double rho = oracle_sqrt(a);
    double sn, cs;
    oracle_sincos_turns(w[2], w[3], &sn, &cs);
    *z0 = rho * cs;
    *z1 = rho * sn;
}

/* ------------------------------------------------------------------ */
/* Problem set-up (parser.cu:46-66, 68-154, 156-185 restated).          */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t n_nz;          /* nonzero rows, file order */
    double *v;
    uint64_t *n;
    uint64_t *start;        /* exclusive prefix sum of n (cell id of first cell) */
    int *kcap;              /* generations a row's cells may reach */
    uint64_t n_cells;
    uint32_t T;
    uint32_t *order;        /* sorted position -> file index */
    double *cum;            /* cumulative proportion in sorted order */
    double *mean;           /* sorted order */
    double *sd;
    double phi;
    double t_max;
    uint64_t seed;
    int kmax;
} oracle_setup;

static void setup_free(oracle_setup *s)
{
    free(s->v); free(s->n); free(s->start); free(s->kcap);
    free(s->order); free(s->cum); free(s->mean); free(s->sd);
    memset(s, 0, sizeof *s);
}

double oracle_resolve_phi(const oracle_problem *p)
{
    double phi = p->phi;
    if (phi > 0.0)
        return phi;
    /* parser.cu:80-96: smallest value among rows with a nonzero count. */
    double best = 0.0;
    int have = 0;
    for (uint64_t i = 0; i < p->n_rows; ++i) {
        if (p->counts[i] > 0 && (!have || p->values[i] < best)) {
            best = p->values[i];
            have = 1;
        }
    }
    return best;
}

static int row_kcap(double v, double phi)
{
    /* proliferation.cu:323: a cell of fluorescence f divides only if f/2 > phi;
     * fluorescence halves exactly each generation. */
    int k = 0;
    double f = v;
    while (f / 2 > phi) {
        f = f / 2;
        ++k;
        if (k > OR_MAX_GENERATION)
            return k;
    }
    return k;
}

static int setup_build(const oracle_problem *p, oracle_setup *s)
{
    memset(s, 0, sizeof *s);
    if (!p || (p->n_rows && (!p->values || !p->counts)))
        return OR_E_INVALID;
    /* t_max = +inf is allowed (every division is in time); NaN and negative
     * values are invalid arguments, like a negative -t (cmdargs.cpp:202). */
    if (isnan(p->t_max) || p->t_max < 0.0)
        return OR_E_INVALID;
    s->t_max = p->t_max;
    s->seed = p->seed;
    uint64_t n_nz = 0, total = 0;
    for (uint64_t i = 0; i < p->n_rows; ++i) {
        if (!isfinite(p->values[i]))
            return OR_E_NONFINITE;
        if (p->counts[i] > 0) {
            ++n_nz;
            if (total + p->counts[i] < total)
                return OR_E_TOO_MANY_CELLS;
            total += p->counts[i];
        }
    }
    if (total >= OR_MAX_CELLS)
        return OR_E_TOO_MANY_CELLS;
    if (p->n_types && (!p->proportion || !p->mean || !p->sd))
        return OR_E_INVALID;
    /* no types: the proportions sum to 0, not 1 (parser.cu:46-66) */
    if (p->n_types == 0)
        return total ? OR_E_PROPORTIONS : OR_OK;
    /* parser.cu:46-66: left fold of proportions in file order. */
    double sum = 0.0;
    for (uint32_t j = 0; j < p->n_types; ++j) {
        if (!isfinite(p->proportion[j]) || !isfinite(p->mean[j]) || !isfinite(p->sd[j]))
            return OR_E_NONFINITE;
        sum = sum + p->proportion[j];
    }
    if (fabs(1.0 - sum) > 1.0 / pow(10.0, 8.0))
        return OR_E_PROPORTIONS;
    for (uint32_t j = 0; j < p->n_types; ++j)
        if (p->mean[j] == 0.0)
            return OR_E_ZERO_MEAN;
    double phi = oracle_resolve_phi(p);
    if (total > 0 && !(phi > 0.0))
        return OR_E_PHI;
    s->phi = phi;
    s->n_nz = n_nz;
    s->n_cells = total;
    s->T = p->n_types;
    s->v = (double *)malloc((n_nz + 1) * sizeof(double));
    s->n = (uint64_t *)malloc((n_nz + 1) * sizeof(uint64_t));
    s->start = (uint64_t *)malloc((n_nz + 1) * sizeof(uint64_t));
    s->kcap = (int *)malloc((n_nz + 1) * sizeof(int));
    s->order = (uint32_t *)malloc(s->T * sizeof(uint32_t));
    s->cum = (double *)malloc(s->T * sizeof(double));
    s->mean = (double *)malloc(s->T * sizeof(double));
    s->sd = (double *)malloc(s->T * sizeof(double));
    uint64_t r = 0, acc = 0;
    int kmax = 0;
    for (uint64_t i = 0; i < p->n_rows; ++i) {
        if (p->counts[i] == 0)
            continue;
        s->v[r] = p->values[i];
        s->n[r] = p->counts[i];
        s->start[r] = acc;
        acc += p->counts[i];
        int kc = row_kcap(p->values[i], phi);
        if (kc > OR_MAX_GENERATION) {
            setup_free(s);
            return OR_E_DEPTH;
        }
        s->kcap[r] = kc;
        if (kc > kmax)
            kmax = kc;
        ++r;
    }
    s->start[n_nz] = acc;
    s->kmax = kmax;
    /* parser.cu:184: types sorted by descending proportion; equal
     * proportions keep file order (stable insertion sort). */
    for (uint32_t j = 0; j < s->T; ++j)
        s->order[j] = j;
    for (uint32_t j = 1; j < s->T; ++j) {
        uint32_t x = s->order[j];
        uint32_t i = j;
        while (i > 0 && p->proportion[s->order[i - 1]] < p->proportion[x]) {
            s->order[i] = s->order[i - 1];
            --i;
        }
        s->order[i] = x;
    }
    /* cell.cu:86-101: running sum of proportions in sorted order. */
    double c = 0.0;
    for (uint32_t j = 0; j < s->T; ++j) {
        uint32_t fj = s->order[j];
        c = c + p->proportion[fj];
        s->cum[j] = c;
        s->mean[j] = p->mean[fj];
        s->sd[j] = p->sd[fj];
    }
    return OR_OK;
}

int oracle_max_generation(const oracle_problem *p)
{
    oracle_setup s;
    int rc = setup_build(p, &s);
    if (rc != OR_OK)
        return rc;
    int k = s.kmax;
    setup_free(&s);
    return k;
}

/* ------------------------------------------------------------------ */
/* Per-cell simulation (cell.cu:25-143, proliferation.cu:317-381).      */
/* ------------------------------------------------------------------ */
typedef struct {
    const oracle_setup *s;
    uint64_t cell;
    uint64_t row;
    uint32_t j;            /* sorted type index */
    double mean, sd;
    int kcap;
    int64_t *tally;        /* this thread's [n_nz][kdim][T] (file type order) */
    int64_t *drops;        /* this thread's [n_nz][kdim] or NULL */
    uint64_t *st;          /* this thread's stats */
} cell_walk;

static void record(cell_walk *w, int k)
{
    const oracle_setup *s = w->s;
    uint64_t kdim = (uint64_t)s->kmax + 1;
    w->tally[(w->row * kdim + (uint64_t)k) * s->T + s->order[w->j]] += 1;
    w->st[ORACLE_STAT_RECORDED] += 1;
}

static void drop(cell_walk *w, int k)
{
    uint64_t kdim = (uint64_t)w->s->kmax + 1;
    if (w->drops)
        w->drops[w->row * kdim + (uint64_t)k] += 1;
    w->st[ORACLE_STAT_DROPPED] += 1;
}

/* cell.cu:106-122: redraw while the timer is <= 0, here from the node's own
 * stream, attempts 1, 2, ...; each block supplies two candidates. */
static double truncated_timer(cell_walk *w, uint64_t node)
{
    uint32_t blk[4];
    double z0, z1;
    for (uint32_t a = 1; a < 65536u; ++a) {
        oracle_block(w->s->seed, w->cell, node, a, blk);
        oracle_normal_pair(blk, &z0, &z1);
        double t = fma(z0, w->sd, w->mean);
        if (t > 0.0)
            return t;
        w->st[ORACLE_STAT_REJECTED] += 1;
        t = fma(z1, w->sd, w->mean);
        if (t > 0.0)
            return t;
        w->st[ORACLE_STAT_REJECTED] += 1;
    }
    return w->mean;
}

static void divide(cell_walk *w, uint64_t node, int k, double s_time);

/* Daughter at generation k whose division would complete at s_time. */
static void child(cell_walk *w, uint64_t node, int k, double s_time)
{
    if (s_time > w->s->t_max)
        record(w, k);                 /* out of time: tallied */
    else if (k == w->kcap)
        drop(w, k);                   /* f/2 <= phi: dropped */
    else
        divide(w, node, k, s_time);
}

static void divide(cell_walk *w, uint64_t node, int k, double s_time)
{
    uint32_t blk[4];
    double z0, z1;
    w->st[ORACLE_STAT_DIVISIONS] += 1;
    oracle_block(w->s->seed, w->cell, node, 0, blk);
    oracle_normal_pair(blk, &z0, &z1);
    double tl = fma(z0, w->sd, w->mean);
    if (!(tl > 0.0)) {
        w->st[ORACLE_STAT_REJECTED] += 1;
        tl = truncated_timer(w, 2 * node);
    }
    double tr = fma(z1, w->sd, w->mean);
    if (!(tr > 0.0)) {
        w->st[ORACLE_STAT_REJECTED] += 1;
        tr = truncated_timer(w, 2 * node + 1);
    }
    child(w, 2 * node, k + 1, s_time + tl);
    child(w, 2 * node + 1, k + 1, s_time + tr);
}

static void root(cell_walk *w)
{
    const oracle_setup *s = w->s;
    uint32_t blk[4];
    oracle_block(s->seed, w->cell, 0, 0, blk);
    double ut = oracle_u01_closed_open(blk[0], 0u);
    double U = oracle_u01_open_closed(blk[1], 0u);
    uint32_t j = 0;
    while (j + 1 < s->T && !(ut < s->cum[j]))
        ++j;
    w->j = j;
    w->mean = s->mean[j];
    w->sd = s->sd[j];
    w->st[ORACLE_STAT_CELLS] += 1;
    double v = s->v[w->row];
    if (w->mean < 0.0) {
        /* quiescent: timer -1, always out of time (proliferation.cu:406-409) */
        if (v >= s->phi)
            record(w, 0);
        else
            w->st[ORACLE_STAT_LOST] += 1;
        return;
    }
    /* first timer (cell.cu:106-122): two candidates from words 2 and 3 of
     * the root block, then node 1's redraw stream */
    const uint32_t pw[4] = {blk[2], 0u, blk[3], 0u};
    double z0, z1;
    oracle_normal_pair(pw, &z0, &z1);
    double tau0 = fma(z0, w->sd, w->mean);
    if (!(tau0 > 0.0)) {
        w->st[ORACLE_STAT_REJECTED] += 1;
        tau0 = fma(z1, w->sd, w->mean);
        if (!(tau0 > 0.0)) {
            w->st[ORACLE_STAT_REJECTED] += 1;
            tau0 = truncated_timer(w, 1);
        }
    }
    double t0 = tau0 * U;
    double s0 = t0 + tau0;
    if (s0 > s->t_max) {
        if (v >= s->phi)
            record(w, 0);
        else
            w->st[ORACLE_STAT_LOST] += 1;   /* value below phi: not in the output bin set */
    } else if (w->kcap == 0) {
        drop(w, 0);
    } else {
        divide(w, 1, 0, s0);
    }
}

int oracle_simulate(const oracle_problem *p, int64_t *tally, int64_t *drops,
                    uint64_t *stats, int n_threads)
{
    oracle_setup s;
    int rc = setup_build(p, &s);
    if (rc != OR_OK)
        return rc;
    uint64_t kdim = (uint64_t)s.kmax + 1;
    uint64_t tsz = s.n_nz * kdim * (uint64_t)s.T;
    uint64_t dsz = s.n_nz * kdim;
    if (tally)
        memset(tally, 0, tsz * sizeof(int64_t));
    if (drops)
        memset(drops, 0, dsz * sizeof(int64_t));
    uint64_t tot[ORACLE_N_STATS] = {0};
    const uint64_t chunk = 2048;
    uint64_t n_chunks = (s.n_cells + chunk - 1) / chunk;
    if (n_threads <= 0)
        n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
    {
        int64_t *lt = (int64_t *)calloc(tsz ? tsz : 1, sizeof(int64_t));
        int64_t *ld = drops ? (int64_t *)calloc(dsz ? dsz : 1, sizeof(int64_t)) : NULL;
        uint64_t lst[ORACLE_N_STATS] = {0};
        cell_walk w;
        memset(&w, 0, sizeof w);
        w.s = &s;
        w.tally = lt;
        w.drops = ld;
        w.st = lst;
#pragma omp for schedule(dynamic, 1)
        for (uint64_t c = 0; c < n_chunks; ++c) {
            uint64_t c0 = c * chunk;
            uint64_t c1 = c0 + chunk < s.n_cells ? c0 + chunk : s.n_cells;
            /* last row whose first cell is <= c0 */
            uint64_t lo = 0, hi = s.n_nz;
            while (hi - lo > 1) {
                uint64_t mid = (lo + hi) / 2;
                if (s.start[mid] <= c0)
                    lo = mid;
                else
                    hi = mid;
            }
            uint64_t row = lo;
            for (uint64_t cell = c0; cell < c1; ++cell) {
                while (cell >= s.start[row + 1])
                    ++row;
                w.cell = cell;
                w.row = row;
                w.kcap = s.kcap[row];
                root(&w);
            }
        }
#pragma omp critical
        {
            if (tally)
                for (uint64_t i = 0; i < tsz; ++i)
                    tally[i] += lt[i];
            if (drops)
                for (uint64_t i = 0; i < dsz; ++i)
                    drops[i] += ld[i];
            for (int i = 0; i < ORACLE_N_STATS; ++i)
                tot[i] += lst[i];
        }
        free(lt);
        free(ld);
    }
    if (stats)
        memcpy(stats, tot, sizeof tot);
    setup_free(&s);
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* Relaxed (differentiable) estimator and its gradient.                 */
/* ------------------------------------------------------------------ */
/*
 * This project's own relaxation of the plain method (no reference
 * counterpart): every type j of an initial cell is followed with weight
 * p_j (exact ratio weighting instead of a type draw); timers are the plain
 * method's own, mean_j + sd_j z with z the first candidate > 0 of the plain
 * stream (the node's pair block, then the daughter's redraw stream, as in
 * divide() and root() above); a node whose division would complete at s
 * divides with weight d = sigmoid((t_max - s) / beta) and is recorded with
 * weight 1 - d, so a node's weight W is the product of its ancestors' d.
 * Fluorescence stays exact: a node at its row's generation cap records
 * W (1 - d) and drops the rest; at generation 0 a row below phi records
 * nothing (proliferation.cu:317-381 with the hard test replaced).
 * Subtrees whose weight W d falls below 1e-15 are not followed.
 *
 * With beta -> 0 and one type the estimator is the plain tally of the
 * same seed.
 *
 * The gradient of L with respect to (mean_j, sd_j, p_j), given
 * g = dL/dh, is accumulated along the same traversal with forward
 * tangents: ds/dmean = (1 + U) + k, ds/dsd = z_0 (1 + U) + z_1 + ... + z_k
 * (the accepted candidates' z; which candidate is accepted is held fixed,
 * the pathwise derivative) and dlog W/dtheta = -sum over ancestors of
 * (1 - d) / beta ds/dtheta.
 */
#define RELAXED_EPS 1e-15

typedef struct {
    const oracle_setup *s;
    const oracle_problem *p;
    double beta;
    uint64_t cell, row;
    uint32_t j, T;
    double pj, mean, sd;
    int kcap;
    int ge_phi;
    const double *g;       /* [n_nz][kdim][T] or NULL */
    double *h;             /* this thread's [n_nz][kdim][T] or NULL */
    double *grad;          /* this thread's [3][T] */
} relaxed_walk;

static double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

/* The timer of a relaxed walk from candidate z: kept when > 0, else the
 * first candidate > 0 of the redraw stream of `node` (attempts 1, 2, ...,
 * two candidates each, as truncated_timer()); *z becomes the accepted
 * candidate (0 after 65535 attempts, where the timer is the mean). */
static double relaxed_timer(const relaxed_walk *w, uint64_t node, double *z)
{
    double t = fma(*z, w->sd, w->mean);
    if (t > 0.0)
        return t;
    uint32_t blk[4];
    double y0, y1;
    for (uint32_t a = 1; a < 65536u; ++a) {
        oracle_block(w->s->seed, w->cell, node, a, blk);
        oracle_normal_pair(blk, &y0, &y1);
        t = fma(y0, w->sd, w->mean);
        if (t > 0.0) {
            *z = y0;
            return t;
        }
        t = fma(y1, w->sd, w->mean);
        if (t > 0.0) {
            *z = y1;
            return t;
        }
    }
    *z = 0.0;
    return w->mean;
}

static void relaxed_node(relaxed_walk *w, uint64_t node, int k, double s, double a, double b,
                         double W, double gm, double gs)
{
    const uint64_t kdim = (uint64_t)w->s->kmax + 1;
    const double d = sigmoid((w->s->t_max - s) / w->beta);
    const double rec = W * (1.0 - d);
    if (k > 0 || w->ge_phi) {
        const uint64_t idx = (w->row * kdim + (uint64_t)k) * w->T + w->j;
        if (w->h)
            w->h[idx] += w->pj * rec;
        if (w->g) {
            const double g = w->g[idx];
            const double q = d * (1.0 - d) / w->beta;
            w->grad[0 * w->T + w->j] += w->pj * g * W * ((1.0 - d) * gm + q * a);
            w->grad[1 * w->T + w->j] += w->pj * g * W * ((1.0 - d) * gs + q * b);
            w->grad[2 * w->T + w->j] += g * rec;
        }
    }
    if (k == w->kcap || W * d < RELAXED_EPS)
        return;
    uint32_t blk[4];
    double z0, z1;
    oracle_block(w->s->seed, w->cell, node, 0, blk);
    oracle_normal_pair(blk, &z0, &z1);
    const double f = (1.0 - d) / w->beta;
    const double Wc = W * d, gmc = gm - f * a, gsc = gs - f * b;
    const double tl = relaxed_timer(w, 2 * node, &z0);
    const double tr = relaxed_timer(w, 2 * node + 1, &z1);
    relaxed_node(w, 2 * node, k + 1, s + tl, a + 1.0, b + z0, Wc, gmc, gsc);
    relaxed_node(w, 2 * node + 1, k + 1, s + tr, a + 1.0, b + z1, Wc, gmc, gsc);
}

static void relaxed_cell(relaxed_walk *w)
{
    uint32_t blk[4];
    oracle_block(w->s->seed, w->cell, 0, 0, blk);
    const double U = oracle_u01_open_closed(blk[1], 0u);
    const uint32_t pw[4] = {blk[2], 0u, blk[3], 0u};
    double z0, z1;
    oracle_normal_pair(pw, &z0, &z1);
    const uint64_t kdim = (uint64_t)w->s->kmax + 1;
    for (uint32_t j = 0; j < w->T; ++j) {
        w->j = j;
        w->pj = w->p->proportion[j];
        w->mean = w->p->mean[j];
        w->sd = w->p->sd[j];
        if (w->mean < 0.0) {
            /* quiescent: always out of time, recorded at generation 0 */
            if (w->ge_phi) {
                const uint64_t idx = (w->row * kdim) * w->T + j;
                if (w->h)
                    w->h[idx] += w->pj;
                if (w->g)
                    w->grad[2 * w->T + j] += w->g[idx];
            }
            continue;
        }
        /* the root's candidates z0 then z1 (cell.cu:106-122), then node 1's
         * redraw stream, as in root() */
        double z = z0;
        double tau0 = fma(z, w->sd, w->mean);
        if (!(tau0 > 0.0)) {
            z = z1;
            tau0 = relaxed_timer(w, 1, &z);
        }
        relaxed_node(w, 1, 0, tau0 * (1.0 + U), 1.0 + U, z * (1.0 + U), 1.0, 0.0, 0.0);
    }
}

int oracle_relaxed(const oracle_problem *p, double beta, const double *g, double *h, double *grad,
                   int n_threads)
{
    if (!(beta > 0.0) || !isfinite(beta))
        return OR_E_INVALID;
    oracle_setup s;
    int rc = setup_build(p, &s);
    if (rc != OR_OK)
        return rc;
    const uint32_t T = p->n_types;
    const uint64_t kdim = (uint64_t)s.kmax + 1;
    const uint64_t hsz = s.n_nz * kdim * (uint64_t)(T ? T : 1);
    if (h)
        memset(h, 0, hsz * sizeof(double));
    if (grad)
        memset(grad, 0, 3 * (size_t)T * sizeof(double));
    if (n_threads <= 0)
        n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
    {
        double *lh = h ? (double *)calloc(hsz ? hsz : 1, sizeof(double)) : NULL;
        double *lg = (double *)calloc(3 * (size_t)(T ? T : 1), sizeof(double));
        relaxed_walk w;
        memset(&w, 0, sizeof w);
        w.s = &s;
        w.p = p;
        w.beta = beta;
        w.T = T;
        w.g = g;
        w.h = lh;
        w.grad = lg;
#pragma omp for schedule(dynamic, 64)
        for (uint64_t row = 0; row < s.n_nz; ++row) {
            w.row = row;
            w.kcap = s.kcap[row];
            w.ge_phi = s.v[row] >= s.phi;
            for (uint64_t c = s.start[row]; c < s.start[row + 1]; ++c) {
                w.cell = c;
                relaxed_cell(&w);
            }
        }
#pragma omp critical
        {
            if (h)
                for (uint64_t i = 0; i < hsz; ++i)
                    h[i] += lh[i];
            if (grad)
                for (uint64_t i = 0; i < 3 * (uint64_t)T; ++i)
                    grad[i] += lg[i];
        }
        free(lh);
        free(lg);
    }
    setup_free(&s);
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* File formats (parser.cu:68-154 read, :156-185 types, :187-217 write). */
/* ------------------------------------------------------------------ */
static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static int read_histogram(const char *path, double **vals, uint64_t **cnts, uint64_t *n)
{
    FILE *f = fopen(path, "r");
    uint64_t cap = 64, m = 0;
    double v;
    uint64_t c;
    *vals = (double *)malloc(cap * sizeof(double));
    *cnts = (uint64_t *)malloc(cap * sizeof(uint64_t));
    if (f) {
        while (fscanf(f, "%lf", &v) == 1 && fscanf(f, "%" SCNu64, &c) == 1) {
            if (m == cap) {
                cap *= 2;
                *vals = (double *)realloc(*vals, cap * sizeof(double));
                *cnts = (uint64_t *)realloc(*cnts, cap * sizeof(uint64_t));
            }
            (*vals)[m] = v;
            (*cnts)[m] = c;
            ++m;
        }
        fclose(f);
    }
    *n = m;
    return OR_OK;
}

static int read_types(const char *path, double **p, double **mu, double **sd, uint32_t *n)
{
    FILE *f = fopen(path, "r");
    uint32_t cap = 8, m = 0;
    double a, b, c;
    *p = (double *)malloc(cap * sizeof(double));
    *mu = (double *)malloc(cap * sizeof(double));
    *sd = (double *)malloc(cap * sizeof(double));
    if (f) {
        while (fscanf(f, "%lf %lf %lf", &a, &b, &c) == 3) {
            if (m == cap) {
                cap *= 2;
                *p = (double *)realloc(*p, cap * sizeof(double));
                *mu = (double *)realloc(*mu, cap * sizeof(double));
                *sd = (double *)realloc(*sd, cap * sizeof(double));
            }
            (*p)[m] = a;
            (*mu)[m] = b;
            (*sd)[m] = c;
            ++m;
        }
        fclose(f);
    }
    *n = m;
    return OR_OK;
}

int oracle_run_files(const char *histogram_path, const char *types_path,
                     double t_max, double phi, int track_ratio, uint64_t seed,
                     const char *output_path, int n_threads)
{
    double *vals = NULL, *prop = NULL, *mu = NULL, *sd = NULL;
    uint64_t *cnts = NULL, n_rows = 0;
    uint32_t n_types = 0;
    read_histogram(histogram_path, &vals, &cnts, &n_rows);
    read_types(types_path, &prop, &mu, &sd, &n_types);
    oracle_problem pr;
    pr.values = vals; pr.counts = cnts; pr.n_rows = n_rows;
    pr.proportion = prop; pr.mean = mu; pr.sd = sd; pr.n_types = n_types;
    pr.t_max = t_max; pr.phi = phi; pr.seed = seed;
    int rc = OR_OK;
    int kmax = oracle_max_generation(&pr);
    double rphi = oracle_resolve_phi(&pr);
    uint64_t n_nz = 0;
    for (uint64_t i = 0; i < n_rows; ++i)
        n_nz += cnts[i] > 0;
    int64_t *tally = NULL;
    double *bins = NULL;
    int64_t *bc = NULL, *bt = NULL;
    uint64_t nb = 0;
    FILE *out = NULL;
    if (kmax < 0) {
        rc = kmax;
        goto done;
    }
    uint64_t kdim = (uint64_t)kmax + 1;
    tally = (int64_t *)calloc(n_nz * kdim * (n_types ? n_types : 1) + 1, sizeof(int64_t));
    rc = oracle_simulate(&pr, tally, NULL, NULL, n_threads);
    if (rc != OR_OK)
        goto done;
    /* output bin set: every value v/2^k >= phi of every nonzero row */
    uint64_t cap = 16;
    bins = (double *)malloc(cap * sizeof(double));
    for (uint64_t i = 0; i < n_rows; ++i) {
        if (cnts[i] == 0)
            continue;
        double f = vals[i];
        while (f >= rphi) {
            if (nb == cap) {
                cap *= 2;
                bins = (double *)realloc(bins, cap * sizeof(double));
            }
            bins[nb++] = f;
            f = f / 2;
        }
    }
    qsort(bins, nb, sizeof(double), cmp_double);
    uint64_t u = 0;
    for (uint64_t i = 0; i < nb; ++i)
        if (u == 0 || bins[i] != bins[u - 1])
            bins[u++] = bins[i];
    nb = u;
    bc = (int64_t *)calloc(nb + 1, sizeof(int64_t));
    bt = (int64_t *)calloc((nb + 1) * (n_types ? n_types : 1), sizeof(int64_t));
    uint64_t r = 0;
    for (uint64_t i = 0; i < n_rows; ++i) {
        if (cnts[i] == 0)
            continue;
        double f = vals[i];
        for (uint64_t k = 0; k < kdim; ++k) {
            int64_t tot = 0;
            for (uint32_t j = 0; j < n_types; ++j)
                tot += tally[(r * kdim + k) * n_types + j];
            if (tot) {
                double *hit = (double *)bsearch(&f, bins, nb, sizeof(double), cmp_double);
                if (hit) {
                    uint64_t b = (uint64_t)(hit - bins);
                    bc[b] += tot;
                    for (uint32_t j = 0; j < n_types; ++j)
                        bt[b * n_types + j] += tally[(r * kdim + k) * n_types + j];
                }
            }
            f = f / 2;
        }
        ++r;
    }
    out = output_path ? fopen(output_path, "w") : stdout;
    if (!out) {
        rc = OR_E_IO;
        goto done;
    }
    for (uint64_t b = 0; b < nb; ++b) {
        if (bc[b] <= 0)
            continue;
        fprintf(out, "%.10g\t%" PRId64, bins[b], bc[b]);
        if (track_ratio)
            for (uint32_t j = 0; j < n_types; ++j)
                fprintf(out, "\t%" PRId64, bt[b * n_types + j]);
        fputc('\n', out);
    }
    if (out != stdout)
        fclose(out);
done:
    free(vals); free(cnts); free(prop); free(mu); free(sd);
    free(tally); free(bins); free(bc); free(bt);
    return rc;
}
