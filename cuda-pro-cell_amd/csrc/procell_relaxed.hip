// procell_relaxed.hip -- the relaxed (differentiable) estimator of the plain
// method's expected tally, and its gradient with respect to each type's
// (mean, sd, proportion).  No reference counterpart: the reference
// (src/simulation/proliferation.cu) only samples.
//
// Relaxation of the plain method (proliferation.cu:26-215, cell.cu:80-125):
//   * every initial cell takes every type j with weight p_j (no type draw);
//   * timers are the plain method's own, mean_j + sd_j z with z the first
//     candidate > 0 of the plain stream: the root's pair from block
//     (cell, 0, 0) words {2, -, 3, -}, then node 1's redraw stream; a node's
//     daughters' pair from block (cell, node, 0), a daughter whose candidate
//     is <= 0 continuing in its own node's redraw stream (with beta -> 0 and
//     one type the estimator is the plain tally of the same seed);
//     the root's phase U comes from word 1 of block (cell, 0, 0);
//   * the hard test "completes by t_max" becomes d = sigmoid((t_max - s)/beta)
//     for a node whose division would complete at s;
//   * a node of weight W records W (1 - d) at (row, generation k, type j) --
//     generation 0 only when the row's value is >= phi -- and its daughters
//     carry weight W d; fluorescence stays value / 2^k and the generation cap
//     is the plain method's;
//   * a node is expanded while k < cap and W d >= 1e-15.
// Along a path the completion time is s_k = tau_0 (1 + U) + tau_1 + ... +
// tau_k, so ds/dmean = (1 + U) + k, ds/dsd = z_0 (1 + U) + z_1 + ... + z_k,
// and dlog W/dtheta = -sum over ancestors of (1 - d)/beta ds/dtheta.  The
// backward pass recomputes every subtree from the Philox counters (nothing
// of the forward pass is stored) and accumulates
//   dL/dmean_j += p_j g W ((1 - d) dlogW/dmean + d (1 - d)/beta ds/dmean),
//   dL/dsd_j   += p_j g W ((1 - d) dlogW/dsd   + d (1 - d)/beta ds/dsd),
//   dL/dp_j    += g W (1 - d)
// for the cotangent g = dL/dh of each recorded entry.
//
// Kernel shape: one wavefront per workgroup, a static range of (cell,
// proliferating type) walks per wavefront balanced by an estimate of their
// cost, one walk per lane at a time.  A lane runs its walk depth first and
// evaluates both daughters of a node as it expands it, so every loop step of
// every busy lane is one expansion (one Philox block, one Box-Muller pair,
// two sigmoids); expandable right daughters wait in a per-lane stack in
// device memory, one 64-byte line per frame laid out [wavefront][lane][level]
// (lanes of a wavefront sit at different levels, so a field-major layout
// would write a partial line per field and lane).  Recorded weights
// go to an LDS window over the rows of the wavefront's range (the cotangent
// window in the backward pass) and reach the tally with FP64 atomics at the
// end.
//
// A plan covers one parameter set (procell_relaxed_*) or many (the batched
// calibration, procell_calib_*): sets' rows, cells and tally blocks are
// concatenated and every wavefront's range lies in one set.  A cell stride k
// (a power of two) restricts a pass to the initial cells of one residue
// class mod k.  A calibration step -- fill, forward pass, loss and cotangent,
// adjoint pass, Adam -- is captured once into a hipGraph and replayed.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "procell_internal.hpp"
#include "procell_stream.hpp"

namespace procell {
namespace {

constexpr double kRelaxEps = 1e-15;        // smallest W d that is still expanded
// Expansions a single walk may take: a walk that needs more is cut short and
// counted, and the call reports PROCELL_E_DEPTH (parameters far outside the
// estimator's domain, e.g. sd a few times the mean with deep rows, where most
// timers lie near 0, would otherwise run for practically ever).
constexpr uint32_t kWalkBudget = 1u << 22;
// Calibration keeps sd <= 0.8 mean (log 0.8): the expected number of
// generation-k nodes, 2^k P(s_k <= t), grows with the timers' density near 0
// (for untruncated normal timers it would diverge beyond sd / mean = 0.85,
// where mean^2 / (2 sd^2) = ln 2), so the bound keeps the walks cheap.
constexpr double kLogMaxSdRatio = -0.22314355131420976;
// A stacked node: s, W, d (both passes), b, dlogW/dmean, dlogW/dsd (adjoint),
// padded to one 64-byte line so that a push or a pop of a lane moves one line.
constexpr uint32_t kFrameDoubles = 8;
// The adjoint pass (without a cell stride) keeps its two most recently pushed
// frames (the six tangent fields) in LDS, field-major so that a wavefront's
// accesses are bank-conflict free; only older frames reach the device stack.
// Its cotangent window is smaller, so that window and cache fit the LDS of
// four wavefronts per SIMD.  256 sets x 2e5 cells: adjoint 13.4 -> 12.6 ms,
// calibration step 23.3 -> 22.5 ms (one slot measured slower).  The Jacobian
// pass without a stride keeps one frame beside a 768-double window: HBM
// written per expansion 26.5 -> 10.6 B, read 25.7 -> 8.7 B, for 1% more time
// (14.09 -> 14.24 ms; two slots beside 384 doubles measured 2% slower, the
// cache in the stride-4 Jacobian pass 7% slower).
// Idle lanes a wavefront tolerates before it claims walks for them: a claim
// costs the wavefront about 50 instructions, so claiming for every idle lane
// at every step costs more than the idle lanes.  256 sets x 2e5 cells
// (with the merged redraw loop): calibration step 23.96 ms (0), 23.38 (4),
// 23.33 (8), 24.06 (16).
constexpr uint32_t kRefillIdle = 8;
constexpr uint32_t kTanFields = 6;
constexpr uint32_t kTanSlots = 2;  // adjoint
constexpr uint32_t kJacSlots = 1;  // Jacobian pass
constexpr uint32_t kWindowDoubles = 1024;  // LDS window of a wavefront (8 KB)
constexpr uint32_t kWindowAdj = 384;       // window of the adjoint pass beside its frame cache (3 KB)
constexpr uint32_t kWindowJac = 768;       // window of the Jacobian pass beside its frame cache (6 KB)
constexpr uint32_t kRowCap = 64;           // rows of a wavefront's range staged in LDS
// Wavefronts of a plan over several sets, in resident wavefronts: short
// ranges cover few rows, so the tallies stay in the LDS window (the
// Jacobian pass keeps three), and the dispatcher evens out the cost
// estimate's errors.  256 sets x 2e5 cells, 8x against 2x: forward pass
// 12.9 -> 9.8 ms, adjoint 15.9 -> 13.2 ms, Jacobian pass 17.5 -> 13.5 ms,
// calibration step 29.8 -> 23.3 ms; 16x measured level with 8x.
constexpr uint32_t kMultiSetWaveFactor = 8;

// One parameter set of a relaxed plan (a single-set plan has one).  Rows,
// initial cells, tally entries and type blocks of the sets are concatenated.
struct RelaxSet {
    uint64_t seed;
    double t_max;
    uint64_t cell_base;    // first initial cell (stream cell ids are relative to it)
    uint64_t h_base;       // first tally entry of the set's [rows][kdim][n_types]
    uint32_t row_base;     // first nonzero row
    uint32_t n_rows;
    uint32_t kdim;         // generations + 1
    uint32_t n_types;
    uint32_t theta_base;   // first entry of the set's [3][n_types] block of theta / gradient
    uint32_t prolif_base;  // first entry in the proliferating-type list
    uint32_t n_prolif;     // proliferating types (mean >= 0 when the plan was made)
    uint32_t quiet_base;
    uint32_t n_quiet;
    uint32_t pad[3];
};

struct RelaxArgs {
    double beta;
    uint32_t n_sets;
    uint32_t n_rows;           // all sets
    uint32_t kdim_max;         // stack levels per lane
    uint32_t pad;
    const RelaxSet *sets;
    const uint32_t *wave_set;  // [n_waves] set of each wavefront
    const uint64_t *wave_unit; // [n_waves][2] first and end walk of the range (set-local)
    const uint32_t *wave_row;  // [n_waves][2] first and last (global) row of the range
    const uint32_t *row_start; // [n_rows + 1] first initial cell of each row (all sets)
    const uint32_t *row_info;  // [n_rows] generation cap | (value >= phi) << 8
    const uint32_t *row_set;   // [n_rows] set of each row
    const uint32_t *prolif;    // file index of each proliferating type, per set
    const uint32_t *quiet;     // file index of each quiescent type, per set
    const double *theta;       // per set [3][n_types]: mean, sd, proportion, file order
    double *stack;             // [n_waves][64 lanes][kdim_max] frames of kFrameDoubles
    double *h;                 // per set [rows][kdim][n_types] relaxed tally (forward)
    const double *g;           // cotangent, layout of h (backward)
    double *grad;              // per set [3][n_types] gradient (backward)
    unsigned long long *expansions;  // [2] expansions of the call (roots drawn + nodes expanded), walks cut
    unsigned long long *set_cuts;    // [n_sets] walks cut per set (NULL: not kept)
    // Cell stride: a pass walks only the initial cells c (set-local id) with
    // (c + phase) & stride_mask == 0; phase = (*step + phase) & stride_mask
    // when step is set (calibration: the step counter, the adjoint pass one
    // class ahead of the forward pass), else phase & stride_mask.
    const unsigned long long *step;
    uint32_t stride_mask;
    uint32_t phase;
    // Jacobian pass: [2][h_total] dh/dmean_j and dh/dsd_j of each entry's own
    // type j (h of type j depends on type j's parameters only)
    double *jac;
    uint64_t h_total;
};

// What a pass of the walk computes.
enum WalkMode : int {
    kForward = 0,   // h
    kAdjoint = 1,   // gradient of <g, h>
    kJacobian = 2,  // h and its derivatives in each type's mean and sd
};

// 1 / x for x >= 1: the hardware reciprocal estimate refined by two Newton
// steps (each squares the relative error), within an ulp or two of the
// quotient.
// exp(x) for -40 <= x <= 700 (below -40, 1 + exp(x) rounds to 1): x = k ln 2
// + r with |r| <= ln 2 / 2 (Cody-Waite reduction, ln 2 in two parts), exp(r)
// by its Taylor polynomial of degree 12 (truncation below 2e-16 relative),
// times 2^k.  Horner steps p = fma(p, r, c) keep each coefficient in its
// register (no copy before a two-address fma).
// p r + c as one three-address fma (the compiler otherwise ties the result
// to a copy of the loop-invariant coefficient)
__device__ __forceinline__ double horner(double p, double r, double c)
{
    double out;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(out) : "v"(p), "v"(r), "v"(c));
    return out;
}

__device__ __forceinline__ double exp_sig(double x)
{
    const double k = __builtin_rint(x * 1.4426950408889634);
    double r = fma(k, -6.93147180369123816490e-01, x);
    r = fma(k, -1.90821492927058770002e-10, r);
    double p = 2.08767569878681e-09;
    p = horner(p, r, 2.505210838544172e-08);
    p = horner(p, r, 2.755731922398589e-07);
    p = horner(p, r, 2.7557319223985893e-06);
    p = horner(p, r, 2.48015873015873e-05);
    p = horner(p, r, 0.0001984126984126984);
    p = horner(p, r, 0.001388888888888889);
    p = horner(p, r, 0.008333333333333333);
    p = horner(p, r, 0.041666666666666664);
    p = horner(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

__device__ __forceinline__ double recip_ge1(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// sigmoid(y / beta) with ib = 1 / beta.  The exponent is capped at 700 so
// that 1 + e stays finite (an infinite one would turn the Newton steps into
// inf x 0); there the sigmoid is below 1e-304 either way.
__device__ __forceinline__ double sigmoid_scaled(double y, double ib)
{
    return recip_ge1(1.0 + exp_sig(fmin(fmax(-(y * ib), -40.0), 700.0)));
}

template <int Mode, bool Strided>
__device__ __forceinline__ void relaxed_walks(const RelaxArgs &a)
{
    constexpr bool Grad = Mode == kAdjoint;
    constexpr bool Jac = Mode == kJacobian;
    constexpr bool Tan = Grad || Jac;  // path tangents carried (six-field frames)
    // LDS: [kWindowDoubles] window | [3][T] gradient sums | [3][T] theta of the set |
    //      u32 [kRowCap + 1] row starts | [kRowCap] row info | [T] proliferating types
    extern __shared__ double lds[];
    const uint32_t lane = threadIdx.x;
    const uint32_t w = blockIdx.x;
    const uint32_t set_id = a.wave_set[w];
    const RelaxSet S = a.sets[set_id];
    const uint32_t T = S.n_types;
    const uint32_t KT = S.kdim * T;
    uint64_t next = a.wave_unit[2 * w];
    uint64_t u_end = a.wave_unit[2 * w + 1];
    if (next >= u_end)
        return;
    const uint32_t r_lo = a.wave_row[2 * w], r_hi = a.wave_row[2 * w + 1];
    const uint32_t win_n = (r_hi - r_lo + 1) * KT;
    // window: h (forward), g (adjoint) or h | dh/dmean | dh/dsd (Jacobian)
    const uint32_t win_t = Jac ? 3 * win_n : win_n;
    constexpr bool kLdsCache = (Grad || Jac) && !Strided;  // frames cached in LDS
    constexpr uint32_t kWin = kLdsCache ? (Jac ? kWindowJac : kWindowAdj) : kWindowDoubles;
    constexpr uint32_t kSlots = Jac ? kJacSlots : kTanSlots;
    const bool in_lds = win_t <= kWin;
    const uint64_t hw0 = S.h_base + (uint64_t)(r_lo - S.row_base) * KT;  // tally index of the window
    double *gsum = lds + kWin;
    double *theta = gsum + 3 * T;
    uint32_t *rs = (uint32_t *)(theta + 3 * T);
    uint32_t *ri = rs + kRowCap + 1;
    uint32_t *prolif = ri + kRowCap;
    const uint32_t nr = r_hi - r_lo + 1;
    const bool rows_lds = nr <= kRowCap;
    if (in_lds)
        for (uint32_t i = lane; i < win_t; i += 64)
            lds[i] = Grad ? a.g[hw0 + i] : 0.0;
    for (uint32_t i = lane; i < 3 * T; i += 64) {
        gsum[i] = 0.0;
        theta[i] = a.theta[S.theta_base + i];
    }
    for (uint32_t i = lane; i < S.n_prolif; i += 64)
        prolif[i] = a.prolif[S.prolif_base + i];
    if (rows_lds) {
        for (uint32_t i = lane; i <= nr; i += 64)
            rs[i] = a.row_start[r_lo + i];
        for (uint32_t i = lane; i < nr; i += 64)
            ri[i] = a.row_info[r_lo + i];
    }
    __syncthreads();
    // row r's start and info are row_start[r - roff], row_info[r - roff]
    const uint32_t roff = rows_lds ? r_lo : 0u;
    const uint32_t *row_start = rows_lds ? rs : a.row_start;
    const uint32_t *row_info = rows_lds ? ri : a.row_info;
    const bool small_units = u_end <= 0xFFFFFFFFull;

    const dev::Coeffs K = dev::coeffs();
    const dev::Tables tb = dev::global_tables();
    const double t_max = S.t_max, ib = 1.0 / a.beta;
    const uint32_t n_prolif = S.n_prolif;
    double *stk = a.stack + ((uint64_t)w * 64 + lane) * a.kdim_max * kFrameDoubles;
    // LDS frame cache of the adjoint pass: field f of slot q of this lane
    double *fc = (double *)(((uintptr_t)(prolif + T) + 7) & ~(uintptr_t)7) + lane;
    auto fcp = [&](uint32_t q, uint32_t f) -> double & { return fc[(q * kTanFields + f) * 64]; };

    bool active = false;
    uint64_t cell = 0, node = 0, pending = 0;
    uint32_t row = r_lo, j = 0, kcap = 0, k = 0;
    bool ge_phi = false;
    double pj = 0.0, mean = 0.0, sd = 0.0, u1 = 0.0;  // u1 = 1 + U of the walk's cell
    double s = 0.0, b = 0.0, W = 0.0, d = 0.0, gm = 0.0, gs = 0.0;
    double acc_m = 0.0, acc_s = 0.0, acc_p = 0.0;
    uint32_t row_cursor = r_lo;  // row of the first walk not yet taken (wave-uniform)
    uint32_t n_exp = 0, n_cut = 0, walk_exp = 0;
    // The most recently pushed frame stays in registers until another push
    // spills it: it is always the deepest pending level, i.e. the next one
    // popped, so most push/pop pairs never touch memory (a store would also
    // hold up the wavefront's next wait on its loads).
    // (forward pass only: the adjoint's six-field frame would cost it a
    // wavefront per SIMD in registers)
    constexpr bool kCache = !Tan;
    bool c_ok = false;
    uint32_t c_lvl = 0;
    double c_s = 0.0, c_W = 0.0, c_d = 0.0;
    // adjoint and Jacobian passes: frames cached in LDS slots (c_n of them, c_top the
    // deepest); the levels of slots 0 and 1 in t_lvl0, t_lvl1
    static_assert(kSlots == 1 || kSlots == 2, "one or two cached frames");
    uint32_t c_n = 0, c_top = 0, t_lvl0 = 0, t_lvl1 = 0;
    const uint32_t smask = Strided ? a.stride_mask : 0u;
    const uint32_t sphase = Strided ? ((a.step ? (uint32_t)a.step[0] : 0u) + a.phase) & smask : 0u;
    // With a cell stride the range is enumerated over the selected cells only:
    // walk index v stands for local cell soff + stride * (v / n_prolif), type
    // v mod n_prolif (soff: the first cell c with (c + phase) a multiple of
    // the stride), so every claim takes a walk.
    const uint32_t sshift = Strided ? (uint32_t)__popc(smask) : 0u;
    const uint64_t soff = Strided ? (uint64_t)((0u - sphase) & smask) : 0ull;
    if (Strided) {
        auto compress = [&](uint64_t u) -> uint64_t {
            const uint64_t c = u / n_prolif, t = u - c * n_prolif;
            const uint64_t below = c > soff ? (c - soff + smask) >> sshift : 0ull;  // selected cells < c
            return below * n_prolif + ((((uint32_t)c + sphase) & smask) == 0u ? t : 0ull);
        };
        next = compress(next);
        u_end = compress(u_end);
    }

    for (;;) {
        // ---- refill idle lanes from the wavefront's range once more than
        //      kRefillIdle are idle (or all are) ----
        {
            const uint64_t idle = __ballot(!active);
            const bool refill = idle && next < u_end && (__popcll(idle) > kRefillIdle || ~idle == 0);
            if (refill && !active) {
                const uint64_t u = next + (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
                const uint64_t wi = small_units ? (uint64_t)((uint32_t)u / n_prolif) : u / n_prolif;
                if (u < u_end) {
                    cell = Strided ? soff + (wi << sshift) : wi;
                    j = prolif[(uint32_t)(u - wi * n_prolif)];
                    const uint64_t gcell = S.cell_base + cell;
                    row = row_cursor;
                    while (row_start[row + 1 - roff] <= gcell)
                        ++row;
                    const uint32_t info = row_info[row - roff];
                    kcap = info & 255u;
                    ge_phi = ((info >> 8) & 1u) != 0;
                    mean = theta[j];
                    sd = theta[T + j];
                    pj = theta[2 * T + j];
                    node = 0;  // the next step draws the root from block (cell, 0, 0)
                    pending = 0;
                    c_ok = false;
                    c_n = 0;
                    walk_exp = 0;
                    active = true;
                }
            }
            if (refill) {
                next = next + (uint64_t)__popcll(idle);
                if (next > u_end)
                    next = u_end;
                if (next < u_end) {
                    const uint64_t wi = small_units ? (uint64_t)((uint32_t)next / n_prolif) : next / n_prolif;
                    const uint64_t c = S.cell_base + (Strided ? soff + (wi << sshift) : wi);
                    while (row_start[row_cursor + 1 - roff] <= c)
                        ++row_cursor;
                }
            }
        }
        if (__ballot(active) == 0)
            break;
        if (!active)
            continue;

        // ---- one expansion: the root (node 0 stands for "not drawn yet") or
        //      both daughters of the current node ----
        const bool pre = node == 0;
        ++n_exp;
        ++walk_exp;
        dev::Words wd = dev::draw_block(S.seed, cell, node, 0u);
        if (pre) {
            u1 = 1.0 + dev::uniform_oc(wd.w1, 0u);
            wd = dev::Words{wd.w2, 0u, wd.w3, 0u};
        }
        double z0, z1;
        dev::normal_pair(wd, z0, z1, K, tb);
        // the plain method's timers (cell.cu:106-122): a candidate <= 0 gives
        // way to the first candidate > 0 of a redraw stream -- the root's to
        // its second candidate, then node 1's stream; a daughter's to its own
        // node's stream; after 65535 blocks the timer is the mean (z = 0)
        {
            bool redoA = !(fma(z0, sd, mean) > 0.0);
            const bool redoB = !pre && !(fma(z1, sd, mean) > 0.0);
            if (__builtin_expect(__ballot(redoA || redoB) != 0, 0)) {
                // the root's second candidate first
                if (pre && redoA && fma(z1, sd, mean) > 0.0) {
                    z0 = z1;
                    redoA = false;
                }
                // one loop for both daughters: each round draws one block per
                // lane, of A's stream while A is pending, then of B's (a lane
                // needing both takes them in turn; the values do not depend on
                // the order); a stream gives up after 65535 blocks (z = 0)
                uint32_t need = (redoA ? 1u : 0u) | (redoB ? 2u : 0u);
                uint32_t att = 1;
                while (__ballot(need != 0) != 0) {
                    if (need != 0) {
                        const bool onA = (need & 1u) != 0;
                        const uint64_t rnode = onA ? (pre ? 1ull : 2 * node) : 2 * node + 1;
                        const dev::Words wr = dev::draw_block(S.seed, cell, rnode, att);
                        double y0, y1;
                        dev::normal_pair(wr, y0, y1, K, tb);
                        const bool ok0 = fma(y0, sd, mean) > 0.0;
                        const bool ok1 = fma(y1, sd, mean) > 0.0;
                        const bool done = ok0 || ok1 || att == 65535u;
                        const double v = ok0 ? y0 : (ok1 ? y1 : 0.0);
                        if (done) {
                            if (onA)
                                z0 = v;
                            else
                                z1 = v;
                            need &= onA ? ~1u : ~2u;
                            att = 1;
                        } else {
                            ++att;
                        }
                    }
                }
            }
        }

        double sA, bA, WA, gmA, gsA, sB = 0.0, bB = 0.0;
        uint32_t kA;
        if (pre) {
            const double tau0 = fma(z0, sd, mean);
            sA = tau0 * u1;
            bA = z0 * u1;
            WA = 1.0;
            gmA = 0.0;
            gsA = 0.0;
            kA = 0;
        } else {
            const double f = (1.0 - d) * ib;
            const double a_cur = u1 + (double)k;
            WA = W * d;
            gmA = gm - f * a_cur;
            gsA = gs - f * b;
            sA = s + fma(z0, sd, mean);
            sB = s + fma(z1, sd, mean);
            bA = b + z0;
            bB = b + z1;
            kA = k + 1;
        }
        const double aA = u1 + (double)kA;
        const double dA = sigmoid_scaled(t_max - sA, ib);
        const double dB = pre ? 0.0 : sigmoid_scaled(t_max - sB, ib);
        const double recA = WA * (1.0 - dA);
        const double recB = WA * (1.0 - dB);
        const bool recordA = kA > 0 || ge_phi;
        const uint32_t col = kA * T + j;
        if (Mode == kForward) {
            // both daughters land in the same (row, generation, type) entry
            const double v = (recordA ? pj * recA : 0.0) + (pre ? 0.0 : pj * recB);
            if (recordA || !pre) {
                if (in_lds)
                    atomicAdd(lds + (row - r_lo) * KT + col, v);
                else
                    unsafeAtomicAdd(a.h + S.h_base + (uint64_t)(row - S.row_base) * KT + col, v);
            }
        } else if (Jac) {
            // the recorded p_j W (1 - d) and its derivatives in mean_j, sd_j:
            // p_j W ((1 - d) dlogW/dtheta + d (1 - d) / beta ds/dtheta)
            double vh = 0.0, vm = 0.0, vs = 0.0;
            if (recordA) {
                const double q = dA * (1.0 - dA) * ib;
                vh = recA;
                vm = WA * ((1.0 - dA) * gmA + q * aA);
                vs = WA * ((1.0 - dA) * gsA + q * bA);
            }
            if (!pre) {
                const double q = dB * (1.0 - dB) * ib;
                vh += recB;
                vm += WA * ((1.0 - dB) * gmA + q * aA);
                vs += WA * ((1.0 - dB) * gsA + q * bB);
            }
            if (recordA || !pre) {
                if (in_lds) {
                    double *e = lds + (row - r_lo) * KT + col;
                    atomicAdd(e, pj * vh);
                    atomicAdd(e + win_n, pj * vm);
                    atomicAdd(e + 2 * win_n, pj * vs);
                } else {
                    const uint64_t e = S.h_base + (uint64_t)(row - S.row_base) * KT + col;
                    unsafeAtomicAdd(a.h + e, pj * vh);
                    unsafeAtomicAdd(a.jac + e, pj * vm);
                    unsafeAtomicAdd(a.jac + a.h_total + e, pj * vs);
                }
            }
        } else {
            const double gv = in_lds ? lds[(row - r_lo) * KT + col]
                                     : a.g[S.h_base + (uint64_t)(row - S.row_base) * KT + col];
            if (recordA) {
                const double q = dA * (1.0 - dA) * ib;
                acc_m += pj * gv * WA * ((1.0 - dA) * gmA + q * aA);
                acc_s += pj * gv * WA * ((1.0 - dA) * gsA + q * bA);
                acc_p += gv * recA;
            }
            if (!pre) {
                const double q = dB * (1.0 - dB) * ib;
                acc_m += pj * gv * WA * ((1.0 - dB) * gmA + q * aA);
                acc_s += pj * gv * WA * ((1.0 - dB) * gsA + q * bB);
                acc_p += gv * recB;
            }
        }
        const bool expA = kA < kcap && WA * dA >= kRelaxEps;
        const bool expB = !pre && kA < kcap && WA * dB >= kRelaxEps;

        // ---- next node: left daughter (right one stacked), the only
        //      expandable daughter, or the deepest stacked node ----
        // a walk past its budget is cut short (counted; the call reports it)
        const bool over = walk_exp >= kWalkBudget;
        if (kCache && expA && expB) {
            if (c_ok) {  // spill the cached frame to its level (no tangents in the forward pass)
                double2 *f = (double2 *)(stk + (uint64_t)c_lvl * kFrameDoubles);
                f[0] = make_double2(c_s, c_W);
                f[1].x = c_d;
            }
            c_ok = true;
            c_lvl = kA;
            c_s = sB;
            c_W = WA;
            c_d = dB;
            pending |= 1ull << kA;
        } else if (kLdsCache && expA && expB) {
            if (c_n == kSlots) {  // spill the oldest cached frame to its level
                const uint32_t bot = kSlots == 1 ? 0u : c_top ^ 1u;
                double2 *f = (double2 *)(stk + (uint64_t)(bot ? t_lvl1 : t_lvl0) * kFrameDoubles);
                f[0] = make_double2(fcp(bot, 0), fcp(bot, 1));
                f[1] = make_double2(fcp(bot, 2), fcp(bot, 3));
                f[2] = make_double2(fcp(bot, 4), fcp(bot, 5));
                --c_n;
            }
            c_top = kSlots == 1 ? 0u : c_top ^ 1u;
            fcp(c_top, 0) = sB;
            fcp(c_top, 1) = WA;
            fcp(c_top, 2) = dB;
            fcp(c_top, 3) = bB;
            fcp(c_top, 4) = gmA;
            fcp(c_top, 5) = gsA;
            t_lvl0 = c_top ? t_lvl0 : kA;
            t_lvl1 = c_top ? kA : t_lvl1;
            ++c_n;
            pending |= 1ull << kA;
        } else if (expA && expB) {
            double2 *f = (double2 *)(stk + (uint64_t)kA * kFrameDoubles);
            f[0] = make_double2(sB, WA);
            f[1] = make_double2(dB, bB);
            f[2] = make_double2(gmA, gsA);
            pending |= 1ull << kA;
        }
        if (!over && (expA || expB)) {
            const bool left = expA;
            node = pre ? 1ull : 2ull * node + (left ? 0ull : 1ull);
            k = kA;
            s = left ? sA : sB;
            b = left ? bA : bB;
            d = left ? dA : dB;
            W = WA;
            gm = gmA;
            gs = gsA;
        } else if (!over && pending) {
            const uint32_t L = 63u - (uint32_t)__clzll(pending);
            pending &= ~(1ull << L);
            node = (node >> (k - L)) | 1ull;
            k = L;
            if (kCache && c_ok) {  // the cached frame is level L
                s = c_s;
                W = c_W;
                d = c_d;
                c_ok = false;
            } else if (kLdsCache && c_n) {  // the deepest cached frame is level L
                s = fcp(c_top, 0);
                W = fcp(c_top, 1);
                d = fcp(c_top, 2);
                b = fcp(c_top, 3);
                gm = fcp(c_top, 4);
                gs = fcp(c_top, 5);
                c_top = kSlots == 1 ? 0u : c_top ^ 1u;
                --c_n;
            } else {
                const double2 *f = (const double2 *)(stk + (uint64_t)L * kFrameDoubles);
                const double2 f0 = f[0];
                s = f0.x;
                W = f0.y;
                if (Tan) {
                    const double2 f1 = f[1], f2 = f[2];
                    d = f1.x;
                    b = f1.y;
                    gm = f2.x;
                    gs = f2.y;
                } else {
                    d = f[1].x;
                }
            }
        } else {
            n_cut += over && (expA || expB || pending) ? 1u : 0u;
            if (Grad) {
                atomicAdd(gsum + j, acc_m);
                atomicAdd(gsum + T + j, acc_s);
                atomicAdd(gsum + 2 * T + j, acc_p);
                acc_m = acc_s = acc_p = 0.0;
            }
            active = false;
        }
    }
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) {
        n_exp += __shfl_xor(n_exp, o);
        n_cut += __shfl_xor(n_cut, o);
    }
    if (lane == 0) {
        atomicAdd(a.expansions, (unsigned long long)n_exp);
        if (n_cut) {
            atomicAdd(a.expansions + 1, (unsigned long long)n_cut);
            if (a.set_cuts)
                atomicAdd(a.set_cuts + set_id, (unsigned long long)n_cut);
        }
    }
    if (Mode == kForward) {
        if (in_lds)
            for (uint32_t i = lane; i < win_n; i += 64) {
                const double v = lds[i];
                if (v != 0.0)
                    unsafeAtomicAdd(a.h + hw0 + i, v);
            }
    } else if (Jac) {
        if (in_lds)
            for (uint32_t i = lane; i < win_t; i += 64) {
                const double v = lds[i];
                const uint32_t part = i / win_n, e = i - part * win_n;
                double *dst = part == 0 ? a.h : a.jac + (uint64_t)(part - 1) * a.h_total;
                if (v != 0.0)
                    unsafeAtomicAdd(dst + hw0 + e, v);
            }
    } else {
        for (uint32_t i = lane; i < 3 * T; i += 64) {
            const double v = gsum[i];
            if (v != 0.0)
                unsafeAtomicAdd(a.grad + S.theta_base + i, v);
        }
    }
}

template <bool Grad>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void relaxed_kernel(RelaxArgs a)
{
    relaxed_walks<Grad ? kAdjoint : kForward, false>(a);
}

// h, dh/dmean and dh/dsd in one pass (optionally with a cell stride).
template <bool Strided>
__global__ __launch_bounds__(64) void relaxed_jacobian_kernel(RelaxArgs a)
{
    relaxed_walks<kJacobian, Strided>(a);
}

// With a cell stride; held to four wavefronts per SIMD (the extra refill
// rounds would otherwise cost the adjoint one).
template <bool Grad>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void relaxed_strided_kernel(RelaxArgs a)
{
    relaxed_walks<Grad ? kAdjoint : kForward, true>(a);
}

// Quiescent types: each initial cell of a row with value >= phi records p_j
// at generation 0 (cell.cu:104-108 gives them timer -1, always out of time).
template <bool Grad>
__global__ void relaxed_quiescent_kernel(RelaxArgs a)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.n_rows)
        return;
    const RelaxSet &S = a.sets[a.row_set[row]];
    if (S.n_quiet == 0 || ((a.row_info[row] >> 8) & 1u) == 0)
        return;
    // cells of the row walked by this pass: local ids c in [lo, hi) with
    // (c + phase) a multiple of the stride
    const uint32_t stride = a.stride_mask + 1u;
    const uint32_t phase = ((a.step ? (uint32_t)a.step[0] : 0u) + a.phase) & a.stride_mask;
    const uint64_t lo = a.row_start[row] - S.cell_base + phase, hi = a.row_start[row + 1] - S.cell_base + phase;
    const double cnt = (double)((hi + stride - 1) / stride - (lo + stride - 1) / stride);
    const uint64_t base = S.h_base + (uint64_t)(row - S.row_base) * S.kdim * S.n_types;
    for (uint32_t q = 0; q < S.n_quiet; ++q) {
        const uint32_t j = a.quiet[S.quiet_base + q];
        if (!Grad)
            a.h[base + j] = cnt * a.theta[S.theta_base + 2 * S.n_types + j];
        else
            unsafeAtomicAdd(a.grad + S.theta_base + 2 * S.n_types + j, cnt * a.g[base + j]);
    }
}

thread_local hipError_t g_relaxed_hip = hipSuccess;

#define RELAX_TRY(expr)                      \
    do {                                     \
        hipError_t e_ = (expr);              \
        if (e_ != hipSuccess) {              \
            g_relaxed_hip = e_;              \
            return PROCELL_E_HIP;            \
        }                                    \
    } while (0)

// mode 0: the forward kernels and the strided passes; 1: the adjoint without
// a stride; 2: the Jacobian pass without a stride (the last two: smaller
// window, frame cache)
size_t relaxed_lds_bytes(uint32_t n_types, int mode)
{
    const uint32_t win = mode == 1 ? kWindowAdj : mode == 2 ? kWindowJac : kWindowDoubles;
    const size_t base = (win + 6ull * n_types) * sizeof(double) + (2ull * kRowCap + 1 + n_types) * sizeof(uint32_t);
    const uint32_t slots = mode == 1 ? kTanSlots : mode == 2 ? kJacSlots : 0u;
    return slots ? (base + 7) / 8 * 8 + (size_t)slots * kTanFields * 64 * sizeof(double) : base;
}

// Expected expansions of one walk of a type with this mean in a row with
// this generation cap: a full binary tree down to the generation the mean
// timer reaches by t_max, plus the relaxed fringe.  Only used to balance the
// wavefronts' ranges.
double walk_cost(uint32_t kcap, double t_max, double mean)
{
    double g = mean > 0.0 ? floor(t_max / mean) : (double)kcap;
    if (!(g >= 0.0))
        g = 0.0;
    g = std::min(g, (double)kcap);
    return ldexp(1.0, (int)g + 1) + 1.0;
}

int relaxed_device(int device, int *n_cu)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
        return PROCELL_E_NO_DEVICE;
    RELAX_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    RELAX_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return PROCELL_E_NO_DEVICE;
    *n_cu = prop.multiProcessorCount;
    return PROCELL_OK;
}

// Device tables of a relaxed plan over n parameter sets.
struct RelaxCore {
    int device = 0;
    uint32_t n_sets = 0;
    std::vector<Setup> setups;
    std::vector<RelaxSet> sets;
    std::vector<uint32_t> quiet_flag;  // per type of every set: 1 = quiescent
    uint32_t n_waves = 0;
    size_t lds_bytes = 0;  // every walk kernel but the two below
    size_t lds_adj = 0;    // the adjoint without a stride
    size_t lds_jac = 0;    // the Jacobian pass without a stride
    // the wavefront layout of the strided and Jacobian passes (equal-cost ranges)
    const uint32_t *eq_set = nullptr;
    const uint64_t *eq_unit = nullptr;
    const uint32_t *eq_row = nullptr;
    bool any_prolif = false, any_quiet = false;
    uint64_t h_total = 0;      // tally entries of all sets
    uint32_t types_total = 0;  // types of all sets
    RelaxArgs args{};
    void *d_block = nullptr;
    double *d_stack = nullptr;
};

void core_free(RelaxCore &c)
{
    (void)hipSetDevice(c.device);
    if (c.d_block)
        (void)hipFree(c.d_block);
    if (c.d_stack)
        (void)hipFree(c.d_stack);
    c.d_block = nullptr;
    c.d_stack = nullptr;
}

// Splits walks [0, n_units) of one set into n_w ranges of equal estimated
// cost (unit u = cell u / n_prolif, type prolif[u % n_prolif]).
void split_set(const Setup &s, const std::vector<double> &unit_cost, uint32_t n_prolif, uint32_t n_w,
               uint32_t row_base, uint64_t *unit_out, uint32_t *row_out, const double *cumw = nullptr)
{
    const uint32_t n_rows = (uint32_t)s.v.size();
    const uint64_t n_units = s.n_cells * (uint64_t)n_prolif;
    double total = 0.0;
    for (uint32_t i = 0; i < n_rows; ++i)
        total += unit_cost[i] * (double)(s.start[i + 1] - s.start[i]) * n_prolif;
    std::vector<uint64_t> cut(n_w + 1, n_units);
    cut[0] = 0;
    uint32_t wv = 1;
    double acc = 0.0;
    for (uint32_t i = 0; i < n_rows && wv < n_w; ++i) {
        const uint64_t u0 = s.start[i] * n_prolif, u1 = s.start[i + 1] * n_prolif;
        const double per = unit_cost[i];
        while (wv < n_w) {
            const double need = (total * (cumw ? cumw[wv - 1] : (double)wv / (double)n_w) - acc) / per;
            if (need >= (double)(u1 - u0))
                break;
            uint64_t c = u0 + (uint64_t)std::max(0.0, ceil(need));
            c = std::min(c, u1);
            cut[wv] = std::max(c, cut[wv - 1]);
            ++wv;
        }
        acc += per * (double)(u1 - u0);
    }
    for (uint32_t v = 1; v <= n_w; ++v)
        cut[v] = std::max(cut[v], cut[v - 1]);
    auto row_of = [&](uint64_t cell) {
        return (uint32_t)(std::upper_bound(s.start.begin(), s.start.end() - 1, cell) - s.start.begin() - 1);
    };
    for (uint32_t v = 0; v < n_w; ++v) {
        unit_out[2 * v] = cut[v];
        unit_out[2 * v + 1] = cut[v + 1];
        row_out[2 * v] = row_out[2 * v + 1] = row_base;
        if (cut[v] < cut[v + 1]) {
            row_out[2 * v] = row_base + row_of(cut[v] / n_prolif);
            row_out[2 * v + 1] = row_base + row_of((cut[v + 1] - 1) / n_prolif);
        }
    }
}

// Validates the sets, sizes the launch (resident wavefronts of the plan's
// device, `wave_factor` times more, so that the dispatcher evens out the
// cost estimate's errors between sets) and uploads the tables.
int core_build(RelaxCore &c, const procell_params *ps, uint32_t n, double beta, bool want_bins,
               uint32_t wave_factor)
{
    if (!ps || n == 0)
        return PROCELL_E_INVALID;
    if (!(beta > 0.0) || !isfinite(beta))
        return PROCELL_E_INVALID;
    for (uint32_t i = 1; i < n; ++i)
        if (ps[i].device != ps[0].device)
            return PROCELL_E_INVALID;
    c.n_sets = n;
    c.setups.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const int rc = build_setup(&ps[i], c.setups[i], want_bins);
        if (rc != PROCELL_OK)
            return rc;
    }
    int n_cu = 0;
    {
        const int rc = relaxed_device(ps[0].device, &n_cu);
        if (rc != PROCELL_OK)
            return rc;
    }
    c.device = ps[0].device;

    // ---- concatenation ----
    std::vector<uint32_t> prolif, quiet, row_start, row_info, row_set;
    c.sets.resize(n);
    uint64_t cells = 0;
    uint32_t rows = 0, kdim_max = 1, t_max_types = 1;
    std::vector<uint64_t> units(n);
    for (uint32_t i = 0; i < n; ++i) {
        const Setup &s = c.setups[i];
        RelaxSet &S = c.sets[i];
        memset(&S, 0, sizeof S);
        S.seed = ps[i].seed;
        S.t_max = ps[i].t_max;
        S.cell_base = cells;
        S.h_base = c.h_total;
        S.row_base = rows;
        S.n_rows = (uint32_t)s.v.size();
        S.kdim = (uint32_t)s.kmax + 1;
        S.n_types = std::max(1u, ps[i].n_types);
        S.theta_base = 3 * c.types_total;
        S.prolif_base = (uint32_t)prolif.size();
        S.quiet_base = (uint32_t)quiet.size();
        for (uint32_t j = 0; j < ps[i].n_types; ++j) {
            const bool q = ps[i].types[j].mean < 0.0;
            (q ? quiet : prolif).push_back(j);
            c.quiet_flag.push_back(q ? 1u : 0u);
        }
        for (uint32_t j = ps[i].n_types; j < S.n_types; ++j)
            c.quiet_flag.push_back(1u);
        S.n_prolif = (uint32_t)prolif.size() - S.prolif_base;
        S.n_quiet = (uint32_t)quiet.size() - S.quiet_base;
        c.any_prolif = c.any_prolif || (S.n_prolif && s.n_cells);
        c.any_quiet = c.any_quiet || S.n_quiet;
        for (uint32_t r = 0; r < S.n_rows; ++r) {
            row_start.push_back((uint32_t)(cells + s.start[r]));
            row_info.push_back(s.row_info[r]);
            row_set.push_back(i);
        }
        units[i] = s.n_cells * (uint64_t)S.n_prolif;
        cells += s.n_cells;
        if (cells > kMaxCells)
            return PROCELL_E_TOO_MANY_CELLS;
        rows += S.n_rows;
        c.h_total += (uint64_t)S.n_rows * S.kdim * S.n_types;
        c.types_total += S.n_types;
        kdim_max = std::max(kdim_max, S.kdim);
        t_max_types = std::max(t_max_types, S.n_types);
    }
    row_start.push_back((uint32_t)cells);

    // ---- wavefronts ----
    c.lds_bytes = relaxed_lds_bytes(t_max_types, 0);
    c.lds_adj = relaxed_lds_bytes(t_max_types, 1);
    c.lds_jac = relaxed_lds_bytes(t_max_types, 2);
    int occ = 0, occ_f = 0;
    int occ_s = 0, occ_sf = 0, occ_j = 0, occ_js = 0;
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, relaxed_kernel<true>, 64, c.lds_adj));
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, relaxed_kernel<false>, 64, c.lds_bytes));
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, relaxed_strided_kernel<true>, 64, c.lds_bytes));
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_sf, relaxed_strided_kernel<false>, 64, c.lds_bytes));
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_j, relaxed_jacobian_kernel<false>, 64, c.lds_jac));
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_js, relaxed_jacobian_kernel<true>, 64, c.lds_bytes));
    occ = std::min(std::min(std::min(occ, occ_f), std::min(occ_s, occ_sf)), std::min(occ_j, occ_js));
    if (occ < 1)
        return PROCELL_E_HIP;
    if (ps[0].tuning.blocks_per_cu)
        occ = std::min(occ, (int)ps[0].tuning.blocks_per_cu);
    const uint64_t budget = (uint64_t)occ * (uint64_t)n_cu * std::max(1u, wave_factor);
    std::vector<std::vector<double>> ucost(n);
    std::vector<double> set_cost(n, 0.0);
    double all_cost = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const Setup &s = c.setups[i];
        const RelaxSet &S = c.sets[i];
        ucost[i].assign(S.n_rows, 1.0);
        if (!units[i])
            continue;
        for (uint32_t r = 0; r < S.n_rows; ++r) {
            double cr = 0.0;
            for (uint32_t q = 0; q < S.n_prolif; ++q)
                cr += walk_cost(s.row_info[r] & 255u, ps[i].t_max, ps[i].types[prolif[S.prolif_base + q]].mean);
            ucost[i][r] = cr / S.n_prolif;
            set_cost[i] += ucost[i][r] * (double)(s.start[r + 1] - s.start[r]) * S.n_prolif;
        }
        all_cost += set_cost[i];
    }
    std::vector<uint32_t> waves_of(n, 0);
    uint64_t n_waves = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!units[i])
            continue;
        double share = all_cost > 0.0 ? set_cost[i] / all_cost * (double)budget : 1.0;
        uint64_t wv = (uint64_t)std::max(1.0, floor(share + 0.5));
        wv = std::min<uint64_t>(wv, (units[i] + 63) / 64);
        waves_of[i] = (uint32_t)std::max<uint64_t>(1, wv);
        n_waves += waves_of[i];
    }
    if (n_waves > 0x7FFFFFFFull)
        return PROCELL_E_TOO_MANY_CELLS;
    c.n_waves = (uint32_t)n_waves;
    const uint32_t NW = std::max(1u, c.n_waves);
    std::vector<uint32_t> wave_set(NW, 0), wave_row(2 * NW, 0);
    std::vector<uint64_t> wave_unit(2 * NW, 0);
    // Two layouts of the same number of wavefronts.  Passes with a cell
    // stride and the Jacobian pass take ranges of equal cost, set by set;
    // the full forward and adjoint passes take ranges that shrink
    // geometrically over rounds (the last 1/8 of the first), launched round
    // by round over all sets, so that the wavefronts dispatched last are
    // short.  256 sets x 2e5 cells: calibration step 23.26 -> 22.48 ms,
    // forward 10.50 -> 10.28, adjoint 13.17 -> 13.03 (the strided passes
    // measured 12 % slower and the Jacobian pass 3 % slower with it: their
    // per-wavefront set-up weighs more in short ranges).
    std::vector<uint32_t> eq_set(NW, 0), eq_row(2 * NW, 0);
    std::vector<uint64_t> eq_unit(2 * NW, 0);
    {
        uint32_t wpos = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!waves_of[i])
                continue;
            split_set(c.setups[i], ucost[i], c.sets[i].n_prolif, waves_of[i], c.sets[i].row_base,
                      &eq_unit[2 * wpos], &eq_row[2 * wpos]);
            for (uint32_t v = 0; v < waves_of[i]; ++v)
                eq_set[wpos + v] = i;
            wpos += waves_of[i];
        }
    }
    struct Piece {
        uint32_t round, set;
        uint64_t u0, u1;
        uint32_t r0, r1;
    };
    std::vector<Piece> pieces;
    pieces.reserve(NW);
    const uint32_t R = std::max(1u, wave_factor);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t nw = waves_of[i];
        if (!nw)
            continue;
        const uint32_t rounds = std::min(nw, R);
        const double q = rounds > 1 ? pow(0.125, 1.0 / (double)(rounds - 1)) : 1.0;
        std::vector<double> wk(nw), cumw(nw);
        double tot = 0.0;
        for (uint32_t k = 0; k < nw; ++k) {
            wk[k] = pow(q, (double)((uint64_t)k * rounds / nw));
            tot += wk[k];
        }
        double acc = 0.0;
        for (uint32_t k = 0; k < nw; ++k) {
            acc += wk[k];
            cumw[k] = k + 1 == nw ? 1.0 : acc / tot;
        }
        std::vector<uint64_t> wu(2 * nw);
        std::vector<uint32_t> wr(2 * nw);
        split_set(c.setups[i], ucost[i], c.sets[i].n_prolif, nw, c.sets[i].row_base, wu.data(), wr.data(), cumw.data());
        for (uint32_t k = 0; k < nw; ++k)
            pieces.push_back(Piece{(uint32_t)((uint64_t)k * rounds / nw), i, wu[2 * k], wu[2 * k + 1], wr[2 * k], wr[2 * k + 1]});
    }
    std::stable_sort(pieces.begin(), pieces.end(), [](const Piece &x, const Piece &y) { return x.round < y.round; });
    for (uint32_t v = 0; v < (uint32_t)pieces.size(); ++v) {
        wave_set[v] = pieces[v].set;
        wave_unit[2 * v] = pieces[v].u0;
        wave_unit[2 * v + 1] = pieces[v].u1;
        wave_row[2 * v] = pieces[v].r0;
        wave_row[2 * v + 1] = pieces[v].r1;
    }

    // ---- upload ----
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_sets = take(c.sets.size() * sizeof(RelaxSet));
    const size_t o_ws = take(wave_set.size() * 4), o_wu = take(wave_unit.size() * 8);
    const size_t o_wr = take(wave_row.size() * 4), o_rs = take(row_start.size() * 4);
    const size_t o_es = take(eq_set.size() * 4), o_eu = take(eq_unit.size() * 8), o_er = take(eq_row.size() * 4);
    const size_t o_ri = take(row_info.size() * 4 + 4), o_rt = take(row_set.size() * 4 + 4);
    const size_t o_pr = take(prolif.size() * 4 + 4), o_qu = take(quiet.size() * 4 + 4);
    const size_t o_ex = take(16 + (size_t)n * 8);  // expansions, walks cut, walks cut per set
    if (hipMalloc(&c.d_block, off) != hipSuccess)
        return PROCELL_E_OUT_OF_MEMORY;
    char *base = (char *)c.d_block;
    auto up = [&](size_t o, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    RELAX_TRY(up(o_sets, c.sets.data(), c.sets.size() * sizeof(RelaxSet)));
    RELAX_TRY(up(o_ws, wave_set.data(), wave_set.size() * 4));
    RELAX_TRY(up(o_wu, wave_unit.data(), wave_unit.size() * 8));
    RELAX_TRY(up(o_wr, wave_row.data(), wave_row.size() * 4));
    RELAX_TRY(up(o_es, eq_set.data(), eq_set.size() * 4));
    RELAX_TRY(up(o_eu, eq_unit.data(), eq_unit.size() * 8));
    RELAX_TRY(up(o_er, eq_row.data(), eq_row.size() * 4));
    RELAX_TRY(up(o_rs, row_start.data(), row_start.size() * 4));
    RELAX_TRY(up(o_ri, row_info.data(), row_info.size() * 4));
    RELAX_TRY(up(o_rt, row_set.data(), row_set.size() * 4));
    RELAX_TRY(up(o_pr, prolif.data(), prolif.size() * 4));
    RELAX_TRY(up(o_qu, quiet.data(), quiet.size() * 4));
    RelaxArgs &a = c.args;
    a.beta = beta;
    a.n_sets = n;
    a.n_rows = rows;
    a.kdim_max = kdim_max;
    a.sets = (const RelaxSet *)(base + o_sets);
    a.wave_set = (const uint32_t *)(base + o_ws);
    a.wave_unit = (const uint64_t *)(base + o_wu);
    a.wave_row = (const uint32_t *)(base + o_wr);
    c.eq_set = (const uint32_t *)(base + o_es);
    c.eq_unit = (const uint64_t *)(base + o_eu);
    c.eq_row = (const uint32_t *)(base + o_er);
    a.row_start = (const uint32_t *)(base + o_rs);
    a.row_info = (const uint32_t *)(base + o_ri);
    a.row_set = (const uint32_t *)(base + o_rt);
    a.prolif = (const uint32_t *)(base + o_pr);
    a.quiet = (const uint32_t *)(base + o_qu);
    a.expansions = (unsigned long long *)(base + o_ex);
    a.set_cuts = a.expansions + 2;
    a.h_total = c.h_total;
    if (c.n_waves) {
        const size_t stack_bytes = (size_t)c.n_waves * 64 * kdim_max * kFrameDoubles * sizeof(double);
        if (hipMalloc((void **)&c.d_stack, stack_bytes) != hipSuccess)
            return PROCELL_E_OUT_OF_MEMORY;
        a.stack = c.d_stack;
    }
    return PROCELL_OK;
}

// Enqueues one pass of a WalkMode: the walk kernel and the quiescent kernel
// (the caller has cleared h, the Jacobian or the gradient, and the expansion
// counter).
hipError_t core_launch(const RelaxCore &c, const RelaxArgs &a_in, int mode, hipStream_t st)
{
    const bool grad = mode == kAdjoint;
    RelaxArgs a = a_in;
    if (a.stride_mask || mode == kJacobian) {
        a.wave_set = c.eq_set;
        a.wave_unit = c.eq_unit;
        a.wave_row = c.eq_row;
    }
    if (c.any_prolif && c.n_waves) {
        const dim3 g(c.n_waves), b(64);
        if (mode == kJacobian) {
            if (a.stride_mask)
                hipLaunchKernelGGL(relaxed_jacobian_kernel<true>, g, b, c.lds_bytes, st, a);
            else
                hipLaunchKernelGGL(relaxed_jacobian_kernel<false>, g, b, c.lds_jac, st, a);
        } else if (a.stride_mask) {
            if (grad)
                hipLaunchKernelGGL(relaxed_strided_kernel<true>, g, b, c.lds_bytes, st, a);
            else
                hipLaunchKernelGGL(relaxed_strided_kernel<false>, g, b, c.lds_bytes, st, a);
        } else if (grad) {
            hipLaunchKernelGGL(relaxed_kernel<true>, g, b, c.lds_adj, st, a);
        } else {
            hipLaunchKernelGGL(relaxed_kernel<false>, g, b, c.lds_bytes, st, a);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    if (c.any_quiet && a.n_rows) {
        const dim3 g((a.n_rows + 255) / 256);
        if (grad)
            hipLaunchKernelGGL(relaxed_quiescent_kernel<true>, g, dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(relaxed_quiescent_kernel<false>, g, dim3(256), 0, st, a);
        return hipGetLastError();
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Batched calibration: per set, the squared L2 distance between the
// normalised relaxed histogram (h summed over types, (row, generation)
// entries merged into the output bins value / 2^k) and a normalised target,
// its cotangent dL/dh, and an Adam step on (log mean, log sd, proportion
// logits).
// ---------------------------------------------------------------------------
struct CalibSet {
    uint32_t bin_base;   // first output bin of the set (targets, per-bin scratch)
    uint32_t n_bins;
    uint32_t ptr_base;   // first entry of the set's bin -> entries CSR offsets (n_bins + 1 of them)
    uint32_t ent_base;   // first (row, generation) entry of the set in ent_bin
    uint32_t n_ent;      // rows x kdim
    uint32_t pad[3];
};

struct CalibArgs {
    uint32_t n_sets;
    uint32_t pad;
    double lr, beta1, beta2, eps;
    const RelaxSet *sets;
    const CalibSet *csets;
    const uint32_t *bin_ptr;   // CSR offsets into bin_ent, per set
    const uint32_t *bin_ent;   // set-local entry ids (row * kdim + k) of each bin
    const int32_t *ent_bin;    // bin of each (row, generation) entry, or -1
    const double *q;           // normalised targets per bin
    const uint32_t *quiet_flag;  // per type of every set
    double *hb;                // per-bin scratch: relaxed histogram, then dL/dhb
    double *loss;              // [n_sets]
    const double *h;
    double *g;                 // dL/dh, written for the adjoint pass (null with jac)
    double *grad;              // dL/dtheta per set [3][T]
    const double *jac;         // dh/dmean | dh/dsd, each laid out like h; null: adjoint route
    uint64_t h_total;
    double *x;                 // per set [3][T]: log mean, log sd, logits
    double *m1, *m2;           // Adam moments, layout of x
    double *theta;             // per set [3][T]
    unsigned long long *step;  // [1] Adam steps taken
};

__device__ double block_sum_256(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    const uint32_t wv = threadIdx.x >> 6, ln = threadIdx.x & 63u;
    __syncthreads();
    if (ln == 0)
        red[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (uint32_t i = 0; i < (blockDim.x >> 6); ++i)
        s += red[i];
    return s;
}

// One workgroup (256 threads) per set.  L = sum_b (hb_b / S - q_b)^2 with
// S = sum_b hb_b; dL/dhb_a = (2 / S) (r_a - sum_b r_b hb_b / S), r = hb / S - q.
__global__ __launch_bounds__(256) void calib_loss_kernel(CalibArgs c)
{
    __shared__ double red[4];
    const uint32_t set = blockIdx.x;
    const RelaxSet &S = c.sets[set];
    const CalibSet &C = c.csets[set];
    const uint32_t T = S.n_types;
    const double *h = c.h + S.h_base;
    double *hb = c.hb + C.bin_base;
    const uint32_t *ptr = c.bin_ptr + C.ptr_base;
    double part = 0.0;
    for (uint32_t b = threadIdx.x; b < C.n_bins; b += blockDim.x) {
        double v = 0.0;
        for (uint32_t e = ptr[b]; e < ptr[b + 1]; ++e) {
            const uint64_t at = (uint64_t)c.bin_ent[e] * T;
            for (uint32_t j = 0; j < T; ++j)
                v += h[at + j];
        }
        hb[b] = v;
        part += v;
    }
    const double tot = block_sum_256(part, red);
    const double inv = tot > 0.0 ? 1.0 / tot : 0.0;
    double pl = 0.0, pc = 0.0;
    for (uint32_t b = threadIdx.x; b < C.n_bins; b += blockDim.x) {
        const double p = hb[b] * inv;
        const double r = p - c.q[C.bin_base + b];
        pl += r * r;
        pc += r * p;
        hb[b] = r;
    }
    const double L = block_sum_256(pl, red);
    const double Cc = block_sum_256(pc, red);
    for (uint32_t b = threadIdx.x; b < C.n_bins; b += blockDim.x)
        hb[b] = tot > 0.0 ? 2.0 * inv * (hb[b] - Cc) : 0.0;
    __syncthreads();
    if (threadIdx.x == 0)
        c.loss[set] = L;
    if (!c.jac) {
        double *g = c.g + S.h_base;
        for (uint32_t e = threadIdx.x; e < C.n_ent; e += blockDim.x) {
            const int32_t b = c.ent_bin[C.ent_base + e];
            const double gv = b >= 0 ? hb[b] : 0.0;
            for (uint32_t j = 0; j < T; ++j)
                g[(uint64_t)e * T + j] = gv;
        }
        return;
    }
    // Jacobian route: dL/dtheta = J^T dL/dh.  An entry depends on its own
    // type's parameters only, and h is linear in p_j (dh/dp_j = h / p_j; at
    // p_j = 0 the entry is left 0: the Adam step multiplies it by p_j).
    const double *jm = c.jac + S.h_base, *js = c.jac + c.h_total + S.h_base;
    const double *th = c.theta + S.theta_base;
    double *gr = c.grad + S.theta_base;
    for (uint32_t j = 0; j < T; ++j) {
        double sm = 0.0, ss = 0.0, sp = 0.0;
        for (uint32_t e = threadIdx.x; e < C.n_ent; e += blockDim.x) {
            const int32_t b = c.ent_bin[C.ent_base + e];
            if (b < 0)
                continue;
            const double gv = hb[b];
            const uint64_t at = (uint64_t)e * T + j;
            sm += gv * jm[at];
            ss += gv * js[at];
            sp += gv * h[at];
        }
        sm = block_sum_256(sm, red);
        ss = block_sum_256(ss, red);
        sp = block_sum_256(sp, red);
        if (threadIdx.x == 0) {
            const double pj = th[2 * T + j];
            gr[j] = sm;
            gr[T + j] = ss;
            gr[2 * T + j] = pj > 0.0 ? sp / pj : 0.0;
        }
    }
}

// theta of one set from its unconstrained parameters.
__device__ void calib_theta(const CalibArgs &c, const RelaxSet &S)
{
    const uint32_t T = S.n_types, tb = S.theta_base;
    const uint32_t *qf = c.quiet_flag + tb / 3;
    double mx = c.x[tb + 2 * T];
    for (uint32_t j = 1; j < T; ++j)
        mx = fmax(mx, c.x[tb + 2 * T + j]);
    double sum = 0.0;
    for (uint32_t j = 0; j < T; ++j)
        sum += exp(c.x[tb + 2 * T + j] - mx);
    for (uint32_t j = 0; j < T; ++j) {
        c.theta[tb + j] = qf[j] ? -1.0 : exp(c.x[tb + j]);
        c.theta[tb + T + j] = qf[j] ? -1.0 : exp(c.x[tb + T + j]);
        c.theta[tb + 2 * T + j] = exp(c.x[tb + 2 * T + j] - mx) / sum;
    }
}

__device__ __forceinline__ void adam_update(double &x, double &m, double &v, double gr, const CalibArgs &c,
                                            double bc1, double bc2)
{
    m = c.beta1 * m + (1.0 - c.beta1) * gr;
    v = c.beta2 * v + (1.0 - c.beta2) * gr * gr;
    const double denom = sqrt(v) / sqrt(bc2) + c.eps;
    x = x - (c.lr / bc1) * (m / denom);
}

// One workgroup; thread s updates set s (and s + blockDim, ...).  Chain rule
// to the unconstrained parameters: d/dlog mean = dL/dmean mean, d/dlog sd =
// dL/dsd sd, d/dlogit_j = p_j (dL/dp_j - sum_k p_k dL/dp_k).
__global__ __launch_bounds__(256) void calib_adam_kernel(CalibArgs c)
{
    const unsigned long long t = c.step[0] + 1ull;
    const double bc1 = 1.0 - pow(c.beta1, (double)t), bc2 = 1.0 - pow(c.beta2, (double)t);
    for (uint32_t set = threadIdx.x; set < c.n_sets; set += blockDim.x) {
        const RelaxSet &S = c.sets[set];
        const uint32_t T = S.n_types, tb = S.theta_base;
        const uint32_t *qf = c.quiet_flag + tb / 3;
        const double *th = c.theta + tb;
        const double *gr = c.grad + tb;
        double dot = 0.0;
        for (uint32_t j = 0; j < T; ++j)
            dot += th[2 * T + j] * gr[2 * T + j];
        for (uint32_t j = 0; j < T; ++j) {
            if (!qf[j]) {
                adam_update(c.x[tb + j], c.m1[tb + j], c.m2[tb + j], gr[j] * th[j], c, bc1, bc2);
                adam_update(c.x[tb + T + j], c.m1[tb + T + j], c.m2[tb + T + j], gr[T + j] * th[T + j], c, bc1,
                            bc2);
                // keep sd <= 0.8 mean (kLogMaxSdRatio): the expected tree
                // grows quickly beyond it
                c.x[tb + T + j] = fmin(c.x[tb + T + j], c.x[tb + j] + kLogMaxSdRatio);
            }
            const uint32_t z = tb + 2 * T + j;
            adam_update(c.x[z], c.m1[z], c.m2[z], th[2 * T + j] * (gr[2 * T + j] - dot), c, bc1, bc2);
        }
        calib_theta(c, S);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        c.step[0] = t;
}

__global__ void calib_init_kernel(CalibArgs c)
{
    for (uint32_t set = blockIdx.x * blockDim.x + threadIdx.x; set < c.n_sets; set += gridDim.x * blockDim.x)
        calib_theta(c, c.sets[set]);
}

}  // namespace
}  // namespace procell

using namespace procell;

struct procell_relaxed {
    RelaxCore core;
};

struct procell_calib {
    RelaxCore core;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    CalibArgs cargs{};
    RelaxArgs fwd{}, bwd{};  // fwd is the Jacobian pass without a cell stride
    bool jac_route = false;
    void *d_calib = nullptr;        // tables of the loss kernel
    void *d_state = nullptr;        // x, m1, m2, theta, grad, h, g, per-bin scratch, loss, step
    unsigned long long *d_clear = nullptr;  // h | grad | expansion counter: cleared every step
    size_t clear_words = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace {

void destroy_relaxed(procell_relaxed *r)
{
    if (!r)
        return;
    core_free(r->core);
    delete r;
}

void destroy_calib(procell_calib *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->core.device);
    if (c->graph_exec)
        (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph)
        (void)hipGraphDestroy(c->graph);
    if (c->d_calib)
        (void)hipFree(c->d_calib);
    if (c->d_state)
        (void)hipFree(c->d_state);
    if (c->own_stream && c->stream)
        (void)hipStreamDestroy(c->stream);
    core_free(c->core);
    delete c;
}

int relaxed_call(procell_relaxed *r, int mode, const double *theta, double *h, const double *g, double *gr,
                 double *jac, void *stream)
{
    RELAX_TRY(hipSetDevice(r->core.device));
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    RelaxArgs a = r->core.args;
    a.theta = theta;
    a.h = h;
    a.g = g;
    a.grad = gr;
    a.jac = jac;
    RELAX_TRY(launch_zero(a.expansions, 2ull + a.n_sets, st));
    if (mode == kAdjoint) {
        RELAX_TRY(launch_zero((unsigned long long *)gr, 3ull * r->core.types_total, st));
    } else {
        RELAX_TRY(launch_zero((unsigned long long *)h, r->core.h_total, st));
        if (mode == kJacobian)
            RELAX_TRY(launch_zero((unsigned long long *)jac, 2ull * r->core.h_total, st));
    }
    RELAX_TRY(core_launch(r->core, a, mode, st));
    return PROCELL_OK;
}

// One calibration step.  Without a cell stride: clear, the Jacobian pass (h
// and dh/dtheta in one walk), loss and gradient J^T dL/dh, Adam.  With one:
// clear, forward, loss and cotangent, adjoint on the next class of cells,
// Adam.
hipError_t calib_enqueue(procell_calib *c, hipStream_t st)
{
    hipError_t e = launch_zero(c->d_clear, c->clear_words, st);
    if (e != hipSuccess)
        return e;
    if ((e = core_launch(c->core, c->fwd, c->jac_route ? kJacobian : kForward, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(calib_loss_kernel, dim3(c->core.n_sets), dim3(256), 0, st, c->cargs);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if (!c->jac_route && (e = core_launch(c->core, c->bwd, kAdjoint, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(calib_adam_kernel, dim3(1), dim3(256), 0, st, c->cargs);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int procell_relaxed_create(const procell_params *p, double beta, procell_relaxed **out)
{
    if (!p || !out)
        return PROCELL_E_INVALID;
    *out = nullptr;
    if (!(beta > 0.0) || !isfinite(beta))
        return PROCELL_E_INVALID;
    procell_relaxed *r = new (std::nothrow) procell_relaxed();
    if (!r)
        return PROCELL_E_OUT_OF_MEMORY;
    const int rc = core_build(r->core, p, 1, beta, false, 1);
    if (rc != PROCELL_OK) {
        destroy_relaxed(r);
        return rc;
    }
    *out = r;
    return PROCELL_OK;
}

int procell_relaxed_create_batch(const procell_params *sets, uint32_t n_sets, double beta, procell_relaxed **out)
{
    if (!sets || !out || n_sets == 0)
        return PROCELL_E_INVALID;
    *out = nullptr;
    procell_relaxed *r = new (std::nothrow) procell_relaxed();
    if (!r)
        return PROCELL_E_OUT_OF_MEMORY;
    const int rc = core_build(r->core, sets, n_sets, beta, false, n_sets > 1 ? kMultiSetWaveFactor : 1u);
    if (rc != PROCELL_OK) {
        destroy_relaxed(r);
        return rc;
    }
    *out = r;
    return PROCELL_OK;
}

int procell_relaxed_set_shape(const procell_relaxed *r, uint32_t set, uint64_t *h_base, uint64_t *n_rows,
                              uint32_t *kdim, uint32_t *n_types)
{
    if (!r || set >= r->core.n_sets)
        return PROCELL_E_INVALID;
    const RelaxSet &S = r->core.sets[set];
    if (h_base)
        *h_base = S.h_base;
    if (n_rows)
        *n_rows = S.n_rows;
    if (kdim)
        *kdim = S.kdim;
    if (n_types)
        *n_types = S.n_types;
    return PROCELL_OK;
}

int procell_relaxed_shape(const procell_relaxed *r, uint64_t *n_rows, uint32_t *kdim, uint32_t *n_types)
{
    if (!r)
        return PROCELL_E_INVALID;
    const RelaxSet &S = r->core.sets[0];
    if (n_rows)
        *n_rows = S.n_rows;
    if (kdim)
        *kdim = S.kdim;
    if (n_types)
        *n_types = S.n_types;
    return PROCELL_OK;
}

int procell_relaxed_forward(procell_relaxed *r, const double *theta, double *h, void *stream)
{
    if (!r || !theta || !h)
        return PROCELL_E_INVALID;
    return relaxed_call(r, kForward, theta, h, nullptr, nullptr, nullptr, stream);
}

int procell_relaxed_backward(procell_relaxed *r, const double *theta, const double *g, double *grad,
                             void *stream)
{
    if (!r || !theta || !g || !grad)
        return PROCELL_E_INVALID;
    return relaxed_call(r, kAdjoint, theta, nullptr, g, grad, nullptr, stream);
}

int procell_relaxed_jacobian(procell_relaxed *r, const double *theta, double *h, double *jac, void *stream)
{
    if (!r || !theta || !h || !jac)
        return PROCELL_E_INVALID;
    return relaxed_call(r, kJacobian, theta, h, nullptr, nullptr, jac, stream);
}

int procell_relaxed_set_stride(procell_relaxed *r, uint32_t stride, uint32_t phase)
{
    if (!r || stride == 0 || (stride & (stride - 1)) != 0 || stride > (1u << 16))
        return PROCELL_E_INVALID;
    r->core.args.stride_mask = stride - 1;
    r->core.args.phase = phase & (stride - 1);
    return PROCELL_OK;
}

int procell_relaxed_expansions(procell_relaxed *r, uint64_t *n)
{
    if (!r || !n)
        return PROCELL_E_INVALID;
    RELAX_TRY(hipSetDevice(r->core.device));
    RELAX_TRY(hipDeviceSynchronize());
    unsigned long long v = 0;
    unsigned long long v2[2] = {0, 0};
    RELAX_TRY(hipMemcpy(v2, r->core.args.expansions, sizeof v2, hipMemcpyDeviceToHost));
    v = v2[0];
    *n = v;
    return v2[1] ? PROCELL_E_DEPTH : PROCELL_OK;
}

int procell_relaxed_set_cuts(procell_relaxed *r, uint64_t *cuts, void *stream)
{
    if (!r || !cuts)
        return PROCELL_E_INVALID;
    RELAX_TRY(hipSetDevice(r->core.device));
    RELAX_TRY(hipMemcpyAsync(cuts, r->core.args.set_cuts, (size_t)r->core.args.n_sets * 8, hipMemcpyDeviceToDevice,
                             (hipStream_t)stream));
    return PROCELL_OK;
}

int procell_relaxed_destroy(procell_relaxed *r)
{
    destroy_relaxed(r);
    return PROCELL_OK;
}

int procell_calib_create(const procell_params *sets, uint32_t n_sets, double beta, const procell_result *targets,
                         const procell_adam *adam, uint32_t cell_stride, procell_calib **out)
{
    if (!sets || !targets || !adam || !out || n_sets == 0)
        return PROCELL_E_INVALID;
    if (cell_stride == 0 || (cell_stride & (cell_stride - 1)) != 0 || cell_stride > (1u << 16))
        return PROCELL_E_INVALID;
    *out = nullptr;
    if (!(adam->lr > 0.0) || !(adam->beta1 >= 0.0 && adam->beta1 < 1.0) || !(adam->beta2 >= 0.0 && adam->beta2 < 1.0) ||
        !(adam->eps > 0.0))
        return PROCELL_E_INVALID;
    procell_calib *c = new (std::nothrow) procell_calib();
    if (!c)
        return PROCELL_E_OUT_OF_MEMORY;
    auto fail = [&](int code) {
        destroy_calib(c);
        return code;
    };
#define CALIB_TRY(expr)                       \
    do {                                      \
        hipError_t e_ = (expr);               \
        if (e_ != hipSuccess) {               \
            g_relaxed_hip = e_;               \
            return fail(PROCELL_E_HIP);       \
        }                                     \
    } while (0)
    int rc = core_build(c->core, sets, n_sets, beta, true, n_sets > 1 ? kMultiSetWaveFactor : 1u);
    if (rc != PROCELL_OK)
        return fail(rc);
    RelaxCore &core = c->core;
    for (uint32_t i = 0; i < n_sets; ++i)
        for (uint32_t j = 0; j < sets[i].n_types; ++j)
            if (sets[i].types[j].mean >= 0.0 && !(sets[i].types[j].mean > 0.0 && sets[i].types[j].sd > 0.0 &&
                                                  sets[i].types[j].proportion > 0.0))
                return fail(PROCELL_E_INVALID);  // log parameters need mean, sd, proportion > 0

    // ---- bins, targets, (row, generation) -> bin maps ----
    std::vector<CalibSet> cs(n_sets);
    std::vector<uint32_t> bin_ptr, bin_ent;
    std::vector<int32_t> ent_bin;
    std::vector<double> q;
    for (uint32_t i = 0; i < n_sets; ++i) {
        const Setup &s = core.setups[i];
        const RelaxSet &S = core.sets[i];
        CalibSet &C = cs[i];
        memset(&C, 0, sizeof C);
        C.bin_base = (uint32_t)q.size();
        C.n_bins = (uint32_t)s.bins.size();
        C.ptr_base = (uint32_t)bin_ptr.size();
        C.ent_base = (uint32_t)ent_bin.size();
        C.n_ent = S.n_rows * S.kdim;
        const procell_result &tr = targets[i];
        if (tr.n_out && (!tr.value || !tr.count))
            return fail(PROCELL_E_INVALID);
        std::vector<double> tq(C.n_bins, 0.0);
        double tsum = 0.0;
        for (uint64_t r = 0; r < tr.n_out; ++r) {
            auto it = std::lower_bound(s.bins.begin(), s.bins.end(), tr.value[r]);
            if (it == s.bins.end() || *it != tr.value[r] || tr.count[r] < 0)
                return fail(PROCELL_E_INVALID);  // a target value outside the set's output bins
            tq[it - s.bins.begin()] += (double)tr.count[r];
            tsum += (double)tr.count[r];
        }
        for (double &x : tq)
            q.push_back(tsum > 0.0 ? x / tsum : 0.0);
        std::vector<std::vector<uint32_t>> members(C.n_bins);
        for (uint32_t r = 0; r < S.n_rows; ++r) {
            const uint32_t kc = s.row_info[r] & 255u;
            const bool ge = ((s.row_info[r] >> 8) & 1u) != 0;
            double f = s.v[r];
            for (uint32_t k = 0; k < S.kdim; ++k) {
                int32_t b = -1;
                if (k <= kc && (k > 0 || ge)) {
                    auto it = std::lower_bound(s.bins.begin(), s.bins.end(), f);
                    if (it != s.bins.end() && *it == f) {
                        b = (int32_t)(it - s.bins.begin());
                        members[b].push_back(r * S.kdim + k);
                    }
                }
                ent_bin.push_back(b);
                f = f / 2;
            }
        }
        for (uint32_t b = 0; b < C.n_bins; ++b) {
            bin_ptr.push_back((uint32_t)bin_ent.size());
            for (uint32_t e : members[b])
                bin_ent.push_back(e);
        }
        bin_ptr.push_back((uint32_t)bin_ent.size());
    }
    // CSR offsets are relative to the whole bin_ent array
    const size_t n_bins_total = q.size();

    // ---- device memory ----
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_cs = take(cs.size() * sizeof(CalibSet)), o_bp = take(bin_ptr.size() * 4 + 4);
    const size_t o_be = take(bin_ent.size() * 4 + 4), o_eb = take(ent_bin.size() * 4 + 4);
    const size_t o_q = take(q.size() * 8 + 8), o_qf = take(core.quiet_flag.size() * 4 + 4);
    if (hipMalloc(&c->d_calib, off) != hipSuccess)
        return fail(PROCELL_E_OUT_OF_MEMORY);
    char *base = (char *)c->d_calib;
    auto up = [&](size_t o, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    CALIB_TRY(up(o_cs, cs.data(), cs.size() * sizeof(CalibSet)));
    CALIB_TRY(up(o_bp, bin_ptr.data(), bin_ptr.size() * 4));
    CALIB_TRY(up(o_be, bin_ent.data(), bin_ent.size() * 4));
    CALIB_TRY(up(o_eb, ent_bin.data(), ent_bin.size() * 4));
    CALIB_TRY(up(o_q, q.data(), q.size() * 8));
    CALIB_TRY(up(o_qf, core.quiet_flag.data(), core.quiet_flag.size() * 4));

    const size_t P = 3ull * core.types_total;
    size_t so = 0;
    auto stake = [&](size_t bytes) {
        const size_t o = so;
        so = (so + bytes + 255) / 256 * 256;
        return o;
    };
    // cleared every step: h | Jacobian (without a stride) | grad | expansions
    // (one contiguous block)
    c->jac_route = cell_stride == 1;
    const size_t jac_words = c->jac_route ? 2ull * core.h_total : 0;
    const size_t s_h = stake(0);
    const size_t clear_bytes = (core.h_total + jac_words) * 8 + P * 8 + 16;
    so = (clear_bytes + 255) / 256 * 256;
    const size_t s_x = stake(P * 8), s_m1 = stake(P * 8), s_m2 = stake(P * 8), s_th = stake(P * 8);
    const size_t s_g = stake(c->jac_route ? 8 : core.h_total * 8 + 8), s_hb = stake(n_bins_total * 8 + 8);
    const size_t s_loss = stake(n_sets * 8), s_step = stake(8);
    if (hipMalloc(&c->d_state, so) != hipSuccess)
        return fail(PROCELL_E_OUT_OF_MEMORY);
    char *sb = (char *)c->d_state;
    std::vector<double> x(P, 0.0);
    for (uint32_t i = 0; i < n_sets; ++i) {
        const RelaxSet &S = core.sets[i];
        const uint32_t T = S.n_types;
        for (uint32_t j = 0; j < sets[i].n_types; ++j) {
            const procell_type &ty = sets[i].types[j];
            if (ty.mean >= 0.0) {
                x[S.theta_base + j] = log(ty.mean);
                x[S.theta_base + T + j] = log(ty.sd);
            }
            x[S.theta_base + 2 * T + j] = ty.proportion > 0.0 ? log(ty.proportion) : -745.0;
        }
        for (uint32_t j = sets[i].n_types; j < T; ++j)
            x[S.theta_base + 2 * T + j] = -745.0;  // padding type of an empty type table
    }
    CALIB_TRY(hipMemcpy(sb + s_x, x.data(), P * 8, hipMemcpyHostToDevice));
    CALIB_TRY(hipMemset(sb + s_m1, 0, P * 8));
    CALIB_TRY(hipMemset(sb + s_m2, 0, P * 8));
    CALIB_TRY(hipMemset(sb + s_step, 0, 8));
    CALIB_TRY(hipMemset(sb, 0, clear_bytes));
    c->d_clear = (unsigned long long *)(sb + s_h);
    c->clear_words = clear_bytes / 8;
    double *d_h = (double *)(sb + s_h);
    double *d_jac = c->jac_route ? d_h + core.h_total : nullptr;
    double *d_grad = d_h + core.h_total + jac_words;
    unsigned long long *d_exp = (unsigned long long *)(d_grad + P);

    CalibArgs &ca = c->cargs;
    ca.n_sets = n_sets;
    ca.lr = adam->lr;
    ca.beta1 = adam->beta1;
    ca.beta2 = adam->beta2;
    ca.eps = adam->eps;
    ca.sets = core.args.sets;
    ca.csets = (const CalibSet *)(base + o_cs);
    ca.bin_ptr = (const uint32_t *)(base + o_bp);
    ca.bin_ent = (const uint32_t *)(base + o_be);
    ca.ent_bin = (const int32_t *)(base + o_eb);
    ca.q = (const double *)(base + o_q);
    ca.quiet_flag = (const uint32_t *)(base + o_qf);
    ca.hb = (double *)(sb + s_hb);
    ca.loss = (double *)(sb + s_loss);
    ca.h = d_h;
    ca.g = (double *)(sb + s_g);
    ca.grad = d_grad;
    ca.jac = d_jac;
    ca.h_total = core.h_total;
    ca.x = (double *)(sb + s_x);
    ca.m1 = (double *)(sb + s_m1);
    ca.m2 = (double *)(sb + s_m2);
    ca.theta = (double *)(sb + s_th);
    ca.step = (unsigned long long *)(sb + s_step);
    c->fwd = core.args;
    c->fwd.step = ca.step;
    c->fwd.stride_mask = cell_stride - 1;
    c->fwd.theta = ca.theta;
    c->fwd.h = d_h;
    c->fwd.expansions = d_exp;
    c->fwd.set_cuts = nullptr;  // a calibration reports the step's total only
    c->fwd.jac = d_jac;
    c->bwd = c->fwd;
    // with a stride, the adjoint walks the next class of cells: the residual
    // (from the forward pass) and the Jacobian then come from independent
    // cells, so their product is an unbiased gradient of the squared loss
    c->bwd.phase = cell_stride > 1 ? 1u : 0u;
    c->bwd.h = nullptr;
    c->bwd.g = ca.g;
    c->bwd.grad = d_grad;

    if (sets[0].stream) {
        c->stream = (hipStream_t)sets[0].stream;
    } else {
        CALIB_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    hipLaunchKernelGGL(calib_init_kernel, dim3((n_sets + 255) / 256), dim3(256), 0, c->stream, ca);
    CALIB_TRY(hipGetLastError());
    CALIB_TRY(hipStreamSynchronize(c->stream));
    *out = c;
    return PROCELL_OK;
#undef CALIB_TRY
}

int procell_calib_step(procell_calib *c, uint32_t n_steps)
{
    if (!c)
        return PROCELL_E_INVALID;
    RELAX_TRY(hipSetDevice(c->core.device));
    if (!c->graph_exec) {
        RELAX_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e = calib_enqueue(c, c->stream);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g)
                (void)hipGraphDestroy(g);
            g_relaxed_hip = e != hipSuccess ? e : e2;
            return PROCELL_E_HIP;
        }
        const hipError_t e3 = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            (void)hipGraphDestroy(g);
            c->graph_exec = nullptr;
            g_relaxed_hip = e3;
            return PROCELL_E_HIP;
        }
        c->graph = g;
    }
    for (uint32_t i = 0; i < n_steps; ++i)
        RELAX_TRY(hipGraphLaunch(c->graph_exec, c->stream));
    return PROCELL_OK;
}

int procell_calib_theta(procell_calib *c, double *theta, uint64_t capacity)
{
    if (!c || !theta || capacity < 3ull * c->core.types_total)
        return PROCELL_E_INVALID;
    RELAX_TRY(hipSetDevice(c->core.device));
    RELAX_TRY(hipStreamSynchronize(c->stream));
    RELAX_TRY(hipMemcpy(theta, c->cargs.theta, 3ull * c->core.types_total * 8, hipMemcpyDeviceToHost));
    return PROCELL_OK;
}

int procell_calib_loss(procell_calib *c, double *loss, uint32_t capacity, uint64_t *steps, uint64_t *expansions)
{
    if (!c || !loss || capacity < c->core.n_sets)
        return PROCELL_E_INVALID;
    RELAX_TRY(hipSetDevice(c->core.device));
    RELAX_TRY(hipStreamSynchronize(c->stream));
    RELAX_TRY(hipMemcpy(loss, c->cargs.loss, c->core.n_sets * 8ull, hipMemcpyDeviceToHost));
    if (steps) {
        unsigned long long s = 0;
        RELAX_TRY(hipMemcpy(&s, c->cargs.step, 8, hipMemcpyDeviceToHost));
        *steps = s;
    }
    unsigned long long ex[2] = {0, 0};
    RELAX_TRY(hipMemcpy(ex, c->fwd.expansions, sizeof ex, hipMemcpyDeviceToHost));
    if (expansions)
        *expansions = ex[0];
    return ex[1] ? PROCELL_E_DEPTH : PROCELL_OK;
}

void *procell_calib_stream(procell_calib *c) { return c ? (void *)c->stream : nullptr; }

int procell_calib_destroy(procell_calib *c)
{
    destroy_calib(c);
    return PROCELL_OK;
}

}  // extern "C"
