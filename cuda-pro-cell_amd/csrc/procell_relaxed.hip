// procell_relaxed.hip -- the relaxed (differentiable) estimator of the plain
// method's expected tally, and its gradient with respect to each type's
// (mean, sd, proportion).  No reference counterpart: the reference
// (src/simulation/proliferation.cu) only samples.
//
// Relaxation of the plain method (proliferation.cu:26-215, cell.cu:80-125):
//   * every initial cell takes every type j with weight p_j (no type draw);
//   * timers are mean_j + sd_j z, untruncated, with z the first candidates of
//     the plain stream's blocks: the root's pair from block (cell, 0, 0)
//     words {2, -, 3, -}, a node's daughters' pair from block (cell, node, 0);
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
// device memory, laid out [wavefront][level][field][lane] so that a
// wavefront's pushes and pops at one level are coalesced.  Recorded weights
// go to an LDS window over the rows of the wavefront's range (the cotangent
// window in the backward pass) and reach the tally with FP64 atomics at the
// end.
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
constexpr uint32_t kFrameFields = 6;       // s, b, W, d, dlogW/dmean, dlogW/dsd
constexpr uint32_t kWindowDoubles = 1024;  // LDS window of a wavefront (8 KB)

struct RelaxArgs {
    uint64_t seed;
    double t_max;
    double beta;
    uint32_t n_rows;
    uint32_t kdim;             // generations + 1
    uint32_t n_types;
    uint32_t n_prolif;         // proliferating types (mean >= 0 at plan creation)
    uint32_t n_quiet;
    uint32_t pad;
    const uint32_t *prolif;    // [n_prolif] file index of each proliferating type
    const uint32_t *quiet;     // [n_quiet] file index of each quiescent type
    const double *theta;       // [3][n_types]: mean, sd, proportion, file order
    const uint32_t *row_start; // [n_rows + 1] first initial cell of each nonzero row
    const uint32_t *row_info;  // [n_rows] generation cap | (value >= phi) << 8
    const uint64_t *wave_unit; // [n_waves + 1] first walk of each wavefront's range
    const uint32_t *wave_row;  // [n_waves][2] first and last row of the range
    double *stack;             // [n_waves][kdim][kFrameFields][64]
    double *h;                 // [n_rows][kdim][n_types] relaxed tally (forward)
    const double *g;           // [n_rows][kdim][n_types] cotangent (backward)
    double *grad;              // [3][n_types] gradient (backward)
    unsigned long long *expansions;  // [1] expansions of the call (roots drawn + nodes expanded)
};

__device__ __forceinline__ double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

template <bool Grad>
__global__ __launch_bounds__(64) void relaxed_kernel(RelaxArgs a)
{
    extern __shared__ double lds[];  // [kWindowDoubles] window | [3][n_types] gradient sums
    const uint32_t lane = threadIdx.x;
    const uint32_t w = blockIdx.x;
    const uint32_t T = a.n_types;
    const uint32_t KT = a.kdim * T;
    uint64_t next = a.wave_unit[w];
    const uint64_t u_end = a.wave_unit[w + 1];
    if (next >= u_end)
        return;
    const uint32_t r_lo = a.wave_row[2 * w], r_hi = a.wave_row[2 * w + 1];
    const uint32_t win_n = (r_hi - r_lo + 1) * KT;
    const bool in_lds = win_n <= kWindowDoubles;
    double *gsum = lds + kWindowDoubles;
    if (in_lds)
        for (uint32_t i = lane; i < win_n; i += 64)
            lds[i] = Grad ? a.g[(uint64_t)r_lo * KT + i] : 0.0;
    if (Grad)
        for (uint32_t i = lane; i < 3 * T; i += 64)
            gsum[i] = 0.0;
    __syncthreads();

    const dev::Coeffs K = dev::coeffs();
    const dev::Tables tb = dev::global_tables();
    const double t_max = a.t_max, beta = a.beta;
    const uint32_t n_prolif = a.n_prolif;
    double *stk = a.stack + (uint64_t)w * a.kdim * kFrameFields * 64 + lane;

    bool active = false;
    uint64_t cell = 0, node = 0, pending = 0;
    uint32_t row = r_lo, j = 0, kcap = 0, k = 0;
    bool ge_phi = false;
    double pj = 0.0, mean = 0.0, sd = 0.0, u1 = 0.0;  // u1 = 1 + U of the walk's cell
    double s = 0.0, b = 0.0, W = 0.0, d = 0.0, gm = 0.0, gs = 0.0;
    double acc_m = 0.0, acc_s = 0.0, acc_p = 0.0;
    uint32_t row_cursor = r_lo;  // row of the first walk not yet taken (wave-uniform)
    uint32_t n_exp = 0;

    for (;;) {
        // ---- refill idle lanes from the wavefront's range ----
        const uint64_t idle = __ballot(!active);
        if (idle && next < u_end) {
            if (!active) {
                const uint64_t u = next + (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (u < u_end) {
                    cell = u / n_prolif;
                    j = a.prolif[u - cell * n_prolif];
                    row = row_cursor;
                    while (a.row_start[row + 1] <= cell)
                        ++row;
                    const uint32_t info = a.row_info[row];
                    kcap = info & 255u;
                    ge_phi = ((info >> 8) & 1u) != 0;
                    mean = a.theta[j];
                    sd = a.theta[T + j];
                    pj = a.theta[2 * T + j];
                    node = 0;  // the next step draws the root from block (cell, 0, 0)
                    pending = 0;
                    active = true;
                }
            }
            next = next + (uint64_t)__popcll(idle);
            if (next > u_end)
                next = u_end;
            if (next < u_end) {
                const uint64_t c = next / n_prolif;
                while (a.row_start[row_cursor + 1] <= c)
                    ++row_cursor;
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
        dev::Words wd = dev::draw_block(a.seed, cell, node, 0u);
        if (pre) {
            u1 = 1.0 + dev::uniform_oc(wd.w1, 0u);
            wd = dev::Words{wd.w2, 0u, wd.w3, 0u};
        }
        double z0, z1;
        dev::normal_pair(wd, z0, z1, K, tb);

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
            const double f = (1.0 - d) / beta;
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
        const double dA = sigmoid((t_max - sA) / beta);
        const double dB = pre ? 0.0 : sigmoid((t_max - sB) / beta);
        const double recA = WA * (1.0 - dA);
        const double recB = WA * (1.0 - dB);
        const bool recordA = kA > 0 || ge_phi;
        const uint32_t col = kA * T + j;
        if (!Grad) {
            // both daughters land in the same (row, generation, type) entry
            const double v = (recordA ? pj * recA : 0.0) + (pre ? 0.0 : pj * recB);
            if (recordA || !pre) {
                if (in_lds)
                    atomicAdd(lds + (row - r_lo) * KT + col, v);
                else
                    unsafeAtomicAdd(a.h + (uint64_t)row * KT + col, v);
            }
        } else {
            const double gv = in_lds ? lds[(row - r_lo) * KT + col] : a.g[(uint64_t)row * KT + col];
            if (recordA) {
                const double q = dA * (1.0 - dA) / beta;
                acc_m += pj * gv * WA * ((1.0 - dA) * gmA + q * aA);
                acc_s += pj * gv * WA * ((1.0 - dA) * gsA + q * bA);
                acc_p += gv * recA;
            }
            if (!pre) {
                const double q = dB * (1.0 - dB) / beta;
                acc_m += pj * gv * WA * ((1.0 - dB) * gmA + q * aA);
                acc_s += pj * gv * WA * ((1.0 - dB) * gsA + q * bB);
                acc_p += gv * recB;
            }
        }
        const bool expA = kA < kcap && WA * dA >= kRelaxEps;
        const bool expB = !pre && kA < kcap && WA * dB >= kRelaxEps;

        // ---- next node: left daughter (right one stacked), the only
        //      expandable daughter, or the deepest stacked node ----
        if (expA && expB) {
            double *f = stk + (uint64_t)kA * kFrameFields * 64;
            f[0 * 64] = sB;
            f[1 * 64] = bB;
            f[2 * 64] = WA;
            f[3 * 64] = dB;
            f[4 * 64] = gmA;
            f[5 * 64] = gsA;
            pending |= 1ull << kA;
        }
        if (expA || expB) {
            const bool left = expA;
            node = pre ? 1ull : 2ull * node + (left ? 0ull : 1ull);
            k = kA;
            s = left ? sA : sB;
            b = left ? bA : bB;
            d = left ? dA : dB;
            W = WA;
            gm = gmA;
            gs = gsA;
        } else if (pending) {
            const uint32_t L = 63u - (uint32_t)__clzll(pending);
            pending &= ~(1ull << L);
            node = (node >> (k - L)) | 1ull;
            k = L;
            const double *f = stk + (uint64_t)L * kFrameFields * 64;
            s = f[0 * 64];
            b = f[1 * 64];
            W = f[2 * 64];
            d = f[3 * 64];
            gm = f[4 * 64];
            gs = f[5 * 64];
        } else {
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
    for (int o = 32; o > 0; o >>= 1)
        n_exp += __shfl_xor(n_exp, o);
    if (lane == 0)
        atomicAdd(a.expansions, (unsigned long long)n_exp);
    if (!Grad) {
        if (in_lds)
            for (uint32_t i = lane; i < win_n; i += 64) {
                const double v = lds[i];
                if (v != 0.0)
                    unsafeAtomicAdd(a.h + (uint64_t)r_lo * KT + i, v);
            }
    } else {
        for (uint32_t i = lane; i < 3 * T; i += 64) {
            const double v = gsum[i];
            if (v != 0.0)
                unsafeAtomicAdd(a.grad + i, v);
        }
    }
}

// Quiescent types: each initial cell of a row with value >= phi records p_j
// at generation 0 (cell.cu:104-108 gives them timer -1, always out of time).
template <bool Grad>
__global__ void relaxed_quiescent_kernel(RelaxArgs a)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)a.n_rows * a.n_quiet)
        return;
    const uint32_t row = (uint32_t)(t / a.n_quiet);
    const uint32_t j = a.quiet[t - (uint64_t)row * a.n_quiet];
    if (((a.row_info[row] >> 8) & 1u) == 0)
        return;
    const double cnt = (double)(a.row_start[row + 1] - a.row_start[row]);
    const uint64_t idx = (uint64_t)row * a.kdim * a.n_types + j;
    if (!Grad)
        a.h[idx] = cnt * a.theta[2 * a.n_types + j];
    else
        unsafeAtomicAdd(a.grad + 2 * a.n_types + j, cnt * a.g[idx]);
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

size_t relaxed_lds_bytes(uint32_t n_types) { return (kWindowDoubles + 3ull * n_types) * sizeof(double); }

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

}  // namespace
}  // namespace procell

using namespace procell;

struct procell_relaxed {
    int device = 0;
    Setup s;
    uint32_t n_waves = 0;
    size_t lds_bytes = 0;
    RelaxArgs args{};
    void *d_block = nullptr;
    double *d_stack = nullptr;
};

namespace {

void destroy_relaxed(procell_relaxed *r)
{
    if (!r)
        return;
    (void)hipSetDevice(r->device);
    if (r->d_block)
        (void)hipFree(r->d_block);
    if (r->d_stack)
        (void)hipFree(r->d_stack);
    delete r;
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

int relaxed_launch(procell_relaxed *r, bool grad, void *stream)
{
    RELAX_TRY(hipSetDevice(r->device));
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    const RelaxArgs &a = r->args;
    const uint64_t n_h = (uint64_t)a.n_rows * a.kdim * a.n_types;
    RELAX_TRY(launch_zero(a.expansions, 1, st));
    if (!grad)
        RELAX_TRY(launch_zero((unsigned long long *)a.h, n_h, st));
    else
        RELAX_TRY(launch_zero((unsigned long long *)a.grad, 3ull * a.n_types, st));
    if (a.n_prolif && r->n_waves) {
        if (grad)
            hipLaunchKernelGGL(relaxed_kernel<true>, dim3(r->n_waves), dim3(64), r->lds_bytes, st, a);
        else
            hipLaunchKernelGGL(relaxed_kernel<false>, dim3(r->n_waves), dim3(64), r->lds_bytes, st, a);
        RELAX_TRY(hipGetLastError());
    }
    const uint64_t nq = (uint64_t)a.n_rows * a.n_quiet;
    if (nq) {
        const dim3 g((unsigned)((nq + 255) / 256));
        if (grad)
            hipLaunchKernelGGL(relaxed_quiescent_kernel<true>, g, dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(relaxed_quiescent_kernel<false>, g, dim3(256), 0, st, a);
        RELAX_TRY(hipGetLastError());
    }
    return PROCELL_OK;
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
    int rc = build_setup(p, r->s, false);
    if (rc != PROCELL_OK) {
        delete r;
        return rc;
    }
    int n_cu = 0;
    rc = relaxed_device(p->device, &n_cu);
    if (rc != PROCELL_OK) {
        delete r;
        return rc;
    }
    r->device = p->device;
    auto fail = [&](int code) {
        destroy_relaxed(r);
        return code;
    };
    const Setup &s = r->s;
    RelaxArgs &a = r->args;
    a.seed = p->seed;
    a.t_max = p->t_max;
    a.beta = beta;
    a.n_rows = (uint32_t)s.v.size();
    a.kdim = (uint32_t)s.kmax + 1;
    a.n_types = p->n_types;
    std::vector<uint32_t> prolif, quiet;
    for (uint32_t j = 0; j < p->n_types; ++j)
        (p->types[j].mean < 0.0 ? quiet : prolif).push_back(j);
    a.n_prolif = (uint32_t)prolif.size();
    a.n_quiet = (uint32_t)quiet.size();

    // ---- wavefronts: resident ones only (each owns a stack) ----
    r->lds_bytes = relaxed_lds_bytes(std::max(1u, a.n_types));
    int occ = 0;
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, relaxed_kernel<true>, 64, r->lds_bytes));
    int occ_f = 0;
    RELAX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, relaxed_kernel<false>, 64, r->lds_bytes));
    occ = std::min(occ, occ_f);
    if (occ < 1)
        return fail(PROCELL_E_HIP);
    if (p->tuning.blocks_per_cu)
        occ = std::min(occ, (int)p->tuning.blocks_per_cu);
    const uint64_t n_units = s.n_cells * (uint64_t)a.n_prolif;
    uint64_t n_waves = (uint64_t)occ * (uint64_t)n_cu;
    n_waves = std::min<uint64_t>(n_waves, std::max<uint64_t>(1, (n_units + 63) / 64));
    r->n_waves = n_units ? (uint32_t)n_waves : 0u;

    // ---- balanced ranges of walks: unit u is (cell u / n_prolif, type
    //      prolif[u % n_prolif]) ----
    const uint32_t W = std::max(1u, r->n_waves);
    std::vector<uint64_t> wave_unit(W + 1, n_units);
    std::vector<uint32_t> wave_row(2 * W, 0);
    if (n_units) {
        std::vector<double> row_cost(a.n_rows);
        double total = 0.0;
        for (uint32_t i = 0; i < a.n_rows; ++i) {
            const uint32_t kc = s.row_info[i] & 255u;
            double c = 0.0;
            for (uint32_t j : prolif)
                c += walk_cost(kc, p->t_max, p->types[j].mean);
            row_cost[i] = c / a.n_prolif;  // per unit
            total += row_cost[i] * (double)(s.start[i + 1] - s.start[i]) * a.n_prolif;
        }
        wave_unit[0] = 0;
        uint32_t wv = 1;
        double acc = 0.0;
        for (uint32_t i = 0; i < a.n_rows && wv < W; ++i) {
            const uint64_t u0 = s.start[i] * a.n_prolif, u1 = s.start[i + 1] * a.n_prolif;
            const double per = row_cost[i];
            while (wv < W) {
                const double target = total * (double)wv / (double)W;
                const double need = (target - acc) / per;
                if (need >= (double)(u1 - u0))
                    break;
                uint64_t cut = u0 + (uint64_t)std::max(0.0, ceil(need));
                cut = std::min(cut, u1);
                wave_unit[wv] = std::max(cut, wave_unit[wv - 1]);
                ++wv;
            }
            acc += per * (double)(u1 - u0);
        }
        for (; wv <= W; ++wv)
            wave_unit[wv] = n_units;
        for (uint32_t v = 1; v <= W; ++v)
            wave_unit[v] = std::max(wave_unit[v], wave_unit[v - 1]);
        auto row_of = [&](uint64_t cell) {
            return (uint32_t)(std::upper_bound(s.start.begin(), s.start.end() - 1, cell) - s.start.begin() - 1);
        };
        for (uint32_t v = 0; v < W; ++v) {
            if (wave_unit[v] >= wave_unit[v + 1])
                continue;
            wave_row[2 * v] = row_of(wave_unit[v] / a.n_prolif);
            wave_row[2 * v + 1] = row_of((wave_unit[v + 1] - 1) / a.n_prolif);
        }
    }

    // ---- device tables ----
    std::vector<uint32_t> row_start(a.n_rows + 1);
    for (uint32_t i = 0; i <= a.n_rows; ++i)
        row_start[i] = (uint32_t)s.start[i];
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_rs = take(row_start.size() * 4), o_ri = take(s.row_info.size() * 4 + 4);
    const size_t o_wu = take(wave_unit.size() * 8), o_wr = take(wave_row.size() * 4);
    const size_t o_pr = take(prolif.size() * 4 + 4), o_qu = take(quiet.size() * 4 + 4);
    const size_t o_ex = take(8);
    if (hipMalloc(&r->d_block, off) != hipSuccess)
        return fail(PROCELL_E_OUT_OF_MEMORY);
    char *base = (char *)r->d_block;
    auto up = [&](size_t o, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    RELAX_TRY(up(o_rs, row_start.data(), row_start.size() * 4));
    RELAX_TRY(up(o_ri, s.row_info.data(), s.row_info.size() * 4));
    RELAX_TRY(up(o_wu, wave_unit.data(), wave_unit.size() * 8));
    RELAX_TRY(up(o_wr, wave_row.data(), wave_row.size() * 4));
    RELAX_TRY(up(o_pr, prolif.data(), prolif.size() * 4));
    RELAX_TRY(up(o_qu, quiet.data(), quiet.size() * 4));
    a.row_start = (const uint32_t *)(base + o_rs);
    a.row_info = (const uint32_t *)(base + o_ri);
    a.wave_unit = (const uint64_t *)(base + o_wu);
    a.wave_row = (const uint32_t *)(base + o_wr);
    a.prolif = (const uint32_t *)(base + o_pr);
    a.quiet = (const uint32_t *)(base + o_qu);
    a.expansions = (unsigned long long *)(base + o_ex);
    if (r->n_waves) {
        const size_t stack_bytes = (size_t)r->n_waves * a.kdim * kFrameFields * 64 * sizeof(double);
        if (hipMalloc((void **)&r->d_stack, stack_bytes) != hipSuccess)
            return fail(PROCELL_E_OUT_OF_MEMORY);
        a.stack = r->d_stack;
    }
    *out = r;
    return PROCELL_OK;
}

int procell_relaxed_shape(const procell_relaxed *r, uint64_t *n_rows, uint32_t *kdim, uint32_t *n_types)
{
    if (!r)
        return PROCELL_E_INVALID;
    if (n_rows)
        *n_rows = r->args.n_rows;
    if (kdim)
        *kdim = r->args.kdim;
    if (n_types)
        *n_types = r->args.n_types;
    return PROCELL_OK;
}

int procell_relaxed_forward(procell_relaxed *r, const double *theta, double *h, void *stream)
{
    if (!r || !theta || !h)
        return PROCELL_E_INVALID;
    r->args.theta = theta;
    r->args.h = h;
    r->args.g = nullptr;
    r->args.grad = nullptr;
    return relaxed_launch(r, false, stream);
}

int procell_relaxed_backward(procell_relaxed *r, const double *theta, const double *g, double *grad,
                             void *stream)
{
    if (!r || !theta || !g || !grad)
        return PROCELL_E_INVALID;
    r->args.theta = theta;
    r->args.h = nullptr;
    r->args.g = g;
    r->args.grad = grad;
    return relaxed_launch(r, true, stream);
}

int procell_relaxed_expansions(procell_relaxed *r, uint64_t *n)
{
    if (!r || !n)
        return PROCELL_E_INVALID;
    RELAX_TRY(hipSetDevice(r->device));
    RELAX_TRY(hipDeviceSynchronize());
    unsigned long long v = 0;
    RELAX_TRY(hipMemcpy(&v, r->args.expansions, sizeof v, hipMemcpyDeviceToHost));
    *n = v;
    return PROCELL_OK;
}

int procell_relaxed_destroy(procell_relaxed *r)
{
    destroy_relaxed(r);
    return PROCELL_OK;
}

}  // extern "C"
