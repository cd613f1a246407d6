// procell_kernels.hip -- gfx950 kernels of the plain proliferation method.
//
// The reference materialises every tree level as a dense array of 40-byte
// cells and recurses with device-side launches (src/simulation/
// proliferation.cu:289-402), creating the initial cells in a separate pass
// (cells_population.cu:74-117).  Here one persistent kernel does the whole
// run and no cell is ever stored in device memory: each wavefront draws its
// own initial cells (root steps, 64 at a time) into a small LDS buffer,
// hands the dividing ones to idle lanes, and every lane walks its cell's
// division tree depth first, keeping only the pending right daughters'
// completion times in its column of an LDS stack.  Work is balanced without
// device-scope coordination: the initial cells are split into several
// segments of about equal expected work per resident workgroup and the
// hardware dispatcher starts the workgroup of the next segment wherever one
// finishes; inside a workgroup, wavefronts claim root steps from an LDS
// cursor, and at the very end idle lanes take pending daughters of busy
// lanes of the same wavefront.  (Claims and work stealing through device
// memory were measured here: a device-scope probe of other workgroups'
// cursors took tens of microseconds under load.)
//
// Tallies (counts by nonzero row, generation, type) accumulate in a u32 LDS
// window over the first rows of the workgroup's segment and are added to the
// int64 device tally at the end; rows outside the window use 64-bit device
// atomics (aggregated per (row, type) for generation 0).
#include "procell_internal.hpp"
#include "procell_stream.hpp"

namespace procell {
namespace {

using dev::Words;

constexpr uint32_t kWave = 64;
constexpr uint32_t kNoWindow = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        x += __shfl_xor(x, off);
    return x;
}

// Index of the r-th set bit of m (r < popcount(m)).
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t r)
{
    uint32_t base = 0;
#pragma unroll
    for (uint32_t width = 32; width >= 1; width >>= 1) {
        const uint64_t low = m & ((width == 64 ? ~0ull : (1ull << width)) - 1ull);
        const uint32_t c = (uint32_t)__popcll(low);
        if (r >= c) {
            r -= c;
            m >>= width;
            base += width;
        } else {
            m = low;
        }
    }
    return base;
}

// Wave-uniform copies (lane 0's value).  The builtin works on int, so each
// half goes back through uint32_t: a widened int would be sign-extended.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }

__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | (uint64_t)uni((uint32_t)x);
}

__global__ void zero_kernel(unsigned long long *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Per-lane event counts.  Recorded and dropped cells are not counted on the
// device: the host takes recorded = the sum of the tally and dropped = cells
// + divisions - recorded - lost (every division adds one cell, and every
// cell ends recorded, dropped or lost).
// LDS copies of the stream's lookup tables (procell_math_tables.inc).
constexpr size_t kMathTableBytes = (128 + 256) * 16;

__device__ __forceinline__ void copy_tables(double (*m_log)[2], double (*m_sc)[2], uint32_t tid, uint32_t bdim)
{
    for (uint32_t i = tid; i < 128; i += bdim) {
        m_log[i][0] = dev::procell_log_table[i][0];
        m_log[i][1] = dev::procell_log_table[i][1];
    }
    for (uint32_t i = tid; i < 256; i += bdim) {
        m_sc[i][0] = dev::procell_sincos_table[i][0];
        m_sc[i][1] = dev::procell_sincos_table[i][1];
    }
}

struct Counts {
    uint32_t div, rej, lost, cells, pool;
};

// Adds a workgroup's counts to counter copy blockIdx.x % kCounterCopies:
// wavefront sums, then one LDS atomic per wavefront and counter, then one
// device atomic per counter.  `red` is kNumCounters u32 of LDS, zeroed and
// followed by a barrier before the first call; every thread calls it.
__device__ __forceinline__ void flush_counts(const Counts &n, unsigned long long *counters, uint32_t lane,
                                             uint32_t *red)
{
    constexpr uint32_t kN = 5;
    const uint32_t v[kN] = {wave_sum(n.pool), wave_sum(n.div), wave_sum(n.lost), wave_sum(n.rej),
                            wave_sum(n.cells)};
    const uint32_t idx[kN] = {kPoolCount, kDivisions, kLost, kRejected, kCells};
    if (lane < kN) {
        uint32_t x = v[0];
#pragma unroll
        for (uint32_t i = 1; i < kN; ++i)
            x = lane == i ? v[i] : x;
        if (x)
            atomicAdd(&red[lane], x);
    }
    __syncthreads();
    if (threadIdx.x < kN) {
        const uint32_t x = red[threadIdx.x];
        uint32_t c = idx[0];
#pragma unroll
        for (uint32_t i = 1; i < kN; ++i)
            c = threadIdx.x == i ? idx[i] : c;
        if (x)
            atomicAdd(&counters[(blockIdx.x % kCounterCopies) * kNumCounters + c], (unsigned long long)x);
    }
}

// Highest / lowest set bit of a nonzero 32- or 64-bit mask.
__device__ __forceinline__ uint32_t top_bit(uint32_t m) { return 31u - (uint32_t)__builtin_clz(m); }
__device__ __forceinline__ uint32_t top_bit(uint64_t m) { return 63u - (uint32_t)__builtin_clzll(m); }
__device__ __forceinline__ uint32_t low_bit(uint32_t m) { return (uint32_t)__builtin_ctz(m); }
__device__ __forceinline__ uint32_t low_bit(uint64_t m) { return (uint32_t)__builtin_ctzll(m); }

// Lane state: the dividing cell a lane currently holds.  N holds a node's
// heap index and the pending-level mask: 32 bits when no tree is deeper than
// 31 generations, else 64.
template <typename N>
struct Lane {
    uint32_t cell;      // initial cell, relative to its set (cell ids < 2^32)
    N node, pend;
    uint32_t rowtf;     // row relative to the set's first row | file type << 24
    double s, mean, sd;
    double tl, tr;      // daughters' timers of the division in progress
    uint32_t k, kcap;
    uint32_t tb_lds;    // window index of (row, generation 0, type), or kNoWindow
    uint32_t redo;      // bit 0 / bit 1: left / right daughter's timer still to redraw
    uint32_t att;       // attempt number of that redraw
    bool active;
};

// ======================================================================
// sim_kernel: workgroup g runs segment g of the initial cells (contiguous,
// of about equal expected work; several per resident workgroup, so that the
// dispatcher balances them over the CUs); its wavefronts work independently:
//  * claims: a wavefront claims the segment's next root step (64 cells)
//    with one LDS atomic on the workgroup's cursor.
//  * root steps (cell.cu:25-143, cells_population.cu:97-117): when the
//    wavefront's buffer of dividing roots is empty and lanes are idle, the
//    64 lanes draw the next 64 initial cells of the segment, one each, whatever
//    the lanes hold: type, truncated first timer, phase.  Cells that never
//    divide are tallied at generation 0 (or lost); dividing ones become
//    48-byte entries of the wavefront's LDS buffer.
//  * hand-outs: idle lanes take buffered entries right before each division
//    step (the one place in the loop where lane state is handed over, so
//    that the compiler keeps the lane state in the same registers on every
//    path: with a second hand-out site in the refill branch the loop carried
//    about 40 register copies per refill, 3.5 % of the run); once the roots
//    are done and the buffer is empty, idle lanes take the shallowest pending
//    right daughter of a busy lane of the wavefront.
//  * refills (root steps, steals) run as a uniform branch of the one loop,
//    once more than idle_limit lanes are idle.
//  * division steps (proliferation.cu:317-381): each busy lane divides its
//    current cell per step: one Philox block and one Box-Muller pair give
//    both daughters' timers; daughters that will not divide are tallied or
//    dropped on the spot, so the lane's next cell is always a dividing one
//    (the left daughter, else the right, else the deepest pending right
//    daughter from the lane's column of the LDS stack).
// LDS: log and sin/cos tables | stream coefficients | type tables ((mean,
// sd): double2 [T]; cum: double [T]; file: u32 [T]) | row tables when they fit (first
// cell: u32 [n_rows + 1]; info: u32 [n_rows]) | counter sums u32 [8] | root
// step cursor u32 [4] | per wavefront: buffered entries QueueEntry [64] |
// window tally u32 [win_rows * kdim * T + 1 spare] | stack double
// [stack_levels * blockDim]
// ======================================================================
__host__ __device__ constexpr size_t sim_wave_lds_bytes()
{
    return (size_t)kWave * sizeof(QueueEntry);
}

constexpr size_t kSimSharedBytes = 48;  // counter sums u32 [8], root step cursor u32 [4]

template <bool RowsLds, typename N, bool GStack>
__global__ void __launch_bounds__(256, 4) sim_kernel(RunArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t bdim = blockDim.x;
    const uint32_t lane = tid & (kWave - 1);
    const uint32_t wave = tid / kWave;
    const uint32_t n_waves = bdim / kWave;
    const uint32_t seg0 = blockIdx.x;
    // the workgroup's parameter set (every segment it drains belongs to it)
    const uint32_t set = a.n_sets > 1 ? a.seg_set[seg0] : 0u;
    const SetInfo S = a.sets[set];
    const uint64_t seed = a.n_sets > 1 ? a.set_seed[set] : a.seed;
    const double t_max = S.t_max;
    const uint32_t T = S.n_types;
    const uint32_t kdim = S.kdim;
    const uint32_t win_len = a.win_rows * kdim * T;
    const uint32_t lds_rows = RowsLds ? a.n_rows : 0u;
    double(*m_log)[2] = (double(*)[2])lds;            // [128] log table
    double(*m_sc)[2] = (double(*)[2])(lds + 128 * 16);  // [256] sin/cos table
    dev::Coeffs *k_lds = (dev::Coeffs *)(lds + kMathTableBytes);
    double2 *t_msd = (double2 *)(lds + kMathTableBytes + (sizeof(dev::Coeffs) + 15) / 16 * 16);  // mean, sd
    double *t_cum = (double *)(t_msd + T);
    uint32_t *t_file = (uint32_t *)(t_cum + T);
    uint32_t *r_start = (uint32_t *)((unsigned char *)t_msd + ((size_t)T * 28 + 15) / 16 * 16);
    uint32_t *r_info = r_start + lds_rows + 1;
    uint32_t *red = (uint32_t *)((unsigned char *)r_start + (RowsLds ? ((size_t)(2 * lds_rows + 1) * 4 + 15) / 16 * 16 : 0));
    uint32_t *cursor = red + kNumCounters;  // root steps of the segment claimed so far
    QueueEntry *ent = (QueueEntry *)((unsigned char *)red + kSimSharedBytes + (size_t)wave * sim_wave_lds_bytes());
    unsigned int *win = (unsigned int *)((unsigned char *)red + kSimSharedBytes + (size_t)n_waves * sim_wave_lds_bytes());
    // the lanes' stacks: in LDS after the window, or (GStack: plans whose
    // deep trees would leave no LDS for the window) the workgroup's block of
    // device memory, same layout
    double *stack = GStack ? a.gstack + (size_t)seg0 * a.stack_levels * bdim
                           : (double *)((unsigned char *)win + ((size_t)win_len * 4 + 4 + 15) / 16 * 16);
    // the lane's stack column: level l (1-based) at stk_pop + l * stride,
    // i.e. at stk_push + (l - 1) * stride (bytes)
    const uint32_t stride = bdim * (uint32_t)sizeof(double);
    unsigned char *const stk_push = (unsigned char *)(stack + tid);
    unsigned char *const stk_pop = stk_push - stride;

    const uint64_t seg_begin = uni64(a.seg_begin[seg0]);
    const uint64_t seg_cells = uni64(a.seg_end[seg0]) - seg_begin;
    const uint32_t row_lo = uni(a.seg_row[seg0]);  // first row of the segment and of the tally window
    if (tid == 0)
        *k_lds = dev::coeffs();
    copy_tables(m_log, m_sc, tid, bdim);
    for (uint32_t i = tid; i < T; i += bdim) {
        t_cum[i] = a.type_cum[S.type_base + i];
        t_msd[i] = make_double2(a.type_mean[S.type_base + i], a.type_sd[S.type_base + i]);
        t_file[i] = a.type_file[S.type_base + i];
    }
    for (uint32_t i = tid; i <= lds_rows && RowsLds; i += bdim)
        r_start[i] = a.row_start[i];
    for (uint32_t i = tid; i < lds_rows; i += bdim)
        r_info[i] = a.row_info[i];
    for (uint32_t i = tid; i <= win_len; i += bdim)
        win[i] = 0u;  // win[win_len] absorbs the adds of lanes outside the window
    if (tid < kNumCounters + 1)
        red[tid] = 0u;  // counter sums and the root step cursor
    __syncthreads();

    // the stream's FP64 coefficients, held in vector registers for the whole
    // kernel (loaded from LDS, so they are not re-materialised as scalar literals)
    const dev::Coeffs K = *k_lds;
    const dev::Tables tb{m_log, m_sc};
    // the row tables' address space is known at compile time (no flat loads)
    const uint32_t *rs = RowsLds ? r_start : a.row_start;
    const uint32_t *ri = RowsLds ? r_info : a.row_info;

    uint32_t rpos = 0, rn = 0;      // buffered dividing roots: entries [rpos, rn) of ent
    uint32_t roots_done = 0;
    // the first eight cumulative proportions of the type search, in scalar
    // registers (-inf for types not searched: the last one and those beyond T;
    // in vector registers they would take the 16 that keep the kernel within
    // the 128 of four wavefronts per SIMD)
    double cum8[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k)
        cum8[k] = __builtin_bit_cast(double, uni64(__builtin_bit_cast(uint64_t, k + 1 < T ? t_cum[k] : -__builtin_inf())));
    // Row boundaries of the wavefront's root steps, in registers: lane j holds
    // the first cell of row rb + 1 + j (~0 past the last row) and the info of
    // row rb + j.  A wavefront's root steps only move forward in its segment,
    // so rb only grows.
    uint32_t rb = row_lo, bnd = 0, rinfo = 0;
    auto load_rows = [&]() {
        const uint32_t r1 = rb + 1 + lane;
        bnd = r1 <= a.n_rows ? rs[r1] : 0xFFFFFFFFu;
        rinfo = rb + lane < a.n_rows ? ri[rb + lane] : 0u;
    };
    load_rows();

    Lane<N> L;
    L.cell = 0;
    L.rowtf = 0;
    L.node = L.pend = 0;
    L.s = L.mean = L.sd = L.tl = L.tr = 0.0;
    L.k = L.kcap = 0;
    L.redo = 0;
    L.att = 1;
    L.tb_lds = kNoWindow;
    L.active = false;
    Counts n = {0, 0, 0, 0, 0};
    uint32_t w_div = 0;  // divisions of the wavefront's common-case steps (scalar)
    const uint32_t idle_limit = a.idle_limit;

    // idle lanes take buffered dividing roots, in lane order
    auto hand_out = [&](uint64_t idle, uint32_t n_idle) {
        const uint32_t rank = lane_rank(idle);
        const uint32_t avail = rn - rpos;
        if (!L.active && rank < avail) {
            const QueueEntry &e = ent[rpos + rank];
            L.cell = e.cell;
            L.s = e.s0;
            L.mean = e.mean;
            L.sd = e.sd;
            L.node = (N)1;
            L.k = 0u;
            L.pend = (N)0;
            L.redo = 0u;
            L.kcap = e.kcap;
            L.tb_lds = e.tb_lds;
            L.rowtf = e.rowtf;
            L.active = true;
        }
        rpos = uni(rpos + min(n_idle, avail));
    };

    for (;;) {
        const uint64_t idle = __ballot(!L.active);
        const uint32_t n_idle = uni((uint32_t)__popcll(idle));
        // ======== refill once more than idle_limit lanes are idle ========
        if (n_idle > idle_limit) {
            // ---- root step: the buffer is empty ----
            if (rpos >= rn && !roots_done) {
                // claim the segment's next root step
                uint32_t step = 0;
                if (lane == 0)
                    step = atomicAdd(cursor, 1u);
                const uint64_t c_off = (uint64_t)uni(__shfl(step, 0)) * kWave;
                if (c_off >= seg_cells)
                    roots_done = 1;
                if (!roots_done) {
                    const uint64_t c_cur = seg_begin + c_off;
                    const uint32_t n_valid = uni((uint32_t)min<uint64_t>(kWave, seg_cells - c_off));
                    const uint64_t c = c_cur + lane;
                    const bool valid = lane < n_valid;
                    // rows of the step's cells (cell ids < 2^32): move rb to the
                    // row of c_cur, then count the boundaries inside the step
                    const uint32_t c32 = (uint32_t)c_cur;
                    for (uint64_t le = __ballot(bnd <= c32); le; le = __ballot(bnd <= c32)) {
                        rb = uni(rb + (uint32_t)__popcll(le));
                        load_rows();
                    }
                    const uint32_t nb = uni((uint32_t)__popcll(__ballot((uint64_t)bnd < c_cur + n_valid)));
                    uint32_t rrel = 0;  // boundaries at or below this lane's cell
                    for (uint32_t j = 0; j < nb; ++j)
                        rrel += lane >= (uint32_t)__builtin_amdgcn_readlane((int)bnd, (int)j) - c32 ? 1u : 0u;
                    const uint32_t row = rb + rrel;
                    const uint32_t info = nb ? (uint32_t)__shfl(rinfo, (int)rrel)
                                             : (uint32_t)__builtin_amdgcn_readfirstlane((int)rinfo);
                    const uint32_t kc = info & 0xFFu;
                    const bool ge_phi = (info >> 8) & 1u;
                    // one block per initial cell: type uniform (word 0), phase
                    // (word 1) and the first timer's Box-Muller pair (words 2, 3)
                    const Words w = dev::draw_block(seed, c - S.cell_base, 0, 0);
                    const double ut = dev::uniform_co_hi(w.w0);
                    const double U = dev::uniform_oc_hi(w.w1);
                    // first type with ut < cumulative proportion (cell.cu:90-101):
                    // scanning from the highest index down, the last hit is the
                    // first; past the last proportion, the last type
                    uint32_t jt = T - 1;
                    if (T > 9) {
                        for (uint32_t k = T - 2; k >= 8; --k)
                            jt = ut < t_cum[k] ? k : jt;
                    }
    #pragma unroll
                    for (int k = 7; k >= 0; --k)
                        jt = ut < cum8[k] ? (uint32_t)k : jt;
                    const double2 msd = t_msd[jt];
                    const double m = msd.x;
                    const double sdv = msd.y;
                    const bool quiescent = m < 0.0;
                    // first timer (cell.cu:106-122): two candidates from the root
                    // block, then node 1's redraw stream
                    // (the second candidate only where the first is <= 0: its
                    // sine is computed there alone)
                    double z0, z1;
                    dev::normal_pair<true>(Words{w.w2, 0u, w.w3, 0u}, z0, z1, K, tb);
                    const double c0 = fma(z0, sdv, m);
                    const bool live = valid && !quiescent;
                    double tau0 = c0;
                    if (__builtin_expect(__ballot(live && !(c0 > 0.0)) != 0, 0)) {
                        if (live && !(c0 > 0.0)) {
                            ++n.rej;
                            const double c1 = fma(z1, sdv, m);
                            tau0 = c1;
                        }
                        if (live && !(tau0 > 0.0)) {
                            // rare: node 1's redraw stream from attempt 1
                            ++n.rej;
                            tau0 = m;
                            for (uint32_t att = 1; att < 65536u; ++att) {
                                const Words wr = dev::draw_block(seed, c - S.cell_base, 1, att);
                                double y0, y1;
                                dev::normal_pair(wr, y0, y1, K, tb);
                                const double d0 = fma(y0, sdv, m);
                                if (d0 > 0.0) { tau0 = d0; break; }
                                ++n.rej;
                                const double d1 = fma(y1, sdv, m);
                                if (d1 > 0.0) { tau0 = d1; break; }
                                ++n.rej;
                            }
                        }
                    }
                    const double t0 = tau0 * U;
                    const double s0 = t0 + tau0;
                    // quiescent or out of time: tallied if its value is in the
                    // output bin set; in time with f/2 <= phi: dropped; else it
                    // divides
                    const bool record = valid && (quiescent || s0 > t_max);
                    const bool rec_ok = record && ge_phi;
                    const bool divides = valid && !record && kc != 0;
                    n.cells += valid ? 1u : 0u;
                    n.lost += record && !ge_phi ? 1u : 0u;
                    n.pool += divides ? 1u : 0u;
                    // generation-0 tallies: the window, else one device atomic
                    // per distinct (row, type) of the step
                    const uint32_t tf = t_file[jt];
                    const uint32_t woff = row - row_lo;
                    const bool in_win = woff < a.win_rows;
                    if (rec_ok && in_win) {
                        atomicAdd(&win[woff * kdim * T + tf], 1u);
                    }
                    uint64_t pend = __ballot(rec_ok && !in_win);
                    while (pend) {
                        const uint32_t leader = (uint32_t)__builtin_ctzll(pend);
                        const uint32_t row_l = uni(__shfl(row, (int)leader));
                        const uint32_t tf_l = uni(__shfl(tf, (int)leader));
                        const uint64_t same = __ballot(rec_ok && !in_win && row == row_l && tf == tf_l);
                        if (lane == leader)
                            atomicAdd(&a.tally[S.tally_base + (uint64_t)(row_l - S.row_base) * kdim * T + tf_l],
                                      (unsigned long long)__popcll(same));
                        pend &= ~same;
                    }
                    // dividing cells: buffered for the hand-outs
                    const uint64_t dmask = __ballot(divides);
                    if (divides) {
                        QueueEntry &e = ent[lane_rank(dmask)];
                        e.s0 = s0;
                        e.mean = m;
                        e.sd = sdv;
                        e.cell = (uint32_t)(c - S.cell_base);
                        e.tb_lds = in_win ? woff * kdim * T + tf : kNoWindow;
                        e.rowtf = (row - S.row_base) | (tf << 24);
                        e.kcap = kc;
                    }
                    rpos = 0;
                    rn = uni((uint32_t)__popcll(dmask));
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // ---- steal pending daughters inside the wavefront ----
            if (rpos >= rn && roots_done) {
                const uint64_t donors = __ballot(L.active && L.pend != 0);
                const uint32_t nd = uni(min((uint32_t)__popcll(donors), n_idle));
                if (nd) {
                    const uint32_t my_lvl = L.pend ? low_bit(L.pend) : 0u;
                    const uint32_t r = lane_rank(idle);
                    const bool thief = !L.active && r < nd;
                    const uint32_t src = thief ? nth_set_bit(donors, r) : lane;
                    const uint32_t lvl = __shfl(my_lvl, (int)src);
                    const N d_node = __shfl(L.node, (int)src);
                    const uint32_t d_k = __shfl(L.k, (int)src);
                    const uint32_t d_cell = __shfl(L.cell, (int)src);
                    const uint32_t d_kcap = __shfl(L.kcap, (int)src);
                    const uint32_t d_tbl = __shfl(L.tb_lds, (int)src);
                    const uint32_t d_rtf = __shfl(L.rowtf, (int)src);
                    const double d_mean = __shfl(L.mean, (int)src);
                    const double d_sd = __shfl(L.sd, (int)src);
                    if (L.active && L.pend && lane_rank(donors) < nd)
                        L.pend &= ~((N)1 << my_lvl);
                    if (thief) {
                        L.s = stack[(size_t)(lvl - 1) * bdim + wave * kWave + src];
                        L.node = (d_node >> (d_k - lvl)) | (N)1;
                        L.k = lvl;
                        L.cell = d_cell;
                        L.kcap = d_kcap;
                        L.tb_lds = d_tbl;
                        L.rowtf = d_rtf;
                        L.mean = d_mean;
                        L.sd = d_sd;
                        L.pend = 0;
                        L.redo = 0;
                        L.active = true;
                    }
                }
            }

            if (rpos >= rn && __ballot(L.active) == 0) {
                if (roots_done)
                    break;
                continue;
            }
        }
        // ---- idle lanes take buffered roots ----
        if (rpos < rn) {
            const uint64_t idle2 = __ballot(!L.active);
            hand_out(idle2, uni((uint32_t)__popcll(idle2)));
        }

        // ======== division step ========
        // Straight-line code: every outcome is a select.  The step's flags
        // are lane masks, so their logic runs on the scalar unit beside the
        // vector work (the vector unit is the bound: 4.9 VALU instructions
        // per division against 1.1 scalar).  Idle lanes run the step too
        // with their results discarded.
        {
            // One Philox block + Box-Muller pair per lane and step: either the
            // pair block of a new division (timers of both daughters) or one
            // block of a daughter's truncated-normal redraw stream
            // (cell.cu:106-122); a lane with a redraw pending does not advance
            // its tree in that step.
            const uint32_t act = L.active ? 1u : 0u;
            // the level a pop at the end of this step reads: a lane pops
            // only when the step pushes nothing, so it is the deepest
            // pending level of the step's start; reading it here overlaps
            // the LDS round trip with the division step
            const uint32_t lvl = top_bit(L.pend | (N)2);  // 1 when nothing is pending (level bits start at 1)
            const double sp = *(const double *)(stk_pop + __umul24(lvl, stride));
            double tl, tr;
            bool cmp;  // division complete this step
            if (__builtin_expect(__ballot(L.redo != 0) == 0, 1)) {
                // no redraw pending in the wavefront (the common case): every
                // busy lane starts a new division with its pair block; every
                // lane's redo is 0 and its attempt number 1 (attempts restart
                // at 1 whenever a redraw ends)
                const Words w = dev::draw_block(seed, (uint64_t)L.cell, L.node, 0u);
                double z0, z1;
                dev::normal_pair(w, z0, z1, K, tb);
                tl = fma(z0, L.sd, L.mean);
                tr = fma(z1, L.sd, L.mean);
                const bool ok = tl > 0.0 && tr > 0.0;
                cmp = L.active && ok;
                w_div += (uint32_t)__popcll(__ballot(L.active));
                if (__builtin_expect(__ballot(L.active && !ok) != 0, 0)) {
                    // a timer <= 0: that daughter continues in its redraw stream
                    const uint32_t r0 = tl > 0.0 ? 0u : 1u;
                    const uint32_t r1 = tr > 0.0 ? 0u : 1u;
                    L.redo = act ? r0 | (r1 << 1) : 0u;
                    n.rej += act * (r0 + r1);
                }
            } else {
                const uint32_t fresh = L.redo == 0 ? 1u : 0u;
                const uint32_t rright = (L.redo & 1u) ? 0u : 1u;   // redraw the right daughter?
                const N bnode = fresh ? L.node : (L.node << 1) | (N)rright;
                const Words w = dev::draw_block(seed, (uint64_t)L.cell, bnode, fresh ? 0u : L.att);
                double z0, z1;
                dev::normal_pair(w, z0, z1, K, tb);
                const double c0 = fma(z0, L.sd, L.mean);
                const double c1 = fma(z1, L.sd, L.mean);
                const uint32_t r0 = c0 > 0.0 ? 0u : 1u;
                const uint32_t r1 = c1 > 0.0 ? 0u : 1u;
                // redraw: first candidate > 0 of this block; after 65535 blocks the mean
                const uint32_t got = (r0 & r1) ^ 1u;
                const uint32_t give_up = (got ^ 1u) & (L.att >= 65535u ? 1u : 0u);
                const double v = give_up ? L.mean : (r0 ? c1 : c0);
                const uint32_t done = got | give_up;
                const uint32_t rbit = rright ? 2u : 1u;
                tl = fresh ? c0 : ((done & (rright ^ 1u)) ? v : L.tl);
                tr = fresh ? c1 : ((done & rright) ? v : L.tr);
                const uint32_t redo = fresh ? (r0 | (r1 << 1)) : (done ? L.redo & ~rbit : L.redo);
                n.div += act & fresh;
                n.rej += act * (fresh ? r0 + r1 : r0 + (r0 & r1));
                L.att = fresh | done ? 1u : L.att + 1u;
                L.redo = act ? redo : 0u;
                cmp = L.active && redo == 0;
            }
            L.tl = tl;
            L.tr = tr;

            // daughters (proliferation.cu:317-381): out of time -> tallied,
            // f/2 <= phi -> dropped, else divides later.  The flags below are
            // lane masks (scalar registers), so their logic runs on the
            // scalar unit.
            const double sl = L.s + tl;
            const double sr = L.s + tr;
            const uint32_t k1 = L.k + 1;
            const bool rec_l = sl > t_max;
            const bool rec_r = sr > t_max;
            const bool can = k1 < L.kcap;
            const bool div_l = !rec_l && can;
            const bool div_r = !rec_r && can;
            const uint32_t nrec = cmp ? (uint32_t)rec_l + (uint32_t)rec_r : 0u;
            const bool in_win = L.tb_lds != kNoWindow;
            // lanes without a completed division add 0 to the spare slot win[win_len]
            atomicAdd(&win[(cmp && in_win) ? L.tb_lds + __umul24(k1, T) : win_len], nrec);
            if (__builtin_expect(__ballot(!in_win && nrec != 0) != 0, 0)) {
                if (!in_win && nrec != 0)
                    atomicAdd(&a.tally[S.tally_base + (uint64_t)(L.rowtf & 0xFFFFFFu) * kdim * T +
                                       (L.rowtf >> 24) + (uint64_t)k1 * T],
                              (unsigned long long)nrec);
            }
            // right daughter's completion time, at level k1; level k1 holds
            // no pending entry of an active lane, and idle lanes have none at
            // all (their L.k is at most 1 or a level they popped, below the
            // stack's depth), so the store is harmless when it is not pushed
            *(double *)(stk_push + __umul24(L.k, stride)) = sr;
            const bool go_l = cmp && div_l;
            const bool go_r = cmp && !div_l && div_r;
            const bool pop = cmp && !div_l && !div_r;
            // (k1 stays below the width of N for a lane that pushes; the
            // mask keeps the shift defined for the others)
            const N pend2 = L.pend | ((N)(go_l && div_r) << (k1 & (sizeof(N) * 8 - 1)));
            const bool has = pend2 != 0;
            // single-level selects only (nested ones turn into branches)
            const bool down = go_l || go_r;
            const N child = (L.node << 1) | (N)go_r;
            const uint32_t sh = pop ? ((L.k - lvl) & (sizeof(N) * 8 - 1)) : 0u;
            const N upn = (L.node >> sh) | (N)pop;  // pending right daughter
            const double s_child = go_l ? sl : sr;
            const double s_up = pop ? sp : L.s;
            const uint32_t k_up = pop ? lvl : L.k;
            const N clear = (N)pop << lvl;
            L.node = down ? child : upn;
            L.s = down ? s_child : s_up;
            L.k = down ? k1 : k_up;
            L.pend = pend2 & ~clear;
            L.active = L.active && !(pop && !has);
        }
    }

    n.div += lane == 0 ? w_div : 0u;
    flush_counts(n, a.counters + (size_t)set * kCounterCopies * kNumCounters, lane, red);
    const uint64_t gbase = S.tally_base + (uint64_t)(row_lo - S.row_base) * kdim * T;
    for (uint32_t i = tid; i < win_len; i += bdim) {
        const uint32_t c = win[i];
        if (c)
            atomicAdd(&a.tally[gbase + i], (unsigned long long)c);
    }
}

__global__ void debug_draws_kernel(uint64_t seed, const uint64_t *cell, const uint64_t *node,
                                   const uint32_t *attempt, uint64_t n, uint32_t *words, double *z)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const Words w = dev::draw_block(seed, cell[i], node[i], attempt[i]);
    words[4 * i + 0] = w.w0;
    words[4 * i + 1] = w.w1;
    words[4 * i + 2] = w.w2;
    words[4 * i + 3] = w.w3;
    double z0, z1;
    dev::normal_pair(w, z0, z1, dev::coeffs(), dev::global_tables());
    z[2 * i] = z0;
    z[2 * i + 1] = z1;
}

}  // namespace

size_t sim_lds_bytes(uint32_t win_rows, uint32_t kdim, uint32_t n_types, uint32_t stack_levels, uint32_t block_size,
                     uint32_t lds_rows)
{
    const size_t tables = kMathTableBytes + (sizeof(dev::Coeffs) + 15) / 16 * 16 + ((size_t)n_types * 28 + 15) / 16 * 16 +
                          (lds_rows ? ((size_t)(2 * lds_rows + 1) * 4 + 15) / 16 * 16 : 0) + kSimSharedBytes;
    const size_t waves = (size_t)(block_size / kWave) * sim_wave_lds_bytes();
    const size_t win = ((size_t)win_rows * kdim * n_types * 4 + 4 + 15) / 16 * 16;
    const size_t stk = (size_t)stack_levels * block_size * sizeof(double);
    return tables + waves + win + stk;
}

hipError_t launch_zero(unsigned long long *p, uint64_t n, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

template <bool GStack>
void launch_sim_as(const RunArgs &a, const LaunchShape &ls, hipStream_t st)
{
    const dim3 g(ls.blocks), b(ls.block_size);
    const bool narrow = a.kdim <= 32;  // heap indices and level masks fit 32 bits
    if (a.rows_in_lds && narrow)
        hipLaunchKernelGGL((sim_kernel<true, uint32_t, GStack>), g, b, ls.lds_bytes, st, a);
    else if (a.rows_in_lds)
        hipLaunchKernelGGL((sim_kernel<true, uint64_t, GStack>), g, b, ls.lds_bytes, st, a);
    else if (narrow)
        hipLaunchKernelGGL((sim_kernel<false, uint32_t, GStack>), g, b, ls.lds_bytes, st, a);
    else
        hipLaunchKernelGGL((sim_kernel<false, uint64_t, GStack>), g, b, ls.lds_bytes, st, a);
}

hipError_t launch_sim(const RunArgs &a, const LaunchShape &ls, hipStream_t st)
{
    if (a.n_cells == 0)
        return hipSuccess;
    if (a.gstack)
        launch_sim_as<true>(a, ls, st);
    else
        launch_sim_as<false>(a, ls, st);
    return hipGetLastError();
}

hipError_t launch_debug_draws(uint64_t seed, const uint64_t *cell, const uint64_t *node,
                              const uint32_t *attempt, uint64_t n, uint32_t *words, double *z,
                              hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(debug_draws_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, seed,
                       cell, node, attempt, n, words, z);
    return hipGetLastError();
}

template <bool GStack>
hipError_t sim_occupancy_as(int block_size, size_t lds_bytes, bool rows_in_lds, uint32_t kdim, int *blocks_per_cu)
{
    const bool narrow = kdim <= 32;
    if (rows_in_lds && narrow)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sim_kernel<true, uint32_t, GStack>,
                                                            block_size, lds_bytes);
    if (rows_in_lds)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sim_kernel<true, uint64_t, GStack>,
                                                            block_size, lds_bytes);
    if (narrow)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sim_kernel<false, uint32_t, GStack>,
                                                            block_size, lds_bytes);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sim_kernel<false, uint64_t, GStack>,
                                                        block_size, lds_bytes);
}

hipError_t sim_occupancy(int block_size, size_t lds_bytes, bool rows_in_lds, uint32_t kdim, bool gstack,
                         int *blocks_per_cu)
{
    return gstack ? sim_occupancy_as<true>(block_size, lds_bytes, rows_in_lds, kdim, blocks_per_cu)
                  : sim_occupancy_as<false>(block_size, lds_bytes, rows_in_lds, kdim, blocks_per_cu);
}

}  // namespace procell
