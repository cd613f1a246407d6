// procell_kernels.hip -- gfx950 kernels of the plain proliferation method.
//
// The reference materialises every tree level as a dense array of 40-byte
// cells and recurses with device-side launches (src/simulation/
// proliferation.cu:289-402), creating the initial cells in a separate pass
// (cells_population.cu:74-117).  Here two kernels run back to back on one
// stream and no cell below the root is ever stored in device memory:
//
//  * root_kernel (cell.cu:25-143).  One initial cell per lane, no divergence:
//    type draw, truncated first timer, phase.  Cells that never divide are
//    tallied (generation 0) or dropped at once; a dividing cell becomes a
//    16-byte queue entry.  The initial cells are split on the host into one
//    contiguous segment per tree-kernel workgroup, of roughly equal expected
//    work, and each segment into units of kUnit cells, one wavefront's work
//    item each.  A unit stages its entries in LDS and appends them to its
//    segment's part of the queue with one atomic, so each segment's entries
//    are contiguous.
//  * tree_kernel (proliferation.cu:317-381).  Persistent workgroups, one per
//    segment.  A wavefront claims entries with one atomic per claim, first
//    from its workgroup's segment, then, once that is drained, from the
//    segment with the most unclaimed entries among 64 probed ones (work
//    stealing across workgroups), which it keeps draining before it probes
//    again.  Claims are a share of what the segment has left (guided
//    self-scheduling), so they shrink towards its end and the wavefronts
//    finish together; the next claim is issued when the last entries of the
//    current one are loaded, so its latency overlaps the division steps.
//    Entries are loaded 64 at a time into two register sets (one entry per
//    lane and set).  Each busy lane divides its current cell
//    per step: one Philox block and one Box-Muller pair give both daughters'
//    timers; daughters that will not divide are tallied or dropped on the
//    spot, so the lane's next cell is always a dividing one (the left
//    daughter, else the right, else the deepest pending right daughter from
//    the lane's column of the LDS stack, else a queue entry).  Once no entry
//    is left anywhere, idle lanes take the shallowest pending right daughter
//    of a busy lane in the same wavefront (work stealing inside the
//    wavefront).
//  * Tallies (counts by nonzero row, generation, type) of the tree kernel
//    accumulate in a u32 LDS window over the first rows of the workgroup's
//    segment and are added to the int64 device tally at the end; rows outside
//    the window use 64-bit device atomics.  The root kernel aggregates its
//    generation-0 tallies per wavefront before one device atomic per
//    distinct (row, type).
#include "procell_internal.hpp"
#include "procell_stream.hpp"

namespace procell {
namespace {

using dev::Words;

constexpr uint32_t kWave = 64;
constexpr uint32_t kNoWindow = 0xFFFFFFFFu;
constexpr uint32_t kProbes = 32;       // segments one steal scan probes

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

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// 32-bit integer mix (the finaliser of MurmurHash3) for spreading probes.
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)x);
}

__global__ void zero_kernel(unsigned long long *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// First row of a range: 64-ary search with one load per lane per round over
// rows [lo, hi) (row_start[lo] <= cell < row_start[hi]).
__device__ __forceinline__ uint32_t find_row_wave(const uint32_t *row_start, uint32_t lo, uint32_t hi,
                                                  uint64_t cell, uint32_t lane)
{
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + kWave - 1) / kWave;
        const uint32_t cand = lo + lane * step;
        const bool le = cand < hi && row_start[cand] <= cell;
        const uint64_t m = __ballot(le);
        const uint32_t top = 63u - (uint32_t)__builtin_clzll(m);  // lane 0 always qualifies
        const uint32_t new_lo = lo + top * step;
        const uint32_t new_hi = new_lo + step < hi ? new_lo + step : hi;
        lo = new_lo;
        hi = new_hi;
    }
    return lo;
}

// Row of the cell c_cur + lane of a root step (lanes < n_valid), given the row
// row_cur of c_cur.  Either the whole step lies in row_cur, or the rows that
// start inside the step are marked in the wavefront's 64 flags (one load per
// lane) and each lane counts the starts at or below it.
__device__ __forceinline__ uint32_t step_row(const uint32_t *row_start, uint32_t n_rows, uint32_t row_cur,
                                             uint64_t c_cur, uint32_t n_valid, uint32_t lane, uint32_t *flags)
{
    const uint64_t next = row_cur + 1 < n_rows ? (uint64_t)row_start[row_cur + 1] : ~0ull;
    if (__builtin_amdgcn_readfirstlane((uint32_t)(next > c_cur + n_valid - 1)))
        return row_cur;
    flags[lane] = 0u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t ri = row_cur + 1 + lane;
    if (ri < n_rows) {
        const uint64_t t = (uint64_t)row_start[ri] - c_cur;  // >= 0: row_cur holds c_cur
        if (t < n_valid)
            flags[t] = 1u;
    }
    __builtin_amdgcn_wave_barrier();
    const uint64_t m = __ballot(flags[lane] != 0u);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    return row_cur + (uint32_t)__popcll(m & upto);
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

// ======================================================================
// root_kernel: persistent wavefronts working independently on units of
// kUnit initial cells (cells [lo, hi) of segment unit_seg[u], fewer at the
// segment's end).  Wavefront i of the grid takes units i, i + W, i + 2W,
// ... (W wavefronts; units have equal cost, and a counter shared by all
// wavefronts would serialise its atomics).  A unit's dividing cells are staged in the wavefront's LDS area and
// appended to the segment's part of the queue with one atomic; its
// generation-0 tallies go to the wavefront's LDS window over the unit's
// first rows.  No workgroup barrier inside the loop, so one wavefront's
// atomic latencies are hidden by the others.
// LDS: log and sin/cos tables | type tables (cum, mean, sd: double [T];
// file: u32 [T]) | row tables when they fit (first cell: u32 [n_rows + 1];
// info: u32 [n_rows]) | counter sums u32 [8] | per wavefront: row-start
// flags u32 [64], generation-0 tally window u32 [root_win_rows(T) * T + 1
// spare], staged entries QueueEntry [kUnit]
// ======================================================================
__host__ __device__ constexpr size_t root_wave_lds_bytes(uint32_t max_types)
{
    return kWave * 4 + ((size_t)root_win_cap(max_types) * 4 + 4 + 15) / 16 * 16 + (size_t)kUnit * sizeof(QueueEntry);
}

// Adds a wavefront's counts to counter copy `copy` of set `set` and clears
// them (batched runs: a wavefront's units belong to different sets).
__device__ __forceinline__ void flush_counts_wave(Counts &n, unsigned long long *counters, uint32_t set,
                                                  uint32_t copy, uint32_t lane)
{
    constexpr uint32_t kN = 5;
    const uint32_t v[kN] = {wave_sum(n.pool), wave_sum(n.div), wave_sum(n.lost), wave_sum(n.rej),
                            wave_sum(n.cells)};
    const uint32_t idx[kN] = {kPoolCount, kDivisions, kLost, kRejected, kCells};
    unsigned long long *c = counters + ((size_t)set * kCounterCopies + copy) * kNumCounters;
    if (lane < kN) {
        uint32_t x = v[0], ci = idx[0];
#pragma unroll
        for (uint32_t i = 1; i < kN; ++i) {
            x = lane == i ? v[i] : x;
            ci = lane == i ? idx[i] : ci;
        }
        if (x)
            atomicAdd(&c[ci], (unsigned long long)x);
    }
    n = Counts{0, 0, 0, 0, 0};
}

// Batched: the units of one wavefront belong to different parameter sets;
// the set's values are loaded when it changes, the type tables are read
// from device memory and counts are flushed per unit.
template <bool RowsLds, bool Batched>
__global__ void __launch_bounds__(256) root_kernel(RunArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t bdim = blockDim.x;
    const uint32_t lane = tid & (kWave - 1);
    const uint32_t wave = tid / kWave;
    const uint32_t waves = bdim / kWave;
    const uint32_t Tl = a.n_types;  // LDS layout: the most types of any set
    const uint32_t lds_rows = RowsLds ? a.n_rows : 0u;
    double(*m_log)[2] = (double(*)[2])lds;            // [128] log table
    double(*m_sc)[2] = (double(*)[2])(lds + 128 * 16);  // [256] sin/cos table
    double *t_cum = (double *)(lds + kMathTableBytes);
    double *t_mean = t_cum + Tl;
    double *t_sd = t_mean + Tl;
    uint32_t *t_file = (uint32_t *)(t_sd + Tl);
    uint32_t *r_start = (uint32_t *)(lds + kMathTableBytes + ((size_t)Tl * 28 + 15) / 16 * 16);
    uint32_t *r_info = r_start + lds_rows + 1;
    uint32_t *red = (uint32_t *)((unsigned char *)r_start + ((size_t)(2 * lds_rows + 1) * 4 + 15) / 16 * 16);
    unsigned char *wlds = (unsigned char *)red + 32 + (size_t)wave * root_wave_lds_bytes(Tl);
    uint32_t *flags = (uint32_t *)wlds;
    uint32_t *rwin = flags + kWave;
    QueueEntry *stage = (QueueEntry *)(wlds + kWave * 4 + ((size_t)root_win_cap(Tl) * 4 + 4 + 15) / 16 * 16);

    // the current parameter set (the only one unless Batched)
    uint32_t set = 0;
    SetInfo S = a.sets[0];
    uint64_t seed = Batched ? a.set_seed[0] : a.seed;
    uint32_t T = S.n_types;
    uint32_t kdim = S.kdim;
    double t_max = S.t_max;

    copy_tables(m_log, m_sc, tid, bdim);
    for (uint32_t i = tid; i < (Batched ? 0u : T); i += bdim) {
        t_cum[i] = a.type_cum[i];
        t_mean[i] = a.type_mean[i];
        t_sd[i] = a.type_sd[i];
        t_file[i] = a.type_file[i];
    }
    for (uint32_t i = tid; i <= lds_rows; i += bdim)
        r_start[i] = a.row_start[i];
    for (uint32_t i = tid; i < lds_rows; i += bdim)
        r_info[i] = a.row_info[i];
    if (tid < kNumCounters)
        red[tid] = 0u;
    __syncthreads();
    // the row tables' address space is known at compile time (no flat loads)
    const uint32_t *rs = RowsLds ? r_start : a.row_start;
    const uint32_t *ri = RowsLds ? r_info : a.row_info;
    const dev::Coeffs K = dev::coeffs();
    const dev::Tables tb{m_log, m_sc};
    Counts n = {0, 0, 0, 0, 0};
    const uint32_t n_w = gridDim.x * waves;
    // the first eight cumulative proportions of the type search (-inf for
    // types that are not searched: the last one and those beyond T)
    double cum8[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k)
        cum8[k] = k + 1 < T ? (Batched ? a.type_cum[k] : t_cum[k]) : -__builtin_inf();

    uint32_t u = blockIdx.x * waves + wave;
    while (u < a.n_units) {
        const uint32_t g = uni(a.unit_seg[u]);
        const uint64_t sb = uni64(a.seg_begin[g]);
        const uint64_t lo = uni64(sb + (uint64_t)(u - a.unit_first[g]) * kUnit);
        const uint64_t hi = uni64(min(lo + kUnit, a.seg_begin[g + 1]));
        if (Batched) {
            const uint32_t su = uni(a.seg_set[g]);
            if (su != set) {
                flush_counts_wave(n, a.counters, set, (blockIdx.x * waves + wave) % kCounterCopies, lane);
                set = su;
                S = a.sets[set];
                seed = a.set_seed[set];
                T = S.n_types;
                kdim = S.kdim;
                t_max = S.t_max;
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k)
                    cum8[k] = k + 1 < T ? a.type_cum[S.type_base + k] : -__builtin_inf();
            }
        }
        const uint32_t rwin_rows = root_win_rows(T);
        const uint32_t rwin_len = rwin_rows * T;
        for (uint32_t i = lane; i <= rwin_len; i += kWave)
            rwin[i] = 0u;  // rwin[rwin_len] absorbs the adds of lanes outside the window
        __builtin_amdgcn_wave_barrier();

        const uint32_t row_lo = uni(find_row_wave(rs, 0, a.n_rows, lo, lane));  // first row of the window
        uint32_t row_cur = row_lo;
        uint32_t cnt = 0;  // entries staged so far
        for (uint64_t c_cur = lo; c_cur < hi; c_cur += kWave) {
            const uint32_t n_valid = uni(hi - c_cur < kWave ? (uint32_t)(hi - c_cur) : kWave);
            const uint64_t c = c_cur + lane;
            const bool valid = lane < n_valid;
            const uint32_t row = step_row(rs, a.n_rows, row_cur, c_cur, n_valid, lane, flags);
            const uint32_t info = ri[valid ? row : row_cur];
            const uint32_t kc = info & 0xFFu;
            const bool ge_phi = (info >> 8) & 1u;
            // one block per initial cell: type uniform (word 0), phase (word
            // 1) and the first timer's Box-Muller pair (words 2, 3)
            const Words w = dev::draw_block(seed, c - S.cell_base, 0, 0);
            const double ut = dev::uniform_co(w.w0, 0u);
            const double U = dev::uniform_oc(w.w1, 0u);
            // first type with ut < cumulative proportion (cell.cu:90-101):
            // scanning from the highest index down, the last hit is the
            // first; the first eight proportions are held in registers
            uint32_t jt = T - 1;
            if (T > 9) {
                for (uint32_t k = T - 2; k >= 8; --k)
                    jt = ut < (Batched ? a.type_cum[S.type_base + k] : t_cum[k]) ? k : jt;
            }
#pragma unroll
            for (int k = 7; k >= 0; --k)
                jt = ut < cum8[k] ? (uint32_t)k : jt;
            const double m = Batched ? a.type_mean[S.type_base + jt] : t_mean[jt];
            const double sdv = Batched ? a.type_sd[S.type_base + jt] : t_sd[jt];
            const bool quiescent = m < 0.0;
            // first timer (cell.cu:106-122): two candidates from the root
            // block, then node 1's redraw stream
            double z0, z1;
            dev::normal_pair(Words{w.w2, 0u, w.w3, 0u}, z0, z1, K, tb);
            const double c0 = fma(z0, sdv, m);
            const double c1 = fma(z1, sdv, m);
            const bool live = valid && !quiescent;
            n.rej += live ? (uint32_t)!(c0 > 0.0) + (uint32_t)(!(c0 > 0.0) && !(c1 > 0.0)) : 0u;
            double tau0 = c0 > 0.0 ? c0 : c1;
            if (__builtin_expect(__ballot(live && !(tau0 > 0.0)) != 0, 0)) {
                if (live && !(tau0 > 0.0)) {
                    // rare: node 1's redraw stream from attempt 1
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
            // quiescent or out of time: tallied if its value is in the output
            // bin set; in time with f/2 <= phi: dropped; else it divides
            const bool record = valid && (quiescent || s0 > t_max);
            const bool rec_ok = record && ge_phi;
            const bool divides = valid && !record && kc != 0;
            n.cells += valid ? 1u : 0u;
            n.lost += record && !ge_phi ? 1u : 0u;
            n.pool += divides ? 1u : 0u;
            // generation-0 tallies: the LDS window, else one device atomic per
            // distinct (row, type) of the step
            const uint32_t tf = Batched ? a.type_file[S.type_base + jt] : t_file[jt];
            const uint32_t woff = row - row_lo;
            const bool in_win = woff < rwin_rows;
            atomicAdd(&rwin[rec_ok && in_win ? woff * T + tf : rwin_len], rec_ok && in_win ? 1u : 0u);
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
            // dividing cells: staged in LDS, appended to the queue at the end
            const uint64_t dmask = __ballot(divides);
            if (divides) {
                QueueEntry &e = stage[cnt + lane_rank(dmask)];
                e.s0 = s0;
                e.cell = (uint32_t)(c - S.cell_base);
                e.row_type = row | (jt << 24);
            }
            cnt = uni(cnt + (uint32_t)__popcll(dmask));
            row_cur = uni(__shfl(row, (int)(n_valid - 1)));
        }
        __builtin_amdgcn_wave_barrier();
        // the segment's entries are contiguous from its first cell's position
        if (cnt) {
            unsigned long long qb = 0;
            if (lane == 0)
                qb = atomicAdd(&a.seg_cur[(size_t)g * kCursorWords + 1], (unsigned long long)cnt);
            const uint64_t base = sb + uni64(__shfl(qb, 0));
            for (uint32_t i = lane; i < cnt; i += kWave)
                a.queue[base + i] = stage[i];
        }
        for (uint32_t i = lane; i < rwin_len; i += kWave) {
            const uint32_t r = rwin[i];
            if (r)
                atomicAdd(&a.tally[S.tally_base + (uint64_t)(row_lo - S.row_base + i / T) * kdim * T + i % T],
                          (unsigned long long)r);
        }
        __builtin_amdgcn_wave_barrier();  // stage and window are reused by the next unit
        u = uni((uint32_t)min<uint64_t>((uint64_t)u + n_w, 0xFFFFFFFFull));
    }
    if (Batched)
        flush_counts_wave(n, a.counters, set, (blockIdx.x * waves + wave) % kCounterCopies, lane);
    else
        flush_counts(n, a.counters, lane, red);
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
    uint64_t cell;
    N node, pend;
    uint64_t tb_glob;   // device tally index of (row, generation 0, type)
    double s, mean, sd;
    double tl, tr;      // daughters' timers of the division in progress
    uint32_t k, kcap;
    uint32_t tb_lds;    // window index of (row, generation 0, type), or kNoWindow
    uint32_t redo;      // bit 0 / bit 1: left / right daughter's timer still to redraw
    uint32_t att;       // attempt number of that redraw
    bool active;
};

// ======================================================================
// tree_kernel: workgroup g works through segment g's queue entries.
// LDS: log and sin/cos tables | stream coefficients | type tables (mean, sd:
// double [T]; file: u32 [T]) | row info when it fits (u32 [n_rows]) |
// counter sums u32 [8] | victim hint u32 [4] | window tally u32 [win_rows *
// kdim * T + 1 spare] | stack double [stack_levels * blockDim]
// ======================================================================
template <bool RowsLds, typename N>
__global__ void __launch_bounds__(256, 4) tree_kernel(RunArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t bdim = blockDim.x;
    const uint32_t lane = tid & (kWave - 1);
    const uint32_t wave = tid / kWave;
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
    double *t_mean = (double *)(lds + kMathTableBytes + (sizeof(dev::Coeffs) + 15) / 16 * 16);
    double *t_sd = t_mean + T;
    uint32_t *t_file = (uint32_t *)(t_sd + T);
    uint32_t *r_info = (uint32_t *)((unsigned char *)t_mean + ((size_t)T * 20 + 15) / 16 * 16);
    uint32_t *red = (uint32_t *)((unsigned char *)r_info + ((size_t)lds_rows * 4 + 15) / 16 * 16);
    uint32_t *hint = red + kNumCounters;   // last victim segment found in this workgroup, + 1 (0: none)
    unsigned int *win = hint + 4;
    double *stack = (double *)((unsigned char *)win + ((size_t)win_len * 4 + 4 + 15) / 16 * 16);

    const uint32_t seg_lo = S.seg_lo;            // the set's segments: [seg_lo, seg_lo + n_seg)
    const uint32_t n_seg = S.seg_hi - S.seg_lo;
    const uint32_t row_lo = a.seg_row[seg0];  // first row of the tally window

    // first claim of the wavefront's own segment, issued before the set-up
    // so its latency overlaps it (lane 0 holds the result)
    unsigned long long pf_raw = 0;
    uint32_t pf_size = a.claim;   // size of the claim in flight
    uint32_t pf_pending = 1;      // a claim of pf_size on segment cur_seg is in flight
    if (lane == 0)
        pf_raw = atomicAdd(&a.seg_cur[(size_t)seg0 * kCursorWords], (unsigned long long)pf_size);
    // the segment being drained: first the own one, then victims
    uint32_t cur_seg = seg0;
    uint64_t cur_begin = uni64(a.seg_begin[seg0]);
    uint64_t cur_total = uni64(a.seg_cur[(size_t)seg0 * kCursorWords + 1]);
    if (tid == 0)
        *k_lds = dev::coeffs();
    copy_tables(m_log, m_sc, tid, bdim);
    for (uint32_t i = tid; i < T; i += bdim) {
        t_mean[i] = a.type_mean[S.type_base + i];
        t_sd[i] = a.type_sd[S.type_base + i];
        t_file[i] = a.type_file[S.type_base + i];
    }
    for (uint32_t i = tid; i < lds_rows; i += bdim)
        r_info[i] = a.row_info[i];
    for (uint32_t i = tid; i <= win_len; i += bdim)
        win[i] = 0u;  // win[win_len] absorbs the adds of lanes outside the window
    if (tid < kNumCounters + 4)
        red[tid] = 0u;  // counter sums and the victim hint
    __syncthreads();

    // the stream's FP64 coefficients, held in vector registers for the whole
    // kernel (loaded from LDS, so they are not re-materialised as scalar literals)
    const dev::Coeffs K = *k_lds;
    const dev::Tables tb{m_log, m_sc};
    const uint32_t *ri = RowsLds ? r_info : a.row_info;

    const uint32_t n_waves = bdim / kWave;
    // claim size for a segment with `rest` unclaimed entries: a share of
    // them (guided self-scheduling, so claims shrink towards the end and the
    // wavefronts finish together), between claim_min and claim
    // Claims from another workgroup's segment are capped at claim_steal:
    // a lagging segment is drained by several thieves at the end, and a
    // large share taken by one of them would leave it working alone.
    auto guided = [&](uint64_t rest) {
        const uint32_t share = (rest >> 32 ? 0xFFFFFFFFu : (uint32_t)rest) / (2u * n_waves);
        const uint32_t cap = cur_seg == seg0 ? a.claim : a.claim_steal;
        return uni(max(min(a.claim_min, cap), min(cap, share)));
    };
    uint32_t n_probe = 0;     // probes so far (varies the probed segments)
    // claimed entries of segment cur_seg not yet loaded: [q_cur, q_end)
    uint64_t q_cur = 0, q_end = 0;
    uint32_t roots_done = 0;
    // two prefetch sets of up to 64 entries, entry i of a set held by lane
    // i: hand-outs read the current set (entries pos .. pos + have - 1);
    // the incoming set (in_have entries) is loaded in one unconditional load
    // per lane and copied into the current set only when that runs empty,
    // long after its loads were issued
    uint32_t pos = 0, have = 0, in_have = 0;
    double c_s0 = 0.0, i_s0 = 0.0;
    uint32_t c_cell = 0, c_rt = 0, i_cell = 0, i_rt = 0;

    Lane<N> L;
    L.cell = L.tb_glob = 0;
    L.node = L.pend = 0;
    L.s = L.mean = L.sd = L.tl = L.tr = 0.0;
    L.k = L.kcap = 0;
    L.redo = 0;
    L.att = 1;
    L.tb_lds = kNoWindow;
    L.active = false;
    Counts n = {0, 0, 0, 0, 0};
    const uint32_t idle_limit = a.idle_limit;

    for (;;) {
        // ======== hand buffered entries to idle lanes (branch-free) ========
        if (have == 0 && in_have) {
            c_s0 = i_s0;
            c_cell = i_cell;
            c_rt = i_rt;
            pos = 0;
            have = in_have;
            in_have = 0;
        }
        uint64_t idle = __ballot(!L.active);
        uint32_t n_idle = uni((uint32_t)__popcll(idle));
        if (have && n_idle) {
            const uint32_t rank = lane_rank(idle);
            const bool take = !L.active && rank < have;
            const uint32_t src = min(pos + rank, kWave - 1);
            const double e_s0 = __shfl(c_s0, (int)src);
            const uint32_t e_cell = __shfl(c_cell, (int)src);
            const uint32_t e_rt = __shfl(c_rt, (int)src);
            const uint32_t row = e_rt & 0xFFFFFFu;
            const uint32_t jt = min(e_rt >> 24, T - 1);
            const uint32_t rr = min(row, a.n_rows - 1);
            const uint32_t tf = t_file[jt];
            const uint32_t off = row - row_lo;
            L.cell = take ? e_cell : L.cell;
            L.s = take ? e_s0 : L.s;
            L.node = take ? (N)1 : L.node;
            L.k = take ? 0u : L.k;
            L.pend = take ? (N)0 : L.pend;
            L.redo = take ? 0u : L.redo;
            L.kcap = take ? (ri[rr] & 0xFFu) : L.kcap;
            L.mean = take ? t_mean[jt] : L.mean;
            L.sd = take ? t_sd[jt] : L.sd;
            L.tb_lds = take ? (off < a.win_rows ? off * kdim * T + tf : kNoWindow) : L.tb_lds;
            L.tb_glob = take ? S.tally_base + (uint64_t)(row - S.row_base) * kdim * T + tf : L.tb_glob;
            L.active = L.active || take;
            const uint32_t taken = n_idle < have ? n_idle : have;
            pos = uni(pos + taken);
            have = uni(have - taken);
        } else if (roots_done && n_idle) {
            // ---- steal pending daughters inside the wavefront ----
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
                const uint64_t d_cell = __shfl(L.cell, (int)src);
                const uint32_t d_kcap = __shfl(L.kcap, (int)src);
                const uint32_t d_tbl = __shfl(L.tb_lds, (int)src);
                const uint64_t d_tbg = __shfl(L.tb_glob, (int)src);
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
                    L.tb_glob = d_tbg;
                    L.mean = d_mean;
                    L.sd = d_sd;
                    L.pend = 0;
                    L.redo = 0;
                    L.active = true;
                }
            }
        }
        // ======== fill the incoming prefetch set once it is empty ========
        // (straight-line in the common case, so a claim issued here is
        // defined once per iteration and not waited for before it is used)
        if (!roots_done && in_have == 0) {
            if (q_cur >= q_end) {
                // take the claim in flight (issued when the previous claim's
                // last entries were loaded)
                uint64_t base = ~0ull;
                if (pf_pending) {
                    base = uni64(__shfl(pf_raw, 0));
                    pf_pending = 0;
                }
                // the segment is drained: first try the victim another
                // wavefront of this workgroup found last (one atomic, no
                // scan); else kProbes lanes probe one segment each (spread
                // over all segments, shifted by a hash of the wavefront and
                // the attempt, so concurrent thieves look at different
                // segments) and the probe with the most entries left
                // becomes the segment to drain, published as the
                // workgroup's hint.  No entries left in any probe ends the
                // roots: a segment's own workgroup always drains it.
                uint32_t tried = 0;
                while (base >= cur_total) {
                    const uint32_t hv = uni(*hint);
                    if (hv != 0 && hv - 1 != cur_seg && hv != tried) {
                        tried = hv;
                        cur_seg = hv - 1;
                        cur_begin = uni64(a.seg_begin[cur_seg]);
                        cur_total = uni64(a.seg_cur[(size_t)cur_seg * kCursorWords + 1]);
                        pf_size = uni(min(a.claim_steal, a.claim));
                        unsigned long long r = 0;
                        if (lane == 0)
                            r = atomicAdd(&a.seg_cur[(size_t)cur_seg * kCursorWords], (unsigned long long)pf_size);
                        base = uni64(__shfl(r, 0));
                        continue;
                    }
                    const uint32_t h = hash32((seg0 * n_waves + wave) * 0x9E3779B9u + n_probe * 0x85EBCA6Bu);
                    n_probe = uni(n_probe + 1);
                    const uint32_t stride = n_seg > kProbes ? n_seg / kProbes : 1u;
                    const uint32_t cand = seg_lo + (h + lane * stride + (h >> 16) % stride) % n_seg;
                    uint64_t left = 0;
                    if (lane < kProbes) {
                        const uint64_t tot = a.seg_cur[(size_t)cand * kCursorWords + 1];
                        const uint64_t used = __hip_atomic_load(&a.seg_cur[(size_t)cand * kCursorWords],
                                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        left = used < tot ? tot - used : 0;
                    }
                    uint32_t best = cand;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) {
                        const uint64_t ol = __shfl_xor(left, off);
                        const uint32_t ob = __shfl_xor(best, off);
                        if (ol > left || (ol == left && ob < best)) {
                            left = ol;
                            best = ob;
                        }
                    }
                    if (uni64(left) == 0) {
                        roots_done = 1;
                        break;
                    }
                    cur_seg = uni(best);
                    if (lane == 0)
                        *hint = cur_seg + 1;
                    cur_begin = uni64(a.seg_begin[cur_seg]);
                    cur_total = uni64(a.seg_cur[(size_t)cur_seg * kCursorWords + 1]);
                    pf_size = guided(uni64(left));
                    unsigned long long r = 0;
                    if (lane == 0)
                        r = atomicAdd(&a.seg_cur[(size_t)cur_seg * kCursorWords], (unsigned long long)pf_size);
                    base = uni64(__shfl(r, 0));
                }
                if (!roots_done) {
                    q_cur = uni64(cur_begin + base);
                    q_end = uni64(cur_begin + min(base + pf_size, cur_total));
                }
            }
            if (!roots_done) {
                // load up to 64 claimed entries (lanes past the last one load
                // it again; their copies are never handed out)
                const uint32_t nf = uni((uint32_t)min<uint64_t>(kWave, q_end - q_cur));
                const QueueEntry e = a.queue[q_cur + min(lane, nf - 1)];
                i_s0 = e.s0;
                i_cell = e.cell;
                i_rt = e.row_type;
                in_have = nf;
                q_cur = uni64(q_cur + nf);
                // the claim's last entries are loaded: issue the next claim
                // on the same segment now, so its latency overlaps the steps,
                // sized by what the segment has left (guided self-scheduling:
                // claims shrink towards the end, so the wavefronts finish
                // together)
                const uint64_t rest = cur_begin + cur_total - q_end;
                if (q_cur >= q_end && rest) {
                    pf_size = guided(rest);
                    pf_pending = 1;
                    if (lane == 0)
                        pf_raw = atomicAdd(&a.seg_cur[(size_t)cur_seg * kCursorWords], (unsigned long long)pf_size);
                }
            }
        }

        if (__ballot(L.active) == 0) {
            if (roots_done && have == 0 && in_have == 0)
                break;
            continue;
        }

        // ======== division steps until enough lanes are idle ========
        // Straight-line code: every outcome is a select, so the step costs a
        // handful of scalar instructions (one scalar unit serves the whole
        // CU).  Idle lanes run it too with their results discarded.
        do {
            // One Philox block + Box-Muller pair per lane and step: either the
            // pair block of a new division (timers of both daughters) or one
            // block of a daughter's truncated-normal redraw stream
            // (cell.cu:106-122); a lane with a redraw pending does not advance
            // its tree in that step.  Flags are 0/1 integers so the step's
            // logic stays in vector instructions.
            const uint32_t act = L.active ? 1u : 0u;
            double tl, tr;
            uint32_t redo;
            if (__builtin_expect(__ballot(L.redo != 0) == 0, 1)) {
                // no redraw pending in the wavefront (the common case): every
                // busy lane starts a new division with its pair block
                const Words w = dev::draw_block(seed, L.cell, L.node, 0u);
                double z0, z1;
                dev::normal_pair(w, z0, z1, K, tb);
                tl = fma(z0, L.sd, L.mean);
                tr = fma(z1, L.sd, L.mean);
                const uint32_t r0 = tl > 0.0 ? 0u : 1u;
                const uint32_t r1 = tr > 0.0 ? 0u : 1u;
                redo = r0 | (r1 << 1);
                n.div += act;
                n.rej += act * (r0 + r1);
                L.att = 1u;
            } else {
                const uint32_t fresh = L.redo == 0 ? 1u : 0u;
                const uint32_t rright = (L.redo & 1u) ? 0u : 1u;   // redraw the right daughter?
                const N bnode = fresh ? L.node : (L.node << 1) | (N)rright;
                const Words w = dev::draw_block(seed, L.cell, bnode, fresh ? 0u : L.att);
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
                redo = fresh ? (r0 | (r1 << 1)) : (done ? L.redo & ~rbit : L.redo);
                n.div += act & fresh;
                n.rej += act * (fresh ? r0 + r1 : r0 + (r0 & r1));
                L.att = fresh | done ? 1u : L.att + 1u;
            }
            L.tl = tl;
            L.tr = tr;
            L.redo = act ? redo : 0u;
            const uint32_t cmp = act & (redo == 0 ? 1u : 0u);  // division complete this step

            // daughters (proliferation.cu:317-381): out of time -> tallied,
            // f/2 <= phi -> dropped, else divides later
            const double sl = L.s + tl;
            const double sr = L.s + tr;
            const uint32_t k1 = L.k + 1;
            const uint32_t rec_l = sl > t_max ? 1u : 0u;
            const uint32_t rec_r = sr > t_max ? 1u : 0u;
            const uint32_t can = k1 < L.kcap ? 1u : 0u;
            const uint32_t div_l = (rec_l ^ 1u) & can;
            const uint32_t div_r = (rec_r ^ 1u) & can;
            const uint32_t nrec = cmp * (rec_l + rec_r);
            const bool in_win = L.tb_lds != kNoWindow;
            // lanes without a completed division add 0 to the spare slot win[win_len]
            atomicAdd(&win[(cmp && in_win) ? L.tb_lds + k1 * T : win_len], nrec);
            if (__builtin_expect(__ballot(!in_win && nrec != 0) != 0, 0)) {
                if (!in_win && nrec != 0)
                    atomicAdd(&a.tally[L.tb_glob + (uint64_t)k1 * T], (unsigned long long)nrec);
            }
            // right daughter's completion time; level k1 holds no pending
            // entry of an active lane (idle lanes have none at all), so the
            // store is harmless when it is not pushed
            stack[(size_t)(act ? k1 - 1 : 0u) * bdim + tid] = sr;
            const uint32_t go_l = cmp & div_l;
            const uint32_t go_r = cmp & (div_l ^ 1u) & div_r;
            const uint32_t pop = cmp & (div_l ^ 1u) & (div_r ^ 1u);
            // (k1 stays below the width of N for a lane that pushes; the
            // mask keeps the shift defined for the others)
            const N pend2 = L.pend | ((N)(go_l & div_r) << (k1 & (sizeof(N) * 8 - 1)));
            const uint32_t has = pend2 != 0 ? 1u : 0u;
            const uint32_t lvl = has ? top_bit(pend2 | (N)1) : 1u;
            const double sp = stack[(size_t)(lvl - 1) * bdim + tid];
            // single-level selects only (nested ones turn into branches)
            const uint32_t down = go_l | go_r;
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
            L.active = act && !(pop && !has);
        } while (uni((uint32_t)__popcll(__ballot(!L.active))) <= idle_limit);
    }

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

size_t root_lds_bytes(uint32_t n_types, uint32_t block_size, uint32_t lds_rows)
{
    return kMathTableBytes + ((size_t)n_types * 28 + 15) / 16 * 16 + ((size_t)(2 * lds_rows + 1) * 4 + 15) / 16 * 16 +
           32 + (size_t)(block_size / kWave) * root_wave_lds_bytes(n_types);
}

size_t tree_lds_bytes(uint32_t win_rows, uint32_t kdim, uint32_t n_types, uint32_t stack_levels,
                      uint32_t block_size, uint32_t lds_rows)
{
    const size_t tables = kMathTableBytes + (sizeof(dev::Coeffs) + 15) / 16 * 16 + ((size_t)n_types * 20 + 15) / 16 * 16 +
                          ((size_t)lds_rows * 4 + 15) / 16 * 16 + (kNumCounters + 4) * 4;
    const size_t win = ((size_t)win_rows * kdim * n_types * 4 + 4 + 15) / 16 * 16;
    const size_t stk = (size_t)stack_levels * block_size * sizeof(double);
    return tables + win + stk;
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

hipError_t launch_root(const RunArgs &a, const LaunchShape &ls, hipStream_t st)
{
    if (a.n_cells == 0)
        return hipSuccess;
    const dim3 g((unsigned)ls.root_blocks), b(ls.root_block_size);
    const bool batched = a.n_sets > 1;
    if (a.rows_in_lds && !batched)
        hipLaunchKernelGGL((root_kernel<true, false>), g, b, ls.root_lds_bytes, st, a);
    else if (!batched)
        hipLaunchKernelGGL((root_kernel<false, false>), g, b, ls.root_lds_bytes, st, a);
    else if (a.rows_in_lds)
        hipLaunchKernelGGL((root_kernel<true, true>), g, b, ls.root_lds_bytes, st, a);
    else
        hipLaunchKernelGGL((root_kernel<false, true>), g, b, ls.root_lds_bytes, st, a);
    return hipGetLastError();
}

hipError_t launch_tree(const RunArgs &a, const LaunchShape &ls, hipStream_t st)
{
    if (a.n_cells == 0)
        return hipSuccess;
    const dim3 g(ls.tree_blocks), b(ls.tree_block_size);
    const bool narrow = a.kdim <= 32;  // heap indices and level masks fit 32 bits
    if (a.rows_in_lds && narrow)
        hipLaunchKernelGGL((tree_kernel<true, uint32_t>), g, b, ls.tree_lds_bytes, st, a);
    else if (a.rows_in_lds)
        hipLaunchKernelGGL((tree_kernel<true, uint64_t>), g, b, ls.tree_lds_bytes, st, a);
    else if (narrow)
        hipLaunchKernelGGL((tree_kernel<false, uint32_t>), g, b, ls.tree_lds_bytes, st, a);
    else
        hipLaunchKernelGGL((tree_kernel<false, uint64_t>), g, b, ls.tree_lds_bytes, st, a);
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

hipError_t tree_occupancy(int block_size, size_t lds_bytes, bool rows_in_lds, uint32_t kdim, int *blocks_per_cu)
{
    const bool narrow = kdim <= 32;
    if (rows_in_lds && narrow)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tree_kernel<true, uint32_t>, block_size,
                                                            lds_bytes);
    if (rows_in_lds)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tree_kernel<true, uint64_t>, block_size,
                                                            lds_bytes);
    if (narrow)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tree_kernel<false, uint32_t>, block_size,
                                                            lds_bytes);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tree_kernel<false, uint64_t>, block_size,
                                                        lds_bytes);
}

hipError_t root_occupancy(int block_size, size_t lds_bytes, bool rows_in_lds, int *blocks_per_cu)
{
    return rows_in_lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, root_kernel<true, false>, block_size,
                                                                      lds_bytes)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, root_kernel<false, false>, block_size,
                                                                      lds_bytes);
}

}  // namespace procell
