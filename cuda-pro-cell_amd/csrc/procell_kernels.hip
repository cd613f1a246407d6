// procell_kernels.hip -- gfx950 kernels of the plain proliferation method.
//
// The reference materialises every tree level as a dense array of 40-byte
// cells and recurses with device-side launches (src/simulation/
// proliferation.cu:289-402).  Here the work is split into two kernels and no
// cell is ever stored:
//
//  roots_kernel  one lane per initial cell (cell.cu:25-143,
//                cells_population.cu:97-117): type draw, truncated first
//                timer, phase; tallies cells that never divide and appends
//                the others (16 B each) to a work queue.
//  tree_kernel   persistent: each lane walks one initial cell's division tree
//                depth first (proliferation.cu:317-381).  A step is one
//                division: one Philox block and one Box-Muller pair give both
//                daughters' timers; daughters that will not divide are
//                tallied or dropped on the spot, so the lane's next node is
//                always a dividing one (the left daughter, else the right,
//                else the deepest pending right daughter from the lane's
//                column of an LDS stack, else a fresh queue entry).  Wavefronts
//                claim queue entries in chunks with one atomic.
//
// Tallies are int64 counts by (nonzero row, generation, type); the host maps
// (row, generation) to the output value row * 2^-generation.
#include "procell_internal.hpp"
#include "procell_stream.hpp"

namespace procell {
namespace {

using dev::Words;

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

__device__ __forceinline__ uint32_t find_row(const uint64_t *row_start, uint32_t n_rows, uint64_t cell)
{
    uint32_t lo = 0, hi = n_rows;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (row_start[mid] <= cell)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Redraw candidates of a node's truncated-normal stream from attempt `first`
// until one timer is > 0 (cell.cu:106-122 redraws while the timer is <= 0).
// Every rejected candidate is counted.
__device__ __forceinline__ double truncated_timer(uint64_t seed, uint64_t cell, uint64_t node,
                                                  double mean, double sd, uint32_t &rejected)
{
    for (uint32_t att = 1; att < 65536u; ++att) {
        const Words w = dev::draw_block(seed, cell, node, att);
        double z0, z1;
        dev::normal_pair(w, z0, z1);
        const double c0 = fma(z0, sd, mean);
        if (c0 > 0.0)
            return c0;
        ++rejected;
        const double c1 = fma(z1, sd, mean);
        if (c1 > 0.0)
            return c1;
        ++rejected;
    }
    return mean;
}

// ---------------------------------------------------------------------------
// Zero fill (used instead of hipMemsetAsync so that runs can be captured).
// ---------------------------------------------------------------------------
__global__ void zero_kernel(unsigned long long *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Initial cells.  Tiles of blockDim consecutive cells; cells of a tile span
// few rows, so generation-0 tallies go through an LDS window of `win_rows`
// rows starting at the tile's first row (global atomics beyond it).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) roots_kernel(RunArgs a, uint32_t win_rows)
{
    extern __shared__ unsigned int win[];
    const uint32_t tid = threadIdx.x;
    const uint32_t bdim = blockDim.x;
    const uint32_t T = a.n_types;
    const uint32_t win_len = win_rows * T;
    const uint64_t n_tiles = (a.n_cells + bdim - 1) / bdim;
    uint32_t n_rec = 0, n_drop = 0, n_lost = 0, n_rej = 0, n_cells = 0;

    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t c0 = tile * bdim;
        const uint32_t row0 = find_row(a.row_start, a.n_rows, c0);
        for (uint32_t i = tid; i < win_len; i += bdim)
            win[i] = 0u;
        __syncthreads();

        const uint64_t cell = c0 + tid;
        bool divides = false;
        double s0 = 0.0;
        uint32_t row = 0, jt = 0;
        if (cell < a.n_cells) {
            ++n_cells;
            row = find_row(a.row_start, a.n_rows, cell);
            const uint32_t info = a.row_info[row];
            const uint32_t kcap = info & 0xFFu;
            const bool ge_phi = (info >> 8) & 1u;
            const Words w = dev::draw_block(a.seed, cell, 0, 0);
            const double ut = dev::uniform_co(w.w0, w.w1);
            const double U = dev::uniform_oc(w.w2, w.w3);
            while (jt + 1 < T && !(ut < a.type_cum[jt]))
                ++jt;
            const double mean = a.type_mean[jt];
            bool record = false;
            if (mean < 0.0) {
                record = true;
            } else {
                const double tau0 = truncated_timer(a.seed, cell, 1, mean, a.type_sd[jt], n_rej);
                const double t0 = tau0 * U;
                s0 = t0 + tau0;
                if (s0 > a.t_max)
                    record = true;
                else if (kcap == 0)
                    ++n_drop;
                else
                    divides = true;
            }
            if (record) {
                if (ge_phi) {
                    ++n_rec;
                    const uint32_t tf = a.type_file[jt];
                    const uint32_t off = row - row0;
                    if (off < win_rows)
                        atomicAdd(&win[off * T + tf], 1u);
                    else
                        atomicAdd(&a.tally[(uint64_t)row * a.kdim * T + tf], 1ull);
                } else {
                    ++n_lost;
                }
            }
        }
        // append dividing cells to the queue: one atomic per wavefront
        const uint64_t dmask = __ballot(divides);
        if (dmask) {
            unsigned long long base = 0;
            if (__lane_id() == 0)
                base = atomicAdd(&a.counters[kPoolCount], (unsigned long long)__popcll(dmask));
            base = __shfl(base, 0);
            if (divides) {
                PoolEntry e;
                e.s0 = s0;
                e.cell = (uint32_t)cell;
                e.row_type = row | (jt << 24);
                a.pool[base + lane_rank(dmask)] = e;
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < win_len; i += bdim) {
            const uint32_t c = win[i];
            if (c) {
                const uint32_t r = row0 + i / T;
                atomicAdd(&a.tally[(uint64_t)r * a.kdim * T + (i % T)], (unsigned long long)c);
            }
        }
        __syncthreads();
    }

    n_rec = wave_sum(n_rec);
    n_drop = wave_sum(n_drop);
    n_lost = wave_sum(n_lost);
    n_rej = wave_sum(n_rej);
    n_cells = wave_sum(n_cells);
    if (__lane_id() == 0) {
        if (n_rec) atomicAdd(&a.counters[kRecorded], (unsigned long long)n_rec);
        if (n_drop) atomicAdd(&a.counters[kDropped], (unsigned long long)n_drop);
        if (n_lost) atomicAdd(&a.counters[kLost], (unsigned long long)n_lost);
        if (n_rej) atomicAdd(&a.counters[kRejected], (unsigned long long)n_rej);
        if (n_cells) atomicAdd(&a.counters[kCells], (unsigned long long)n_cells);
    }
}

// ---------------------------------------------------------------------------
// Division trees, depth first, one tree per lane at a time.
// LDS: stack[level * blockDim + tid] = completion time of the pending right
// daughter at `level` (bit `level` of the lane's `pend` mask).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) tree_kernel(RunArgs a, uint32_t chunk)
{
    extern __shared__ double stack[];
    const uint32_t tid = threadIdx.x;
    const uint32_t bdim = blockDim.x;
    const uint32_t T = a.n_types;
    const uint64_t pool_n = __atomic_load_n(&a.counters[kPoolCount], __ATOMIC_RELAXED);

    uint64_t q_cur = 0, q_end = 0;
    bool q_done = false;

    bool active = false;
    uint64_t cell = 0, node = 0, pend = 0;
    uint32_t k = 0, kcap = 0;
    uint64_t tally_base = 0;  // (row * kdim) * T + type_file
    double s = 0.0, mean = 0.0, sd = 0.0;
    uint32_t n_div = 0, n_rec = 0, n_drop = 0, n_rej = 0;

    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (idle != 0 && !q_done) {
            if (q_cur >= q_end) {
                unsigned long long base = 0;
                if (__lane_id() == 0)
                    base = atomicAdd(&a.counters[kPoolNext], (unsigned long long)chunk);
                base = __shfl(base, 0);
                if (base >= pool_n) {
                    q_done = true;
                } else {
                    q_cur = base;
                    q_end = base + chunk < pool_n ? base + chunk : pool_n;
                }
            }
            if (!q_done) {
                const uint64_t avail = q_end - q_cur;
                const uint32_t rank = lane_rank(idle);
                if (!active && rank < avail) {
                    const PoolEntry e = a.pool[q_cur + rank];
                    const uint32_t row = e.row_type & 0xFFFFFFu;
                    const uint32_t jt = e.row_type >> 24;
                    cell = e.cell;
                    s = e.s0;
                    node = 1;
                    k = 0;
                    pend = 0;
                    kcap = a.row_info[row] & 0xFFu;
                    mean = a.type_mean[jt];
                    sd = a.type_sd[jt];
                    tally_base = (uint64_t)row * a.kdim * T + a.type_file[jt];
                    active = true;
                }
                const uint64_t taken = (uint64_t)__popcll(idle);
                q_cur += taken < avail ? taken : avail;
            }
        }
        if (__ballot(active) == 0)
            break;
        if (!active)
            continue;

        // one division: both daughters' timers from one block
        ++n_div;
        double tl, tr;
        {
            const Words w = dev::draw_block(a.seed, cell, node, 0);
            double z0, z1;
            dev::normal_pair(w, z0, z1);
            tl = fma(z0, sd, mean);
            tr = fma(z1, sd, mean);
        }
        if (!(tl > 0.0)) {
            ++n_rej;
            tl = truncated_timer(a.seed, cell, 2 * node, mean, sd, n_rej);
        }
        if (!(tr > 0.0)) {
            ++n_rej;
            tr = truncated_timer(a.seed, cell, 2 * node + 1, mean, sd, n_rej);
        }
        const double sl = s + tl;
        const double sr = s + tr;
        const uint32_t k1 = k + 1;
        const bool rec_l = sl > a.t_max;
        const bool rec_r = sr > a.t_max;
        const bool div_l = !rec_l && k1 < kcap;
        const bool div_r = !rec_r && k1 < kcap;
        const uint32_t nrec = (uint32_t)rec_l + (uint32_t)rec_r;
        n_drop += (uint32_t)(!rec_l && !div_l) + (uint32_t)(!rec_r && !div_r);
        if (nrec) {
            n_rec += nrec;
            atomicAdd(&a.tally[tally_base + (uint64_t)k1 * T], (unsigned long long)nrec);
        }
        if (div_l) {
            if (div_r) {
                stack[k1 * bdim + tid] = sr;
                pend |= 1ull << k1;
            }
            node = 2 * node;
            k = k1;
            s = sl;
        } else if (div_r) {
            node = 2 * node + 1;
            k = k1;
            s = sr;
        } else if (pend) {
            const uint32_t lvl = 63u - (uint32_t)__builtin_clzll(pend);
            pend &= ~(1ull << lvl);
            s = stack[lvl * bdim + tid];
            node = (node >> (k - lvl)) | 1ull;
            k = lvl;
        } else {
            active = false;
        }
    }

    n_div = wave_sum(n_div);
    n_rec = wave_sum(n_rec);
    n_drop = wave_sum(n_drop);
    n_rej = wave_sum(n_rej);
    if (__lane_id() == 0) {
        if (n_div) atomicAdd(&a.counters[kDivisions], (unsigned long long)n_div);
        if (n_rec) atomicAdd(&a.counters[kRecorded], (unsigned long long)n_rec);
        if (n_drop) atomicAdd(&a.counters[kDropped], (unsigned long long)n_drop);
        if (n_rej) atomicAdd(&a.counters[kRejected], (unsigned long long)n_rej);
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
    dev::normal_pair(w, z0, z1);
    z[2 * i] = z0;
    z[2 * i + 1] = z1;
}

}  // namespace

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

hipError_t launch_roots(const RunArgs &a, const LaunchShape &ls, hipStream_t st)
{
    if (a.n_cells == 0)
        return hipSuccess;
    const size_t lds = (size_t)ls.roots_window_rows * a.n_types * sizeof(unsigned int);
    hipLaunchKernelGGL(roots_kernel, dim3(ls.roots_blocks), dim3(ls.roots_block_size), lds, st, a,
                       ls.roots_window_rows);
    return hipGetLastError();
}

hipError_t launch_tree(const RunArgs &a, const LaunchShape &ls, hipStream_t st)
{
    if (a.n_cells == 0)
        return hipSuccess;
    hipLaunchKernelGGL(tree_kernel, dim3(ls.tree_blocks), dim3(ls.tree_block_size), ls.tree_lds_bytes,
                       st, a, ls.chunk);
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

hipError_t tree_occupancy(int block_size, size_t lds_bytes, int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, tree_kernel, block_size, lds_bytes);
}

hipError_t roots_occupancy(int block_size, size_t lds_bytes, int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, roots_kernel, block_size, lds_bytes);
}

}  // namespace procell
