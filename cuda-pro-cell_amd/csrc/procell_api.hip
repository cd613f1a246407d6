// procell_api.hip -- host side of libprocell: validation, problem set-up,
// device buffers, launches and the merge of tallies into the output
// histogram.  Replaces the host halves of create_cells_population
// (src/simulation/cells_population.cu:18-54) and proliferate
// (src/simulation/proliferation.cu:26-215), and the bin-set construction of
// load_fluorescences (src/io/parser.cu:68-154).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "procell_internal.hpp"

namespace procell {

}  // namespace procell

using namespace procell;

// ---------------------------------------------------------------------------
// Plan: device buffers for one problem (or a batch of parameter sets), reused
// across runs.
// ---------------------------------------------------------------------------
constexpr uint32_t kSeedSlots = 4;  // pinned host staging slots for batch seeds
// Segments per resident workgroup unless tuned.  1e7-cell histogram, 3
// workgroups of 256 per CU, segment sizes decreasing geometrically: 1 ->
// 0.321 ms per run, 2 -> 0.289, 3 -> 0.294, 4 -> 0.309, 8 -> 0.340 (more
// segments: more set-up and window flushes).  A batch of many small sets
// balances better with finer segments.
constexpr uint32_t kSegmentsPerWorkgroup = 2;
constexpr uint32_t kBatchSegmentsPerWorkgroup = 8;
// most lane-stack levels a batch keeps in LDS (a 256-lane workgroup: 2 KB per level)
constexpr uint32_t kLdsStackLevels = 11;

struct procell_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool timing = false;
    std::vector<int32_t> track_ratio;  // per set
    std::vector<Setup> setups;         // per set
    std::vector<size_t> tally_base, tally_len;  // per set, in tally entries
    LaunchShape shape{};
    RunArgs args{};
    void *d_block = nullptr;  // one allocation holding every table
    bool own_block = true;    // false: d_block is the static scratch array
    size_t tally_total = 0;
    size_t zero_words = 0;    // counters and tally, cleared per run
    uint64_t *d_seeds = nullptr;       // [n_sets] seeds of the current batch run (in d_block)
    uint64_t *h_seeds = nullptr;       // pinned [kSeedSlots][n_sets]
    hipEvent_t seed_ev[kSeedSlots] = {};
    hipGraph_t graph = nullptr;        // batch runs: clearing, root and tree kernels
    hipGraphExec_t graph_exec = nullptr;
    std::vector<hipEvent_t> kev;   // PROCELL_FLAG_TIMING: start/stop of each run's kernels (ring)
    uint64_t runs = 0;             // runs since the last kernel-time read
    size_t last_slot = 0;          // event slot of the last timed run
    bool timed = false;            // a run has recorded its events
    uint64_t batch_runs = 0;
    bool ran = false;
    // copies of the inputs (the autotuner re-creates the layout from them)
    std::vector<procell_params> params;
    std::vector<std::vector<procell_row>> rows;
    std::vector<std::vector<procell_type>> types;
};

constexpr uint32_t kTimedRuns = 256;

// Device memory of procell_simulate_noalloc, one copy per GPU (see
// scratch_address below for how host code reaches it).
constexpr size_t kScratchBytes = PROCELL_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local hipError_t g_last_hip = hipSuccess;

#define HIP_TRY(expr)                        \
    do {                                     \
        hipError_t e_ = (expr);              \
        if (e_ != hipSuccess) {              \
            g_last_hip = e_;                 \
            return PROCELL_E_HIP;            \
        }                                    \
    } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int check_device(int device, int *n_cu, size_t *lds_per_cu = nullptr, size_t *lds_per_block = nullptr)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
        return PROCELL_E_NO_DEVICE;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return PROCELL_E_NO_DEVICE;
    *n_cu = prop.multiProcessorCount;
    if (lds_per_cu)
        *lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
    if (lds_per_block)
        *lds_per_block = prop.sharedMemPerBlock ? prop.sharedMemPerBlock : 65536;
    return PROCELL_OK;
}

void destroy_plan(procell_plan *pl)
{
    if (!pl)
        return;
    (void)hipSetDevice(pl->device);
    if (pl->graph_exec)
        (void)hipGraphExecDestroy(pl->graph_exec);
    if (pl->graph)
        (void)hipGraphDestroy(pl->graph);
    for (hipEvent_t &e : pl->kev)
        if (e)
            (void)hipEventDestroy(e);
    for (hipEvent_t &e : pl->seed_ev)
        if (e)
            (void)hipEventDestroy(e);
    if (pl->h_seeds)
        (void)hipHostFree(pl->h_seeds);
    if (pl->d_block && pl->own_block)
        (void)hipFree(pl->d_block);
    if (pl->own_stream && pl->stream)
        (void)hipStreamDestroy(pl->stream);
    delete pl;
}

// A plan over n_sets parameter sets (1: a plain run).  Rows, cells, type
// entries, tally entries and segments of the sets are concatenated in set
// order; the device code finds a set's values through its SetInfo.
// With `arena` (device memory of `arena_bytes`, already on ps[0].device) the
// plan's device block is carved from it instead of allocated, no stream or
// event is created (the plan runs on ps[0].stream, NULL: the default stream)
// and timing is off.
int create_plan(const procell_params *ps, uint32_t n_sets, procell_plan **out, void *arena = nullptr,
                size_t arena_bytes = 0)
{
    if (!ps || !out || n_sets == 0)
        return PROCELL_E_INVALID;
    *out = nullptr;
    const procell_tuning &tu = ps[0].tuning;
    if (tu.block_size && (tu.block_size % 64 || tu.block_size > 256))
        return PROCELL_E_INVALID;
    for (uint32_t i = 1; i < n_sets; ++i)
        if (ps[i].device != ps[0].device)
            return PROCELL_E_INVALID;
    procell_plan *pl = new (std::nothrow) procell_plan();
    if (!pl)
        return PROCELL_E_OUT_OF_MEMORY;
    pl->setups.resize(n_sets);
    pl->track_ratio.resize(n_sets);
    for (uint32_t i = 0; i < n_sets; ++i) {
        const int rc = build_setup(&ps[i], pl->setups[i], true);
        if (rc != PROCELL_OK) {
            delete pl;
            return rc;
        }
        pl->track_ratio[i] = ps[i].track_ratio;
    }
    pl->params.assign(ps, ps + n_sets);
    pl->rows.resize(n_sets);
    pl->types.resize(n_sets);
    for (uint32_t i = 0; i < n_sets; ++i) {
        pl->rows[i].assign(ps[i].rows, ps[i].rows + ps[i].n_rows);
        pl->types[i].assign(ps[i].types, ps[i].types + ps[i].n_types);
        pl->params[i].rows = pl->rows[i].data();
        pl->params[i].types = pl->types[i].data();
    }
    int n_cu = 0;
    size_t lds_per_cu = 0, lds_per_block = 0;
    {
        const int rc = check_device(ps[0].device, &n_cu, &lds_per_cu, &lds_per_block);
        if (rc != PROCELL_OK) {
            delete pl;
            return rc;
        }
    }
    pl->device = ps[0].device;
    pl->timing = !arena && (tu.flags & PROCELL_FLAG_TIMING) != 0;
    if (arena && n_sets != 1) {
        delete pl;
        return PROCELL_E_INVALID;
    }
    auto fail = [&](int code) {
        destroy_plan(pl);
        return code;
    };

    // ---- concatenation of the sets ----
    std::vector<SetInfo> info(n_sets);
    uint64_t n_cells = 0, n_rows64 = 0, n_types_tot = 0;
    uint32_t T = 1, kdim = 1;
    pl->tally_base.resize(n_sets);
    pl->tally_len.resize(n_sets);
    for (uint32_t i = 0; i < n_sets; ++i) {
        const Setup &s = pl->setups[i];
        const uint32_t Ti = s.n_types ? s.n_types : 1;
        const uint32_t ki = (uint32_t)s.kmax + 1;
        SetInfo &si = info[i];
        memset(&si, 0, sizeof si);
        si.t_max = s.t_max;
        si.cell_base = n_cells;
        si.tally_base = pl->tally_total;
        si.kdim = ki;
        si.n_types = Ti;
        si.type_base = (uint32_t)n_types_tot;
        si.row_base = (uint32_t)n_rows64;
        pl->tally_base[i] = pl->tally_total;
        pl->tally_len[i] = (size_t)s.v.size() * ki * Ti;
        pl->tally_total += pl->tally_len[i];
        n_cells += s.n_cells;
        n_rows64 += s.v.size();
        n_types_tot += Ti;
        T = std::max(T, Ti);
        kdim = std::max(kdim, ki);
    }
    if (n_cells > kMaxCells || n_rows64 > kMaxRows)
        return fail(PROCELL_E_TOO_MANY_CELLS);
    const uint32_t n_rows = (uint32_t)n_rows64;

    // ---- launch shape: block size, resident workgroups, LDS split ----
    LaunchShape &ls = pl->shape;
    RunArgs &a = pl->args;
    a.stack_levels = kdim > 3 ? kdim - 1 : 2;  // levels 1 .. kmax, at least 2 (idle lanes store at level 2)
    // A batch whose deepest set needs more than kLdsStackLevels stack levels
    // keeps the lanes' stacks in device memory: in LDS they would be sized for
    // that set and leave no room for the tally window at more than two
    // workgroups per CU (512-set batch of 128-1024 bins, up to 17 generations:
    // 2.65 ms per run at 2 workgroups per CU, 5.9 ms at 3 or 4 without a window).
    const bool gstack = n_sets > 1 && a.stack_levels > kLdsStackLevels;
    const uint32_t lds_levels = gstack ? 0u : a.stack_levels;
    // row tables (8 bytes per row) go to LDS when they are small and do not
    // cost a resident workgroup (the root steps read them when a wavefront's
    // rows advance, so from device memory they cost little)
    uint32_t lds_rows = (size_t)(2 * n_rows + 1) * 4 <= 16384 ? n_rows : 0u;
    int bs = tu.block_size ? (int)tu.block_size : 256;
    int occ = 0;
    for (;;) {
        const size_t base = sim_lds_bytes(0, kdim, T, lds_levels, (uint32_t)bs, lds_rows);
        int occ_v = 0;
        if (sim_occupancy(bs, base, lds_rows != 0, kdim, gstack, &occ_v) != hipSuccess)
            return fail(PROCELL_E_HIP);
        if (occ_v >= 1) {
            occ = occ_v;
            break;
        }
        if (bs == 64)
            return fail(PROCELL_E_DEPTH);
        bs -= 64;  // deep trees: a smaller workgroup keeps the stack within LDS
    }
    // by default at most 16 resident wavefronts per CU (4 per SIMD, the
    // kernel's register budget): 4 workgroups of 256 per CU against 3 measured
    // 0.2592 -> 0.2470 ms on the 1e7-cell histogram workload (more waves hide
    // the division step's dependency waits; the smaller tally windows cost
    // less than that)
    const int want = tu.blocks_per_cu ? (int)tu.blocks_per_cu : std::max(1, 16 * 64 / bs);
    if (lds_rows && occ < want) {
        int occ_g = 0;
        if (sim_occupancy(bs, sim_lds_bytes(0, kdim, T, lds_levels, (uint32_t)bs, 0), false, kdim, gstack, &occ_g) !=
            hipSuccess)
            return fail(PROCELL_E_HIP);
        if (occ_g > occ) {
            lds_rows = 0;
            occ = occ_g;
        }
    }
    a.rows_in_lds = lds_rows ? 1u : 0u;
    occ = std::min<int>(occ, want);
    ls.block_size = bs;
    const size_t base = sim_lds_bytes(0, kdim, T, lds_levels, (uint32_t)bs, lds_rows);
    const size_t budget = std::min(lds_per_cu / (size_t)occ, lds_per_block);
    const size_t row_bytes = (size_t)kdim * T * 4;
    uint32_t win_rows = budget > base + 16 ? (uint32_t)((budget - base - 16) / row_bytes) : 0;

    // ---- segments: one per workgroup, each within one set ----
    // `per_wg` segments per resident workgroup in total, of about equal
    // expected work: workgroups beyond the resident ones start as others
    // finish, so the dispatcher balances the segments over the CUs.  In a
    // batch each set gets segments in proportion to its expected work (at
    // least one).
    const uint32_t resident = (uint32_t)((int64_t)n_cu * occ);
    const uint32_t per_wg = tu.segments ? std::min<uint32_t>(tu.segments, 64)
                                        : (n_sets == 1 ? kSegmentsPerWorkgroup : kBatchSegmentsPerWorkgroup);
    std::vector<uint32_t> g_set(n_sets, 1);
    if (n_sets == 1) {
        g_set[0] = (uint32_t)std::min<uint64_t>((uint64_t)resident * per_wg,
                                                std::max<uint64_t>(1, (pl->setups[0].n_cells + 63) / 64));
    } else {
        std::vector<double> w(n_sets);
        double wt = 0.0;
        for (uint32_t i = 0; i < n_sets; ++i) {
            w[i] = set_cost(pl->setups[i]);
            wt += w[i];
        }
        const double target = (double)per_wg * (double)resident;
        for (uint32_t i = 0; i < n_sets; ++i)
            g_set[i] = wt > 0.0 ? (uint32_t)std::max(1.0, std::floor(w[i] / wt * target + 0.5)) : 1u;
    }
    // Segment sizes: a single set's decrease geometrically in launch order
    // (rounds of `resident` segments, each round's segments a factor q
    // smaller, the last ones 1/8 of the first), so the last workgroups
    // dispatched are short and the CUs finish together; a batch's sets are
    // split evenly and all segments are launched largest first.
    // Within a round, segment k goes to CU slot (k mod resident) / n_cu: the
    // dispatcher fills every CU's first slot, then its second, and the
    // instruction arbiter favours the older workgroups of a CU (measured with
    // three per CU: 1175, 1570 and 2050 cycles per division step for the
    // oldest, middle and youngest), so a slot's segments are sized by its
    // issue rate, 1.32 per slot of age (1e7-cell histogram workload: 0.2617
    // -> 0.2587 ms and 0.2644 -> 0.2603 ms on two boxes).
    struct Seg {
        double cost;
        uint32_t set;
        uint64_t b0, b1;  // cells, relative to the set
        uint32_t row;
    };
    std::vector<Seg> segs;
    for (uint32_t i = 0; i < n_sets; ++i) {
        const Setup &s = pl->setups[i];
        const uint32_t gi = g_set[i];
        std::vector<double> wseg(gi, 1.0);
        if (n_sets == 1 && per_wg > 1 && gi == resident * per_wg) {
            const double q = pow(0.125, 1.0 / (double)(per_wg - 1));
            for (uint32_t k = 0; k < gi; ++k) {
                const uint32_t slot = (k % resident) / (uint32_t)n_cu;  // 0: oldest on its CU
                wseg[k] = pow(q, (double)(k / resident)) * pow(1.32, (double)(occ - 1 - (int)slot));
            }
        }
        std::vector<double> cum(gi);
        double acc = 0.0, tot = 0.0;
        for (double x : wseg)
            tot += x;
        for (uint32_t k = 0; k < gi; ++k) {
            acc += wseg[k];
            cum[k] = k + 1 == gi ? 1.0 : acc / tot;
        }
        std::vector<uint64_t> b;
        std::vector<uint32_t> r;
        split_segments(s, cum, b, r);
        const double sc = set_cost(s);
        for (uint32_t k = 0; k < gi; ++k)
            segs.push_back(Seg{sc * wseg[k] / tot, i, b[k], b[k + 1], r[k]});
    }
    std::stable_sort(segs.begin(), segs.end(), [](const Seg &x, const Seg &y) { return x.cost > y.cost; });
    std::vector<uint64_t> seg_begin, seg_end;
    std::vector<uint32_t> seg_row, seg_set;
    for (const Seg &sg : segs) {
        const SetInfo &si = info[sg.set];
        seg_begin.push_back(si.cell_base + sg.b0);
        seg_end.push_back(si.cell_base + sg.b1);
        seg_row.push_back(si.row_base + sg.row);
        seg_set.push_back(sg.set);
    }
    const uint32_t g = (uint32_t)seg_row.size();
    // no more window rows than a segment spans
    uint32_t span = 1;
    for (uint32_t k = 0; k < g; ++k) {
        const Setup &s = pl->setups[seg_set[k]];
        const SetInfo &si = info[seg_set[k]];
        if (!s.n_cells)
            continue;
        const uint64_t first = seg_begin[k] - si.cell_base;
        const uint64_t end = seg_end[k] - si.cell_base;
        const uint64_t lc = end > first ? end - 1 : first;
        const uint64_t lcc = lc < s.n_cells ? lc : s.n_cells - 1;
        const uint32_t lr = (uint32_t)(std::upper_bound(s.start.begin(), s.start.end() - 1, lcc) - s.start.begin()) - 1;
        span = std::max(span, lr - (seg_row[k] - si.row_base) + 1);
    }
    win_rows = std::min(win_rows, span);
    // u32 window counters must not overflow: bound the leaves any window can see
    for (uint32_t k = 0; k < g && win_rows; ++k) {
        const Setup &s = pl->setups[seg_set[k]];
        const uint32_t r0 = seg_row[k] - info[seg_set[k]].row_base;
        const uint32_t nr = (uint32_t)s.v.size();
        double bound = 0.0;
        for (uint32_t r = r0; r < std::min<uint32_t>(nr, r0 + win_rows); ++r)
            bound += (double)(s.start[r + 1] - s.start[r]) * ldexp(1.0, (int)(s.row_info[r] & 0xFFu));
        if (bound >= 4.0e9)
            win_rows = 0;
    }
    // the window must not cost a resident workgroup (LDS allocation granularity)
    for (;;) {
        ls.lds_bytes = sim_lds_bytes(win_rows, kdim, T, lds_levels, (uint32_t)bs, lds_rows);
        int occ_w = 0;
        if (sim_occupancy(bs, ls.lds_bytes, lds_rows != 0, kdim, gstack, &occ_w) != hipSuccess)
            return fail(PROCELL_E_HIP);
        // keep a margin for the per-workgroup LDS allocation granularity
        const bool fits = (size_t)occ * align_up(ls.lds_bytes, 2048) <= lds_per_cu;
        if ((occ_w >= occ && fits) || win_rows == 0)
            break;
        win_rows = win_rows > 8 ? win_rows - win_rows / 8 : 0;
    }
    ls.blocks = (int)g;
    a.win_rows = win_rows;
    a.idle_limit = tu.idle_lanes ? std::min<uint32_t>(tu.idle_lanes, 63) : 12;

    // ---- one device allocation: counters | tally | tables ----
    size_t off = 0;
    const size_t o_cnt = off; off += (size_t)n_sets * kCounterCopies * kNumCounters * 8;
    const size_t o_tally = off; off += pl->tally_total * 8;
    pl->zero_words = off / 8;
    off = align_up(off, 256);
    const size_t o_rs = off; off = align_up(off + (size_t)(n_rows + 1) * 4, 256);
    const size_t o_ri = off; off = align_up(off + (size_t)(n_rows + 1) * 4, 256);
    const size_t o_cum = off; off = align_up(off + (size_t)n_types_tot * 8, 256);
    const size_t o_mean = off; off = align_up(off + (size_t)n_types_tot * 8, 256);
    const size_t o_sd = off; off = align_up(off + (size_t)n_types_tot * 8, 256);
    const size_t o_tf = off; off = align_up(off + (size_t)n_types_tot * 4, 256);
    const size_t o_sb = off; off = align_up(off + (size_t)g * 8, 256);
    const size_t o_se = off; off = align_up(off + (size_t)g * 8, 256);
    const size_t o_sr = off; off = align_up(off + (size_t)g * 4, 256);
    const size_t o_ss = off; off = align_up(off + (size_t)g * 4, 256);
    const size_t o_info = off; off = align_up(off + (size_t)n_sets * sizeof(SetInfo), 256);
    const size_t o_seed = off; off = align_up(off + (size_t)n_sets * 8, 256);
    const size_t o_gs = off; off = align_up(off + (gstack ? (size_t)g * a.stack_levels * bs * sizeof(double) : 0), 256);
    if (arena) {
        if (off > arena_bytes)
            return fail(PROCELL_E_SCRATCH);
        pl->stream = (hipStream_t)ps[0].stream;
        pl->d_block = arena;
        pl->own_block = false;
    } else {
        if (ps[0].stream) {
            pl->stream = (hipStream_t)ps[0].stream;
        } else {
            if (hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) != hipSuccess)
                return fail(PROCELL_E_HIP);
            pl->own_stream = true;
        }
        if (hipMalloc(&pl->d_block, off) != hipSuccess)
            return fail(PROCELL_E_OUT_OF_MEMORY);
    }
    char *dbase = (char *)pl->d_block;
    a.n_sets = n_sets;
    a.n_cells = n_cells;
    a.n_rows = n_rows;
    a.n_types = T;
    a.kdim = kdim;
    a.counters = (unsigned long long *)(dbase + o_cnt);
    a.tally = (unsigned long long *)(dbase + o_tally);
    a.gstack = gstack ? (double *)(dbase + o_gs) : nullptr;

    a.row_start = (const uint32_t *)(dbase + o_rs);
    a.row_info = (const uint32_t *)(dbase + o_ri);
    a.type_cum = (const double *)(dbase + o_cum);
    a.type_mean = (const double *)(dbase + o_mean);
    a.type_sd = (const double *)(dbase + o_sd);
    a.type_file = (const uint32_t *)(dbase + o_tf);
    a.seg_begin = (const uint64_t *)(dbase + o_sb);
    a.seg_end = (const uint64_t *)(dbase + o_se);
    a.seg_row = (const uint32_t *)(dbase + o_sr);
    a.seg_set = (const uint32_t *)(dbase + o_ss);
    a.sets = (const SetInfo *)(dbase + o_info);
    a.set_seed = (const uint64_t *)(dbase + o_seed);
    pl->d_seeds = (uint64_t *)(dbase + o_seed);
    // concatenated host tables (cells < 2^32)
    std::vector<uint32_t> start32, info32;
    std::vector<double> cum, mean, sd;
    std::vector<uint32_t> tfile;
    start32.reserve(n_rows + 1);
    for (uint32_t i = 0; i < n_sets; ++i) {
        const Setup &s = pl->setups[i];
        for (size_t r = 0; r < s.v.size(); ++r) {
            start32.push_back((uint32_t)(info[i].cell_base + s.start[r]));
            info32.push_back(s.row_info[r]);
        }
        const uint32_t Ti = info[i].n_types;
        for (uint32_t j = 0; j < Ti; ++j) {
            cum.push_back(j < s.type_cum.size() ? s.type_cum[j] : 1.0);
            mean.push_back(j < s.type_mean.size() ? s.type_mean[j] : -1.0);
            sd.push_back(j < s.type_sd.size() ? s.type_sd[j] : -1.0);
            tfile.push_back(j < s.type_file.size() ? s.type_file[j] : 0u);
        }
    }
    start32.push_back((uint32_t)n_cells);
    info32.push_back(0u);
    auto up = [&](size_t o, const void *src, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(dbase + o, src, bytes, hipMemcpyHostToDevice, pl->stream) == hipSuccess;
    };
    if (!up(o_rs, start32.data(), start32.size() * 4) || !up(o_ri, info32.data(), info32.size() * 4) ||
        !up(o_cum, cum.data(), cum.size() * 8) || !up(o_mean, mean.data(), mean.size() * 8) ||
        !up(o_sd, sd.data(), sd.size() * 8) || !up(o_tf, tfile.data(), tfile.size() * 4) ||
        !up(o_sb, seg_begin.data(), seg_begin.size() * 8) || !up(o_se, seg_end.data(), seg_end.size() * 8) || !up(o_sr, seg_row.data(), seg_row.size() * 4) ||
        !up(o_ss, seg_set.data(), seg_set.size() * 4) || !up(o_info, info.data(), info.size() * sizeof(SetInfo)))
        return fail(PROCELL_E_HIP);
    if (hipStreamSynchronize(pl->stream) != hipSuccess)
        return fail(PROCELL_E_HIP);
    if (pl->timing) {
        pl->kev.assign(2 * kTimedRuns, nullptr);
        for (hipEvent_t &e : pl->kev)
            if (hipEventCreate(&e) != hipSuccess)
                return fail(PROCELL_E_HIP);
    }
    if (n_sets > 1) {
        if (hipHostMalloc((void **)&pl->h_seeds, (size_t)kSeedSlots * n_sets * 8) != hipSuccess)
            return fail(PROCELL_E_OUT_OF_MEMORY);
        for (hipEvent_t &e : pl->seed_ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return fail(PROCELL_E_HIP);
    }
    *out = pl;
    return PROCELL_OK;
}

// The run's kernels: clearing, then the simulation.
hipError_t launch_run(procell_plan *pl)
{
    hipError_t e = launch_zero(pl->args.counters, pl->zero_words, pl->stream);
    if (e == hipSuccess)
        e = launch_sim(pl->args, pl->shape, pl->stream);
    return e;
}

// Batch runs replay one graph of the three kernels; it is captured on the
// first run (the kernels read the seeds from device memory, so the graph
// does not change between runs) and again after the launch knobs change.
int capture_graph(procell_plan *pl)
{
    if (pl->graph_exec)
        return PROCELL_OK;
    HIP_TRY(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t e = launch_run(pl);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(pl->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph)
            (void)hipGraphDestroy(graph);
        g_last_hip = e != hipSuccess ? e : e2;
        return PROCELL_E_HIP;
    }
    pl->graph = graph;
    HIP_TRY(hipGraphInstantiate(&pl->graph_exec, pl->graph, nullptr, nullptr, 0));
    return PROCELL_OK;
}

void drop_graph(procell_plan *pl)
{
    if (pl->graph_exec)
        (void)hipGraphExecDestroy(pl->graph_exec);
    if (pl->graph)
        (void)hipGraphDestroy(pl->graph);
    pl->graph_exec = nullptr;
    pl->graph = nullptr;
}

// One run: every set gets seeds[i] (or `seed` when seeds is null).
int run_plan(procell_plan *pl, const uint64_t *seeds, uint64_t seed, bool use_graph = true)
{
    HIP_TRY(hipSetDevice(pl->device));
    const uint32_t n_sets = pl->args.n_sets;
    // timing events bracket the whole run: an event recorded between two
    // kernels of the run would add a gap of several microseconds there
    const size_t slot = (size_t)(pl->runs % kTimedRuns);
    if (n_sets == 1) {
        pl->args.seed = seeds ? seeds[0] : seed;
        if (pl->timing)
            HIP_TRY(hipEventRecord(pl->kev[2 * slot], pl->stream));
        HIP_TRY(launch_run(pl));
    } else {
        // stage the seeds in a pinned slot whose previous copy has run, and
        // copy them to the device ahead of the kernels
        const uint32_t k = (uint32_t)(pl->batch_runs % kSeedSlots);
        HIP_TRY(hipEventSynchronize(pl->seed_ev[k]));
        uint64_t *h = pl->h_seeds + (size_t)k * n_sets;
        for (uint32_t i = 0; i < n_sets; ++i)
            h[i] = seeds ? seeds[i] : seed;
        HIP_TRY(hipMemcpyAsync(pl->d_seeds, h, (size_t)n_sets * 8, hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipEventRecord(pl->seed_ev[k], pl->stream));
        ++pl->batch_runs;
        if (use_graph) {
            const int rc = capture_graph(pl);
            if (rc != PROCELL_OK)
                return rc;
        }
        if (pl->timing)
            HIP_TRY(hipEventRecord(pl->kev[2 * slot], pl->stream));
        if (use_graph)
            HIP_TRY(hipGraphLaunch(pl->graph_exec, pl->stream));
        else
            HIP_TRY(launch_run(pl));
    }
    if (pl->timing)
        HIP_TRY(hipEventRecord(pl->kev[2 * slot + 1], pl->stream));
    if (pl->timing) {
        pl->last_slot = slot;
        pl->timed = true;
    }
    ++pl->runs;
    pl->ran = true;
    return PROCELL_OK;
}

int plan_result(procell_plan *pl, uint32_t set, procell_result *r)
{
    if (!r || !pl->ran || set >= pl->args.n_sets)
        return PROCELL_E_INVALID;
    const Setup &s = pl->setups[set];
    const size_t tally_len = pl->tally_len[set];
    if (r->capacity && !(r->value && r->count))
        return PROCELL_E_INVALID;
    HIP_TRY(hipSetDevice(pl->device));
    std::vector<unsigned long long> cnt_all((size_t)kCounterCopies * kNumCounters), cnt(kNumCounters, 0);
    std::vector<int64_t> tally(tally_len);
    HIP_TRY(hipMemcpyAsync(cnt_all.data(), pl->args.counters + (size_t)set * kCounterCopies * kNumCounters,
                           cnt_all.size() * 8, hipMemcpyDeviceToHost, pl->stream));
    if (tally_len)
        HIP_TRY(hipMemcpyAsync(tally.data(), pl->args.tally + pl->tally_base[set], tally_len * 8,
                               hipMemcpyDeviceToHost, pl->stream));
    HIP_TRY(hipStreamSynchronize(pl->stream));
    for (size_t i = 0; i < cnt_all.size(); ++i)
        cnt[i % kNumCounters] += cnt_all[i];
    procell_stats &st = r->stats;
    fill_stats(cnt.data(), tally.data(), tally_len, st);
    st.workgroups = (uint32_t)pl->shape.blocks;
    st.block_size = (uint32_t)pl->shape.block_size;
    st.lds_bytes = (uint32_t)pl->shape.lds_bytes;
    st.window_rows = pl->args.win_rows;
    if (pl->timing && pl->timed) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, pl->kev[2 * pl->last_slot], pl->kev[2 * pl->last_slot + 1]));
        st.ms_tree = ms;
        st.ms_run = ms;
    }
    if (r->raw_tally) {
        if (r->raw_tally_len < tally_len)
            return PROCELL_E_CAPACITY;
        memcpy(r->raw_tally, tally.data(), tally_len * 8);
    }
    return merge_tally(s, tally.data(), r);
}

// In host code `scratch` names a host-side stand-in, not device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device with that device current, and only that address is handed
// to kernels and copies.
constexpr int kScratchDevices = 64;

std::mutex g_scratch_mu[kScratchDevices];  // held for a whole call on that device
std::mutex g_scratch_addr_mu;
void *g_scratch_ptr[kScratchDevices] = {};

// Device address of `device`'s scratch array; the device must be current.
int scratch_address(int device, void **ptr)
{
    std::lock_guard<std::mutex> lk(g_scratch_addr_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return PROCELL_E_HIP;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return PROCELL_OK;
}

}  // namespace

extern "C" {

int procell_abi_version(void) { return PROCELL_ABI_VERSION; }

const char *procell_strerror(int code)
{
    switch (code) {
    case PROCELL_OK: return "success";
    case PROCELL_E_INVALID: return "invalid argument";
    case PROCELL_E_PROPORTIONS: return "proportion distribution of cell types does not sum to 1";
    case PROCELL_E_ZERO_MEAN: return "a cell type has mean 0";
    case PROCELL_E_NONFINITE: return "non-finite value in the input";
    case PROCELL_E_PHI: return "phi must be > 0";
    case PROCELL_E_DEPTH: return "division tree deeper than 62 generations (phi too small)";
    case PROCELL_E_TOO_MANY_CELLS: return "input exceeds 2^32-1 cells, 2^24 rows or 256 types";
    case PROCELL_E_CAPACITY: return "result buffer too small";
    case PROCELL_E_HIP: return hipGetErrorString(g_last_hip);
    case PROCELL_E_NO_DEVICE: return "no usable gfx950 device";
    case PROCELL_E_IO: return "cannot write output file";
    case PROCELL_E_OUT_OF_MEMORY: return "out of memory";
    case PROCELL_E_SCRATCH: return "input needs more device memory than the 16 MiB static scratch of procell_simulate_noalloc";
    default: return "unknown error";
    }
}

int procell_plan_create(const procell_params *p, procell_plan **plan) { return create_plan(p, 1, plan); }

int procell_plan_run(procell_plan *plan, uint64_t seed)
{
    if (!plan)
        return PROCELL_E_INVALID;
    return run_plan(plan, nullptr, seed);
}

int procell_plan_result(procell_plan *plan, procell_result *r)
{
    if (!plan)
        return PROCELL_E_INVALID;
    return plan_result(plan, 0, r);
}

int procell_batch_create(const procell_params *sets, uint32_t n_sets, procell_plan **plan)
{
    return create_plan(sets, n_sets, plan);
}

int procell_batch_run(procell_plan *plan, const uint64_t *seeds)
{
    if (!plan || !seeds)
        return PROCELL_E_INVALID;
    return run_plan(plan, seeds, 0);
}

int procell_batch_result(procell_plan *plan, uint32_t set, procell_result *r)
{
    if (!plan)
        return PROCELL_E_INVALID;
    return plan_result(plan, set, r);
}

int procell_batch_autotune(procell_plan *plan, const uint64_t *seeds, uint32_t reps, procell_tuning *chosen)
{
    return procell_plan_autotune(plan, seeds, reps, chosen);
}

// Device time of warm-up + reps runs (eager launches) of a plan.
static int time_plan(procell_plan *pl, const uint64_t *seeds, uint32_t reps, hipEvent_t e0, hipEvent_t e1, float *ms)
{
    const bool timing = pl->timing;
    pl->timing = false;
    int rc = run_plan(pl, seeds, 0, false);  // warm-up
    if (rc == PROCELL_OK && hipEventRecord(e0, pl->stream) != hipSuccess)
        rc = PROCELL_E_HIP;
    for (uint32_t i = 0; rc == PROCELL_OK && i < reps; ++i)
        rc = run_plan(pl, seeds, 0, false);
    if (rc == PROCELL_OK &&
        (hipEventRecord(e1, pl->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
         hipEventElapsedTime(ms, e0, e1) != hipSuccess))
        rc = PROCELL_E_HIP;
    pl->timing = timing;
    return rc;
}

int procell_plan_autotune(procell_plan *plan, const uint64_t *seeds, uint32_t reps, procell_tuning *chosen)
{
    if (!plan || !seeds || !chosen || reps == 0)
        return PROCELL_E_INVALID;
    HIP_TRY(hipSetDevice(plan->device));
    // layouts: resident workgroups per CU (sets the LDS tally window's
    // width) x segments per resident workgroup; for each, the idle-lane
    // threshold, which needs no new layout
    static const uint32_t kBlocks[] = {2, 3, 4};
    static const uint32_t kSegments[] = {1, 2, 3, 4, 8};
    static const uint32_t kIdle[] = {6, 8, 12, 16};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return PROCELL_E_HIP;
    }
    const uint32_t n_sets = plan->args.n_sets;
    std::vector<procell_params> ps = plan->params;
    float best = -1.f;
    procell_plan *best_plan = nullptr;
    procell_tuning best_tu{};
    int rc = PROCELL_OK;
    for (uint32_t bpc : kBlocks) {
        for (uint32_t seg : kSegments) {
            for (procell_params &p : ps) {
                p.tuning.blocks_per_cu = bpc;
                p.tuning.segments = seg;
                p.tuning.idle_lanes = 0;
                p.tuning.flags &= ~PROCELL_FLAG_TIMING;  // candidates need no timing ring
                p.stream = (void *)plan->stream;         // and run on the plan's stream
            }
            procell_plan *cand = nullptr;
            rc = create_plan(ps.data(), n_sets, &cand);
            if (rc != PROCELL_OK)
                break;
            cand->own_stream = false;
            bool kept = false;
            for (uint32_t idle : kIdle) {
                cand->args.idle_limit = idle;
                float ms = 0.f;
                rc = time_plan(cand, seeds, reps, e0, e1, &ms);
                if (rc != PROCELL_OK)
                    break;
                if (best < 0.f || ms < best) {
                    best = ms;
                    best_tu = ps[0].tuning;
                    best_tu.idle_lanes = idle;
                    best_tu.block_size = (uint32_t)cand->shape.block_size;
                    if (best_plan && best_plan != cand)
                        destroy_plan(best_plan);
                    best_plan = cand;
                    kept = true;
                }
            }
            if (best_plan == cand)
                best_plan->args.idle_limit = best_tu.idle_lanes;
            if (!kept)
                destroy_plan(cand);
            if (rc != PROCELL_OK)
                break;
        }
        if (rc != PROCELL_OK)
            break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != PROCELL_OK || !best_plan) {
        if (best_plan)
            destroy_plan(best_plan);
        return rc != PROCELL_OK ? rc : PROCELL_E_INVALID;
    }
    // the plan takes the fastest candidate's layout; its own stream, events
    // and timing ring stay
    drop_graph(plan);
    std::swap(plan->shape, best_plan->shape);
    std::swap(plan->args, best_plan->args);
    std::swap(plan->d_block, best_plan->d_block);
    std::swap(plan->tally_total, best_plan->tally_total);
    std::swap(plan->zero_words, best_plan->zero_words);
    std::swap(plan->d_seeds, best_plan->d_seeds);
    std::swap(plan->tally_base, best_plan->tally_base);
    std::swap(plan->tally_len, best_plan->tally_len);
    std::swap(plan->setups, best_plan->setups);
    best_tu.flags = plan->params[0].tuning.flags;
    for (uint32_t i = 0; i < n_sets; ++i)
        plan->params[i].tuning = best_tu;
    destroy_plan(best_plan);
    plan->runs = 0;
    plan->ran = false;
    *chosen = best_tu;
    return PROCELL_OK;
}

int procell_plan_destroy(procell_plan *plan)
{
    destroy_plan(plan);
    return PROCELL_OK;
}

void *procell_plan_stream(procell_plan *plan) { return plan ? (void *)plan->stream : nullptr; }

int procell_plan_kernel_times(procell_plan *plan, double *ms, uint32_t max_runs, uint32_t *n_runs)
{
    if (!plan || !n_runs || (max_runs && !ms))
        return PROCELL_E_INVALID;
    if (!plan->timing)
        return PROCELL_E_INVALID;
    HIP_TRY(hipSetDevice(plan->device));
    HIP_TRY(hipStreamSynchronize(plan->stream));
    const uint64_t avail = plan->runs < kTimedRuns ? plan->runs : kTimedRuns;
    const uint32_t n = (uint32_t)(avail < max_runs ? avail : max_runs);
    for (uint32_t i = 0; i < n; ++i) {
        const size_t slot = (size_t)((plan->runs - n + i) % kTimedRuns);
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, plan->kev[2 * slot], plan->kev[2 * slot + 1]));
        ms[i] = t;
    }
    *n_runs = n;
    plan->runs = 0;
    return PROCELL_OK;
}

int procell_simulate(const procell_params *p, procell_result *r)
{
    if (!p || !r)
        return PROCELL_E_INVALID;
    procell_params q = *p;
    q.stream = nullptr;
    procell_plan *pl = nullptr;
    int rc = create_plan(&q, 1, &pl);
    if (rc != PROCELL_OK)
        return rc;
    rc = run_plan(pl, nullptr, p->seed);
    if (rc == PROCELL_OK)
        rc = plan_result(pl, 0, r);
    destroy_plan(pl);
    return rc;
}

int procell_simulate_noalloc(const procell_params *p, procell_result *r)
{
    if (!p || !r)
        return PROCELL_E_INVALID;
    int n_cu = 0;
    int rc = check_device(p->device, &n_cu);  // makes the device current
    if (rc != PROCELL_OK)
        return rc;
    if (p->device >= kScratchDevices)
        return PROCELL_E_INVALID;
    std::lock_guard<std::mutex> lk(g_scratch_mu[p->device]);
    void *ptr = nullptr;
    rc = scratch_address(p->device, &ptr);
    if (rc != PROCELL_OK)
        return rc;
    procell_params q = *p;
    q.tuning.flags &= ~PROCELL_FLAG_TIMING;
    procell_plan *pl = nullptr;
    rc = create_plan(&q, 1, &pl, ptr, kScratchBytes);
    if (rc != PROCELL_OK)
        return rc;
    rc = run_plan(pl, nullptr, p->seed);
    if (rc == PROCELL_OK)
        rc = plan_result(pl, 0, r);
    destroy_plan(pl);
    return rc;
}

int procell_debug_draws(int32_t device, uint64_t seed, const uint64_t *cell, const uint64_t *node,
                        const uint32_t *attempt, uint64_t n, uint32_t *words, double *z)
{
    if (n && (!cell || !node || !attempt || !words || !z))
        return PROCELL_E_INVALID;
    int n_cu = 0;
    int rc = check_device(device, &n_cu);
    if (rc != PROCELL_OK)
        return rc;
    if (n == 0)
        return PROCELL_OK;
    void *d = nullptr;
    const size_t bytes = n * (8 + 8 + 4 + 16 + 16);
    if (hipMalloc(&d, bytes) != hipSuccess)
        return PROCELL_E_OUT_OF_MEMORY;
    char *b = (char *)d;
    uint64_t *dc = (uint64_t *)b;
    uint64_t *dn = (uint64_t *)(b + n * 8);
    double *dz = (double *)(b + n * 16);
    uint32_t *dw = (uint32_t *)(b + n * 32);
    uint32_t *da = (uint32_t *)(b + n * 48);
    hipError_t e = hipMemcpy(dc, cell, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn, node, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(da, attempt, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_debug_draws(seed, dc, dn, da, n, dw, dz, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(words, dw, n * 16, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(z, dz, n * 16, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        g_last_hip = e;
        return PROCELL_E_HIP;
    }
    return PROCELL_OK;
}

}  // extern "C"
