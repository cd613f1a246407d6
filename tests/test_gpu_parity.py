"""GPU parity: libprocell's HIP kernels against the CPU oracle, bit for bit.

Every comparison is on the raw tally (counts by nonzero row, generation and
type), the merged output histogram and the event counters, all of which must
be identical: the GPU and the oracle evaluate the same Philox stream with the
same FP64 operation sequence.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATS = ("divisions", "recorded", "dropped", "lost", "rejected", "cells")


def _problem_pair(procell, oracle, values, counts, proportion, mean, sd, t_max, phi, seed, **kw):
    gp = procell.Problem(values, counts, proportion, mean, sd, t_max, phi, seed, **kw)
    op = oracle.Problem(values, counts, proportion, mean, sd, t_max, phi, seed)
    return gp, op


def _merge_oracle(oracle, op, tally):
    """Output histogram from the oracle's raw tally (value -> count, per type)."""
    phi = oracle.resolve_phi(op)
    vals = op.values[op.counts > 0]
    out = {}
    for r, v in enumerate(vals):
        f = float(v)
        for k in range(tally.shape[1]):
            row = tally[r, k]
            if row.sum():
                assert f >= phi
                acc = out.setdefault(f, np.zeros(tally.shape[2], dtype=np.int64))
                acc += row
            f = f / 2
    keys = sorted(out)
    return (np.array(keys, dtype=np.float64), np.array([out[k].sum() for k in keys], dtype=np.int64),
            np.array([out[k] for k in keys], dtype=np.int64).reshape(len(keys), tally.shape[2]))


def assert_same(procell, oracle, gp, op, check_output=True):
    res = procell.simulate(gp)
    tally, _, st = oracle.simulate(op)
    assert res.raw_tally.shape == tally.shape
    diff = np.argwhere(res.raw_tally != tally)
    assert diff.size == 0, f"{len(diff)} tally cells differ, first {diff[:5].tolist()}"
    for k in STATS:
        assert res.stats[k] == st[k], (k, res.stats[k], st[k])
    if check_output:
        v, c, tc = _merge_oracle(oracle, op, tally)
        assert np.array_equal(res.value, v)
        assert np.array_equal(res.count, c)
        assert np.array_equal(res.type_count, tc[:, : res.type_count.shape[1]])
    return res, st


def test_device_draws_match_oracle(procell_lib, oracle):
    rng = np.random.default_rng(7)
    n = 1 << 14
    cells = rng.integers(0, 2**48, n, dtype=np.uint64)
    nodes = rng.integers(0, 2**63, n, dtype=np.uint64)
    nodes[:64] = np.arange(64, dtype=np.uint64)
    attempts = rng.integers(0, 2**16, n, dtype=np.uint32)
    attempts[64:128] = 0
    seed = 0x0123456789ABCDEF
    words, z = procell_lib.debug_draws(seed, cells, nodes, attempts)
    for i in range(0, n, 7):
        w = oracle.block(seed, int(cells[i]), int(nodes[i]), int(attempts[i]))
        assert list(words[i]) == w, i
        z0, z1 = oracle.normal_pair(w)
        assert z[i, 0] == z0 and z[i, 1] == z1, (i, z[i], z0, z1)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_small_histogram_bitexact(procell_lib, oracle, seed):
    from procell import workloads

    w = workloads.small_histogram(seed)
    gp = procell_lib.Problem.from_workload(w, seed)
    op = oracle.Problem.from_workload(w, seed)
    assert_same(procell_lib, oracle, gp, op)


def test_large_histogram_bitexact(procell_lib, oracle):
    from procell import workloads

    w = workloads.large_histogram(0)
    gp = procell_lib.Problem.from_workload(w, 0)
    op = oracle.Problem.from_workload(w, 0)
    res, st = assert_same(procell_lib, oracle, gp, op)
    assert st["divisions"] > 1e7


@pytest.mark.parametrize("knobs", [dict(block_size=64, segments=1), dict(block_size=128, segments=7),
                                   dict(block_size=256, segments=64, blocks_per_cu=1),
                                   dict(block_size=192, segments=33, blocks_per_cu=2, idle_lanes=1),
                                   dict(block_size=256, segments=3, blocks_per_cu=3, idle_lanes=20),
                                   dict(block_size=256, segments=2, blocks_per_cu=4, idle_lanes=12)])
def test_launch_knobs_do_not_change_results(procell_lib, oracle, knobs):
    from procell import workloads

    w = workloads.scaled(workloads.small_histogram(5), 0.5)
    gp = procell_lib.Problem.from_workload(w, 99, **knobs)
    op = oracle.Problem.from_workload(w, 99)
    assert_same(procell_lib, oracle, gp, op)


def test_empty_and_zero_rows(procell_lib, oracle):
    gp, op = _problem_pair(procell_lib, oracle, [1.0, 8.0], [0, 0], [1.0], [10.0], [2.0], 100.0, 1.0, 3)
    res = procell_lib.simulate(gp)
    assert len(res.value) == 0 and res.stats["cells"] == 0
    gp, op = _problem_pair(procell_lib, oracle, [], [], [1.0], [10.0], [2.0], 100.0, 1.0, 3)
    res = procell_lib.simulate(gp)
    assert len(res.value) == 0


def test_all_quiescent_is_identity(procell_lib, oracle):
    vals = [3.0, 17.5, 100.0, 2.0]
    cnts = [5, 0, 11, 4]
    gp, op = _problem_pair(procell_lib, oracle, vals, cnts, [0.25, 0.75], [-1.0, -2.0], [-1.0, 3.0],
                           500.0, 2.5, 11)
    res, st = assert_same(procell_lib, oracle, gp, op)
    # rows >= phi recorded unchanged; the 2.0 row is below phi and lost
    assert res.value.tolist() == [3.0, 100.0]
    assert res.count.tolist() == [5, 11]
    assert st["lost"] == 4 and st["divisions"] == 0


def test_t_max_zero(procell_lib, oracle):
    gp, op = _problem_pair(procell_lib, oracle, [64.0, 8.0], [1000, 500], [0.6, 0.4], [5.0, 9.0],
                           [1.0, 2.0], 0.0, 1.0, 5)
    res, st = assert_same(procell_lib, oracle, gp, op)
    assert st["divisions"] == 0
    assert res.count.sum() == 1500


def test_zero_sd_full_trees(procell_lib, oracle):
    # deterministic timers: every lineage of a cell divides in lock step
    gp, op = _problem_pair(procell_lib, oracle, [2.0**20, 3.0 * 2**15], [40, 30], [1.0], [1.0], [0.0],
                           14.5, 1.0, 8)
    res, st = assert_same(procell_lib, oracle, gp, op)
    assert st["divisions"] > 1e5


def test_deep_trees_shrink_workgroup(procell_lib, oracle):
    # phi far below the values: 39 generations allowed, so the per-lane
    # stack is 40 levels deep and the tree kernel runs with smaller workgroups
    gp, op = _problem_pair(procell_lib, oracle, [2.0**40, 3.0 * 2**30], [100, 100], [1.0], [1.0], [1.0],
                           12.0, 1.0, 9)
    res, st = assert_same(procell_lib, oracle, gp, op)
    assert res.raw_tally.shape[1] == 40


def test_heavy_truncation(procell_lib, oracle):
    # sd >> mean: about half of all normal draws are rejected
    gp, op = _problem_pair(procell_lib, oracle, [1000.0, 50.0], [3000, 2000], [0.5, 0.5], [2.0, 30.0],
                           [40.0, 60.0], 60.0, 1.0, 21)
    res, st = assert_same(procell_lib, oracle, gp, op)
    assert st["rejected"] > 1000


@pytest.mark.parametrize("top", [1.0, 3.0, 7.0, 15.0])
def test_shallow_generation_caps(procell_lib, oracle, top):
    """Values a few halvings above phi = 1 (generation caps 0..3, so the
    per-lane stack is one to three levels deep and idle lanes park on its
    last level); t_max 9 against mean timers of 4 and 6, so cells end both
    out of time (recorded) and at the cap (dropped)."""
    vals = np.array([1.0, 2.0, 3.0, 5.0, 8.0, 15.0]) * (top / 15.0) + 1.0
    gp, op = _problem_pair(procell_lib, oracle, vals, [3000, 2000, 1500, 1000, 800, 500],
                           [0.6, 0.4], [4.0, 6.0], [1.5, 2.0], 9.0, 1.0, 17)
    assert_same(procell_lib, oracle, gp, op)


def test_rows_below_phi_and_duplicate_values(procell_lib, oracle):
    # 40 and 20 halve onto shared values; 3 is below phi; unsorted file order
    gp, op = _problem_pair(procell_lib, oracle, [40.0, 3.0, 20.0, 160.0, 20.0], [300, 200, 100, 50, 7],
                           [0.2, 0.5, 0.3], [30.0, 12.0, -1.0], [6.0, 4.0, -1.0], 70.0, 4.0, 77)
    assert_same(procell_lib, oracle, gp, op)


def test_type_ties_and_rounding_of_proportions(procell_lib, oracle):
    gp, op = _problem_pair(procell_lib, oracle, [500.0], [20000], [0.1, 0.3, 0.3, 0.3 + 5e-9],
                           [11.0, 13.0, 17.0, -1.0], [2.0, 3.0, 4.0, -1.0], 90.0, 2.0, 4)
    assert_same(procell_lib, oracle, gp, op)


@pytest.mark.parametrize("seed", [0, 2**32 - 1, 2**64 - 1])
def test_seed_words(procell_lib, oracle, seed):
    gp, op = _problem_pair(procell_lib, oracle, [256.0, 77.0], [5000, 3000], [0.7, 0.3], [20.0, 35.0],
                           [5.0, 9.0], 120.0, 3.0, seed)
    assert_same(procell_lib, oracle, gp, op)


def test_plan_reuse_gives_same_results(procell_lib, oracle):
    from procell import workloads

    w = workloads.scaled(workloads.small_histogram(1), 0.3)
    gp = procell_lib.Problem.from_workload(w, 0)
    with procell_lib.Plan(gp) as plan:
        outs = []
        for seed in (5, 6, 5):
            plan.run(seed)
            outs.append(plan.result())
    assert np.array_equal(outs[0].raw_tally, outs[2].raw_tally)
    assert not np.array_equal(outs[0].raw_tally, outs[1].raw_tally)
    tally, _, _ = oracle.simulate(oracle.Problem.from_workload(w, 6))
    assert np.array_equal(outs[1].raw_tally, tally)


def test_many_rows_use_device_memory_row_tables(procell_lib, oracle):
    # more rows than fit the LDS row tables: row lookups go to device memory
    rng = np.random.default_rng(12)
    n = 5000
    values = np.exp(rng.uniform(np.log(2.0), np.log(5e4), n))
    counts = rng.integers(0, 40, n).astype(np.uint64)
    gp, op = _problem_pair(procell_lib, oracle, values, counts, [0.6, 0.3, 0.1], [25.0, 40.0, -1.0],
                           [8.0, 10.0, -1.0], 110.0, 3.0, 31)
    assert_same(procell_lib, oracle, gp, op)


def test_sparse_rows_many_row_starts_per_step(procell_lib, oracle):
    # one or two cells per row: every 64-cell step spans dozens of rows
    values = [1.5 * 1.01**i for i in range(1500)]
    counts = [1 + (i % 2) for i in range(1500)]
    gp, op = _problem_pair(procell_lib, oracle, values, counts, [0.5, 0.5], [12.0, 30.0], [3.0, 9.0],
                           70.0, 2.0, 17)
    assert_same(procell_lib, oracle, gp, op)


def test_cell_ids_beyond_two_to_the_31(procell_lib):
    """More than 2^31 initial cells (quiescent, so the run is root steps
    only): cell ids and segment bounds past 2^31 must stay exact in the
    kernel's wave-uniform 64-bit values.  The output is the input restricted
    to values >= phi (closed form): every quiescent cell is tallied at its
    own value, or lost below phi."""
    big = 2**31 + 12345
    values = [60.0, 55.0, 10.0]
    counts = [big - 9000, 9000, 7]
    prob = procell_lib.Problem(values, counts, [1.0], [-1.0], [-1.0], 100.0, 50.0, 3)
    print("cells", sum(counts), flush=True)
    res = procell_lib.simulate(prob)
    st = res.stats
    assert st["cells"] == sum(counts) and st["lost"] == 7 and st["divisions"] == 0
    assert res.value.tolist() == [55.0, 60.0]
    assert res.count.tolist() == [9000, big - 9000]
    assert res.type_count.reshape(-1).tolist() == [9000, big - 9000]
    assert res.raw_tally[:, 0, 0].tolist() == [big - 9000, 9000, 0]


def test_autotuned_plan_matches_oracle(procell_lib, oracle):
    """procell_plan_autotune rebuilds the plan's layout (workgroups per CU,
    segments, idle-lane threshold); runs after it equal the oracle bit for
    bit, and the timing ring still works."""
    from procell import workloads

    w = workloads.scaled(workloads.small_histogram(2), 0.3)
    gp, op = _problem_pair(procell_lib, oracle, w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max, 0.0, 77,
                           flags=procell_lib.FLAG_TIMING)
    with procell_lib.Plan(gp) as plan:
        chosen = plan.autotune([77], reps=1)
        print("chosen", chosen, flush=True)
        assert chosen["blocks_per_cu"] in (2, 3, 4) and chosen["segments"] in (1, 2, 3, 4, 8)
        assert chosen["idle_lanes"] in (6, 8, 12, 16)
        for seed in (77, 78):
            plan.run(seed)
            res = plan.result()
            tally, _, st = oracle.simulate(oracle.Problem(w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max,
                                                          0.0, seed))
            assert np.array_equal(res.raw_tally, tally), seed
            for k in STATS:
                assert res.stats[k] == st[k], (seed, k)
        assert len(plan.kernel_times()) == 2
        with pytest.raises(ValueError):
            plan.autotune([1, 2])
