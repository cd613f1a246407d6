"""GPU: the batched plain method equals single runs of its sets, bit for bit.

A batch concatenates heterogeneous parameter sets (rows, cells, types,
t_max, phi) and runs them with one launch of each kernel, replayed from a
hipGraph.  Set i of a batch run with seeds[i] must equal a single-set plan of
the same parameters run with seeds[i]: raw tally, merged histogram, per-type
columns and event counters.  Graph replays are checked one by one, with
eager runs and synchronisations between them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATS = ("divisions", "recorded", "dropped", "lost", "rejected", "cells")


def _sets(procell, seed, n, scale):
    from procell import workloads

    return [procell.Problem.from_workload(w, 0) for w in workloads.parameter_sets(seed, n, scale)]


def _same(got, ref, what=""):
    assert got.raw_tally.shape == ref.raw_tally.shape, what
    diff = np.argwhere(got.raw_tally != ref.raw_tally)
    assert diff.size == 0, f"{what}: {len(diff)} tally cells differ, first {diff[:5].tolist()}"
    assert np.array_equal(got.value, ref.value), what
    assert np.array_equal(got.count, ref.count), what
    assert np.array_equal(got.type_count, ref.type_count), what
    for k in STATS:
        assert got.stats[k] == ref.stats[k], (what, k, got.stats[k], ref.stats[k])


def _single(procell, prob, seed):
    with procell.Plan(prob) as pl:
        pl.run(seed)
        return pl.result()


def test_batch_equals_single_runs_over_graph_replays(procell_lib):
    probs = _sets(procell_lib, 1, 12, 0.02)
    print("batch of", len(probs), "sets", flush=True)
    with procell_lib.Batch(probs) as b:
        for rep in range(3):
            seeds = [1000 * rep + 7 * i + 1 for i in range(len(probs))]
            b.run(seeds)
            got = [b.result(i) for i in range(len(probs))]
            assert any(g.stats["divisions"] > 0 for g in got)
            for i, p in enumerate(probs):
                _same(got[i], _single(procell_lib, p, seeds[i]), f"replay {rep} set {i}")


def test_batch_edge_sets(procell_lib):
    """All-quiescent, t_max 0, deep trees and a one-row set in one batch."""
    P = procell_lib.Problem
    probs = [
        P([8.0, 3.0], [30, 20], [1.0], [-1.0], [-1.0], 40.0, 1.0),                   # all quiescent
        P([100.0, 37.5], [40, 50], [0.5, 0.5], [10.0, 20.0], [3.0, 4.0], 0.0, 1.0),  # t_max 0
        P([2.0**20], [3], [1.0], [2.0], [0.5], 60.0, 1.0),                           # 20 generations
        P([64.0, 5.0, 2.5], [600, 300, 200], [0.5, 0.5], [12.0, 30.0], [11.0, 40.0], 90.0, 2.0),
        P([1000.0], [500], [0.25, 0.75], [5.0, 7.0], [6.0, 2.0], 38.0, 10.0),       # heavy truncation
    ]
    seeds = [11, 12, 13, 14, 15]
    with procell_lib.Batch(probs) as b:
        b.run(seeds)
        for i, p in enumerate(probs):
            _same(b.result(i), _single(procell_lib, p, seeds[i]), f"set {i}")


def test_batch_of_one_set_equals_plan(procell_lib):
    from procell import workloads

    prob = procell_lib.Problem.from_workload(workloads.small_histogram(2), 0)
    with procell_lib.Batch([prob]) as b:
        b.run([77])
        _same(b.result(0), _single(procell_lib, prob, 77))


def test_batch_matches_oracle(procell_lib, oracle):
    from procell import workloads

    ws = workloads.parameter_sets(5, 3, 0.02)
    probs = [procell_lib.Problem.from_workload(w, 0) for w in ws]
    seeds = [21, 22, 23]
    with procell_lib.Batch(probs) as b:
        b.run(seeds)
        for i, w in enumerate(ws):
            op = oracle.Problem.from_workload(w, seeds[i])
            tally, _, st = oracle.simulate(op)
            res = b.result(i)
            assert np.array_equal(res.raw_tally, tally), i
            for k in STATS:
                assert res.stats[k] == st[k], (i, k)


@pytest.mark.parametrize("knobs", [(2, 8, 16), (3, 8, 12), (4, 2, 6)])
def test_deep_batch_stacks_in_device_memory(procell_lib, oracle, knobs):
    """A batch whose deepest set needs more lane-stack levels than a batch
    keeps in LDS (here 14 and 15 levels) runs with the lanes' stacks in
    device memory; every set, shallow or deep, still equals the oracle and
    its own single run (which keeps its stack in LDS), under several launch
    knobs (windows, occupancy, segment counts and hand-out thresholds)."""
    from procell import workloads

    bpc, seg, idle = knobs
    kw = dict(blocks_per_cu=bpc, segments=seg, idle_lanes=idle)
    deep = [
        (["16384", "8192", "5000", "900"], [300, 200, 400, 500], [0.6, 0.4], [6.0, 9.0], [2.0, 3.0], 80.0, 1.0),
        (["65536"], [50], [1.0], [4.0], [1.0], 60.0, 1.0),
    ]
    ws = workloads.parameter_sets(7, 4, 0.02)
    seeds = [31 + i for i in range(len(deep) + len(ws))]
    print("deep batch, knobs", knobs, flush=True)
    probs, ops = [], []
    for i, (v, c, pr, m, sd, tm, phi) in enumerate(deep):
        v = [float(x) for x in v]
        probs.append(procell_lib.Problem(v, c, pr, m, sd, tm, phi, 0, **kw))
        ops.append(oracle.Problem(v, c, pr, m, sd, tm, phi, seeds[i]))
    for i, w in enumerate(ws):
        probs.append(procell_lib.Problem.from_workload(w, 0, **kw))
        ops.append(oracle.Problem.from_workload(w, seeds[len(deep) + i]))
    with procell_lib.Batch(probs) as b:
        b.run(seeds)
        for i, op in enumerate(ops):
            res = b.result(i)
            tally, _, st = oracle.simulate(op)
            assert np.array_equal(res.raw_tally, tally), (knobs, i)
            for k in STATS:
                assert res.stats[k] == st[k], (knobs, i, k)
            if i < len(deep):
                assert res.stats["divisions"] > 100000, (knobs, i)
                _same(res, _single(procell_lib, probs[i], seeds[i]), f"knobs {knobs} set {i}")


def test_autotune_keeps_results(procell_lib):
    probs = _sets(procell_lib, 3, 8, 0.02)
    seeds = [5 + i for i in range(len(probs))]
    with procell_lib.Batch(probs) as b:
        chosen = b.autotune(seeds, reps=1)
        print("chosen", chosen, flush=True)
        assert chosen["idle_lanes"] in (6, 8, 12, 16) and chosen["blocks_per_cu"] in (2, 3, 4)
        assert chosen["segments"] in (1, 2, 3, 4, 8)
        for rep in range(2):  # the graph is captured again for the new layout
            b.run([s + rep for s in seeds])
            for i, p in enumerate(probs):
                _same(b.result(i), _single(procell_lib, p, seeds[i] + rep), f"replay {rep} set {i}")


def test_full_size_batch_sample_equals_single_runs(procell_lib, oracle):
    """The 512-set batch at full size (5.2e7 cells): 16 sets spread over the
    batch equal their single runs bit for bit, and the two smallest of them
    equal the oracle."""
    from procell import workloads

    ws = workloads.parameter_sets(0, 512)
    probs = [procell_lib.Problem.from_workload(w, 0) for w in ws]
    seeds = [9000 + i for i in range(len(ws))]
    pick = list(range(0, 512, 32))
    print("full-size batch:", sum(w.n_cells for w in ws), "cells", flush=True)
    with procell_lib.Batch(probs) as b:
        b.run(seeds)
        got = {i: b.result(i) for i in pick}
    for i in pick:
        _same(got[i], _single(procell_lib, probs[i], seeds[i]), f"set {i}")
    for i in sorted(pick, key=lambda i: ws[i].n_cells)[:2]:
        tally, _, st = oracle.simulate(oracle.Problem.from_workload(ws[i], seeds[i]))
        assert np.array_equal(got[i].raw_tally, tally), i
        for k in STATS:
            assert got[i].stats[k] == st[k], (i, k)
