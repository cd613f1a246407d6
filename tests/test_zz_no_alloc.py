"""procell_simulate_noalloc (GPU): the entry point that allocates no device
memory, against procell_simulate on the same inputs, bit for bit.

Both run the same kernels; the no-alloc variant carves its device buffers
from the library's static per-GPU scratch array, which keeps the previous
call's contents.  Two calls in a row, the second on a smaller input, show
that nothing of the first call leaks into the second (counters and tally are
cleared per run, tables and segments rewritten).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATS = ("divisions", "recorded", "dropped", "lost", "rejected", "cells", "pool",
         "workgroups", "block_size", "lds_bytes", "window_rows")


def _same(ref, got):
    assert np.array_equal(ref.value, got.value)
    assert np.array_equal(ref.count, got.count)
    assert np.array_equal(ref.type_count, got.type_count)
    assert np.array_equal(ref.raw_tally, got.raw_tally)
    for k in STATS:
        assert ref.stats[k] == got.stats[k], (k, ref.stats[k], got.stats[k])


def test_noalloc_matches_simulate_on_two_calls(procell_lib):
    from procell import workloads

    first = workloads.scaled(workloads.large_histogram(0), 0.1, "hist1024_1e6")
    second = workloads.small_histogram(1)  # 64 rows, 3 types, ~1e5 cells
    probs = [procell_lib.Problem.from_workload(first, 101), procell_lib.Problem.from_workload(second, 202)]
    assert np.prod(probs[1].tally_shape()) < np.prod(probs[0].tally_shape())
    assert second.n_cells < first.n_cells
    refs = [procell_lib.simulate(p) for p in probs]
    print("noalloc: first call hist1024_1e6 seed 101, then hist64_1e5 seed 202", flush=True)
    got = [procell_lib.simulate_noalloc(p) for p in probs]
    for r, g in zip(refs, got):
        assert r.stats["divisions"] > 0
        _same(r, g)


def test_noalloc_after_a_batch_of_other_work(procell_lib):
    """A plan run between two no-alloc calls changes nothing in either."""
    from procell import workloads

    w = workloads.scaled(workloads.small_histogram(3), 0.5, "hist64_5e4")
    p = procell_lib.Problem.from_workload(w, 7)
    ref = procell_lib.simulate(p)
    _same(ref, procell_lib.simulate_noalloc(p))
    with procell_lib.Plan(procell_lib.Problem.from_workload(workloads.small_histogram(4), 8)) as plan:
        plan.run(8)
        assert plan.result().stats["divisions"] > 0
    _same(ref, procell_lib.simulate_noalloc(p))


def test_noalloc_rejects_inputs_beyond_its_scratch(procell_lib):
    """An input whose device buffers exceed the static array fails before any kernel runs."""
    n = 60000
    values = np.geomspace(1.0, 1e6, n)
    counts = np.ones(n, dtype=np.uint64)
    p = procell_lib.Problem(values, counts, np.array([0.5, 0.3, 0.2]), np.array([40.0, 60.0, -1.0]),
                            np.array([10.0, 15.0, -1.0]), 100.0, 1.0, 3)
    rows, kdim, n_types = p.tally_shape()
    assert rows * kdim * n_types * 8 > 16 * 1024 * 1024
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.simulate_noalloc(p)
    assert e.value.code == procell_lib.E_SCRATCH
    # the array is still usable afterwards
    from procell import workloads

    q = procell_lib.Problem.from_workload(workloads.small_histogram(5), 9)
    _same(procell_lib.simulate(q), procell_lib.simulate_noalloc(q))
