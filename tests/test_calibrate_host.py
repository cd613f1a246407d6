"""CPU: host-side helpers of the calibration binding (no GPU needed).

output_bins must give the plain method's output bin set (the README fixture:
every v / 2^k >= phi, distinct, ascending) and map each (nonzero row,
generation) to the bin of v / 2^k, -1 where nothing can be recorded; pack /
unpack must round-trip; the Nelder-Mead coroutine must find the minimum of a
quadratic; plain_loss must vanish at the target.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_output_bins_match_readme_fixture(procell_lib):
    from procell import calibrate

    v, c = procell_lib.read_histogram(os.path.join(GOLDEN, "readme_histogram.txt"))
    p, m, s = procell_lib.read_types(os.path.join(GOLDEN, "readme_types.txt"))
    prob = procell_lib.Problem(v, c, p, m, s, 168.0)
    with open(os.path.join(GOLDEN, "readme_bins.json")) as f:
        g = json.load(f)
    rows, kdim, _ = prob.tally_shape()
    bins, idx = calibrate.output_bins(prob, kdim)
    assert bins.tolist() == g["bins"]
    assert idx.shape == (rows, kdim)
    # nonzero rows: 8.144 (phi itself: generation 0 only) and 9823.85 (generations 0..10)
    assert idx[0].tolist() == [0] + [-1] * (kdim - 1)
    assert [bins[i] for i in idx[1]] == [9823.85 / 2**k for k in range(kdim)]


def test_bins_skip_generation_zero_below_phi(procell_lib):
    from procell import calibrate

    prob = procell_lib.Problem([3.0, 20.0], [5, 5], [1.0], [10.0], [2.0], 50.0, phi=4.0)
    rows, kdim, _ = prob.tally_shape()
    bins, idx = calibrate.output_bins(prob, kdim)
    assert bins.tolist() == [5.0, 10.0, 20.0]
    assert idx[0].tolist() == [-1] * kdim               # 3 < phi and 1.5 < phi: nothing recorded
    assert idx[1].tolist()[:2] == [2, 1] and idx[1, 2] == 0   # 20, 10, 5 (5 > 4: a division stops there)


def test_pack_unpack_round_trip():
    from procell import calibrate

    mean, sd, p = np.array([50.0, -1.0, 30.0]), np.array([10.0, -1.0, 5.0]), np.array([0.2, 0.3, 0.5])
    x = calibrate.pack(mean, sd, p)
    assert len(x) == 2 * 2 + 3
    m2, s2, p2 = calibrate.unpack(x, mean >= 0)
    assert np.allclose(m2, mean) and np.allclose(s2, sd) and np.allclose(p2, p)


def test_nelder_mead_finds_quadratic_minimum():
    from procell import calibrate

    target = np.array([1.0, -2.0, 0.5, 3.0])
    f = lambda x: float(((x - target) ** 2 * np.array([1.0, 4.0, 0.5, 2.0])).sum())
    co = calibrate.nelder_mead(np.zeros(4), 0.5)
    x = next(co)
    best = (np.inf, None)
    for _ in range(1500):
        fx = f(x)
        if fx < best[0]:
            best = (fx, x.copy())
        x = co.send(fx)
    assert np.allclose(best[1], target, atol=1e-4), best


def test_plain_loss_zero_at_target():
    from procell import calibrate

    bins = np.array([1.0, 2.0, 4.0])
    q = np.array([1.0, 3.0, 4.0]) / 8.0
    assert calibrate.plain_loss(bins, np.array([1.0, 2.0, 4.0]), np.array([2, 6, 8]), q) == 0.0
    assert calibrate.plain_loss(bins, np.array([1.0, 4.0]), np.array([2, 8]), q) > 0.0
