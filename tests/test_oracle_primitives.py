"""CPU checks that pin the oracle's random stream and FP64 math.

The oracle (oracle/procell_oracle.c) is the parity reference of the GPU path;
these tests pin it against what does not depend on it:
  * Philox4x32-10 against the Random123 known-answer vectors,
  * the counter layout of a draw against a direct Philox call,
  * the 53-bit uniforms against exact rational arithmetic,
  * log, sqrt and sin/cos of 2 pi u against mpmath at 300 bits (error bounds
    in units in the last place: 1.5, 0.5 and 2),
  * Box-Muller pairs against the composition of those primitives and, over
    many draws, against the standard normal distribution (moments and a
    Kolmogorov-Smirnov test).
"""
import json
import math
import os
import random
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# inputs that stress the mantissa fold, the ends of the intervals and the
# quarter-turn reduction
EDGE_WORDS = [(0, 0), (0xFFFFFFFF, 0xFFFFFFFF), (0, 2048), (1, 0), (0x80000000, 0), (0xB504F333, 0),
              (0xB504F334, 0), (0x7FFFFFFF, 0xFFFFFFFF), (0, 0x7FF), (0xFFFFFFFF, 0), (0x20000000, 0),
              (0x1FFFFFFF, 0xFFFFFFFF), (0x40000000, 0), (0x60000000, 0), (0xE0000000, 0), (0xDFFFFFFF, 0xFFFFF800)]


def _words(n, seed):
    rng = random.Random(seed)
    return EDGE_WORDS + [(rng.getrandbits(32), rng.getrandbits(32)) for _ in range(n)]


def test_philox_known_answers(oracle):
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.json")) as f:
        kats = json.load(f)
    assert len(kats) == 3
    for k in kats:
        ctr = [int(x, 16) for x in k["ctr"]]
        key = [int(x, 16) for x in k["key"]]
        out = [int(x, 16) for x in k["out"]]
        assert oracle.philox(ctr, key) == out


def test_draw_counter_layout(oracle):
    """block(seed, cell, node, attempt) is Philox with counter
    (node lo, node hi, cell lo, cell[47:32] | attempt << 16) and key (seed lo, seed hi)."""
    rng = random.Random(7)
    for _ in range(200):
        seed = rng.getrandbits(64)
        cell = rng.getrandbits(48)
        node = rng.getrandbits(63) if rng.random() < 0.5 else rng.randrange(4)
        att = rng.getrandbits(16)
        ctr = [node & 0xFFFFFFFF, node >> 32, cell & 0xFFFFFFFF, ((cell >> 32) & 0xFFFF) | (att << 16)]
        key = [seed & 0xFFFFFFFF, seed >> 32]
        assert oracle.block(seed, cell, node, att) == oracle.philox(ctr, key)
    # distinct (node, attempt) give distinct blocks
    seen = {tuple(oracle.block(5, 9, n, a)) for n in range(16) for a in range(4)}
    assert len(seen) == 64


def test_uniforms_exact(oracle):
    for hi, lo in _words(2000, 1):
        m = (hi << 21) + (lo >> 11)
        assert Fraction(oracle.u01_closed_open(hi, lo)) == Fraction(m, 2**53)
        assert Fraction(oracle.u01_open_closed(hi, lo)) == Fraction(m + 1, 2**53)
    assert oracle.u01_closed_open(0, 0) == 0.0
    assert oracle.u01_closed_open(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0**-53
    assert oracle.u01_open_closed(0, 0) == 2.0**-53
    assert oracle.u01_open_closed(0xFFFFFFFF, 0xFFFFFFFF) == 1.0


def _ulps(got, exact):
    import mpmath

    return float(abs(mpmath.mpf(got) - exact) / math.ulp(float(exact)))


def test_log_within_one_and_a_half_ulp(oracle):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 300
    worst = 0.0
    for hi, lo in _words(4000, 2):
        u = oracle.u01_open_closed(hi, lo)
        got = oracle.log(u)
        if u == 1.0:
            assert got == 0.0
            continue
        worst = max(worst, _ulps(got, mpmath.log(mpmath.mpf(u))))
    # powers of two, the fold at 1.5, the table intervals next to 1
    for x in [2.0**-53, 0.5, 0.25, 0.75, 0.7499999999999999, 1.0 - 2.0**-53, 0.999999, 1.0 - 2.0**-8,
              1.0 - 2.0**-8 - 2.0**-40, 1e-10, 0.5 + 2.0**-52]:
        worst = max(worst, _ulps(oracle.log(x), mpmath.log(mpmath.mpf(x))))
    # near 1 the result is exact to about an ulp of the result itself
    for k in range(1, 2000):
        x = 1.0 - k * 977 * 2.0**-60
        worst = max(worst, _ulps(oracle.log(x), mpmath.log(mpmath.mpf(x))))
    assert worst <= 1.5, worst
    assert oracle.log(1.0) == 0.0


def test_sqrt_correctly_rounded(oracle):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 300
    assert oracle.sqrt(0.0) == 0.0
    rng = random.Random(3)
    xs = [2.0**-53 * 2, 1.0, 2.0, 4.0, 73.6827, 2.0 * 36.7368005696771] + [rng.uniform(0, 80) for _ in range(3000)]
    xs += [-2.0 * oracle.log(oracle.u01_open_closed(hi, lo)) for hi, lo in _words(1000, 4)]
    for x in xs:
        if x == 0.0:
            continue
        got = oracle.sqrt(x)
        assert _ulps(got, mpmath.sqrt(mpmath.mpf(x))) <= 0.5 + 1e-9, x
        assert got == math.sqrt(x)


def test_sincos_turns_within_two_ulp(oracle):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 300
    worst = 0.0
    for hi, lo in _words(4000, 5):
        u = mpmath.mpf(oracle.u01_closed_open(hi, lo))
        s, c = oracle.sincos_turns(hi, lo)
        for got, ex in ((s, mpmath.sin(2 * mpmath.pi * u)), (c, mpmath.cos(2 * mpmath.pi * u))):
            if abs(ex) < 1e-60:   # u on a quarter turn: the reduction makes it exact
                assert got == 0.0
            else:
                worst = max(worst, _ulps(got, ex))
    assert worst <= 2.0, worst
    # quarter turns exactly
    assert oracle.sincos_turns(0, 0) == (0.0, 1.0)
    assert oracle.sincos_turns(0x40000000, 0) == (1.0, 0.0)
    assert oracle.sincos_turns(0x80000000, 0) == (0.0, -1.0)
    assert oracle.sincos_turns(0xC0000000, 0) == (-1.0, 0.0)


def test_normal_pair_is_box_muller(oracle):
    """z0 = sqrt(-2 log u1) cos(2 pi u2), z1 = ... sin(...), with u1 from words
    0-1 on (0,1] and u2 from words 2-3 on [0,1)."""
    rng = random.Random(6)
    for _ in range(500):
        w = [rng.getrandbits(32) for _ in range(4)]
        z0, z1 = oracle.normal_pair(w)
        rho = oracle.sqrt(-2.0 * oracle.log(oracle.u01_open_closed(w[0], w[1])))
        s, c = oracle.sincos_turns(w[2], w[3])
        assert z0 == rho * c and z1 == rho * s
        # against libm to a loose relative bound (independent formula)
        u1 = oracle.u01_open_closed(w[0], w[1])
        u2 = oracle.u01_closed_open(w[2], w[3])
        r = math.sqrt(-2.0 * math.log(u1))
        assert abs(z0 - r * math.cos(2 * math.pi * u2)) <= 1e-13 * max(1.0, r)
        assert abs(z1 - r * math.sin(2 * math.pi * u2)) <= 1e-13 * max(1.0, r)


def test_normal_pairs_are_standard_normal(oracle):
    scipy_stats = pytest.importorskip("scipy.stats")
    z = []
    for node in range(1, 20001):
        z.extend(oracle.normal_pair(oracle.block(11, 3, node, 0)))
    z = np.array(z)
    n = len(z)
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert abs(np.mean(z**3)) < 5 * math.sqrt(15.0 / n)
    assert scipy_stats.kstest(z, "norm").pvalue > 1e-4
    # the two members of a pair are uncorrelated
    assert abs(np.corrcoef(z[0::2], z[1::2])[0, 1]) < 5 / math.sqrt(n / 2)


def test_math_tables_shared_and_regenerable():
    """The oracle and the device code include byte-identical copies of the
    log and sin/cos tables, and the generator reproduces them."""
    import importlib.util

    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    a = open(os.path.join(repo, "oracle", "procell_math_tables.inc")).read()
    b = open(os.path.join(repo, "cuda-pro-cell_amd", "csrc", "procell_math_tables.inc")).read()
    assert a == b
    pytest.importorskip("mpmath")
    spec = importlib.util.spec_from_file_location("gen_math_tables", os.path.join(repo, "oracle", "gen_math_tables.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert gen.render() == a
