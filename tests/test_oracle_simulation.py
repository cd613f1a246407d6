"""CPU checks that pin the oracle's simulation (S1-S6 of the plain method).

The reference's division rule (src/simulation/proliferation.cu:319-381,
404-410), initial cells (src/simulation/cell.cu:25-143,
cells_population.cu:97-117) and input rules (src/io/parser.cu:46-154) are
pinned here by
  * a second, independent scalar implementation (Python, exact fma through
    fractions) compared tally-for-tally on tiny inputs,
  * closed forms: all-quiescent and t_max = 0 (output = input rows >= phi),
    sd = 0 (leaf counts 2^K per cell with K from the cell's own phase),
  * invariants on larger inputs: recorded + dropped + lost = cells + divisions,
    per-row mass sum_k (recorded + dropped) 2^-k = count, no tallies beyond a
    row's generation cap, thread-count invariance,
  * the truncation redraw rate against Phi(-mean/sd),
  * distributional agreement with an independent numpy simulation of the same
    rules (different random generator), per generation and type.
"""
import math
from fractions import Fraction

import numpy as np
import pytest


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _problem(oracle, values, counts, prop, mean, sd, t_max, phi=0.0, seed=0):
    return oracle.Problem(values, counts, prop, mean, sd, t_max, phi, seed)


class Brute:
    """Scalar restatement of S2-S6 on top of the stream primitives:
    root block (cell, 0, 0): type uniform from word 0, phase from word 1,
    first-timer candidates from the Box-Muller pair of words 2, 3 (32 bits
    each); pair block (cell, n, 0) for the timers of
    daughters 2n and 2n+1; redraw blocks (cell, n, a >= 1) for node n, two
    candidates each; the root's first timer continues in node 1's redraw
    stream when both of its candidates are <= 0."""

    def __init__(self, oracle, values, counts, prop, mean, sd, t_max, phi, seed):
        self.o = oracle
        self.t_max = t_max
        self.seed = seed
        if phi <= 0:
            phi = min(v for v, c in zip(values, counts) if c > 0)
        self.phi = phi
        self.rows = [(v, c) for v, c in zip(values, counts) if c > 0]
        self.kcap = []
        for v, _ in self.rows:
            k, f = 0, v
            while f / 2 > phi:
                f /= 2
                k += 1
            self.kcap.append(k)
        self.kdim = max(self.kcap) + 1 if self.kcap else 1
        T = len(prop)
        self.T = T
        self.order = sorted(range(T), key=lambda j: -prop[j])   # stable
        self.cum, c = [], 0.0
        for j in self.order:
            c = c + prop[j]
            self.cum.append(c)
        self.mean, self.sd = mean, sd
        self.tally = np.zeros((len(self.rows), self.kdim, max(T, 1)), dtype=np.int64)
        self.drops = np.zeros((len(self.rows), self.kdim), dtype=np.int64)
        self.st = dict(divisions=0, recorded=0, dropped=0, lost=0, rejected=0, cells=0)

    def timer(self, cell, node, mu, sd):
        for a in range(1, 65536):
            w = self.o.block(self.seed, cell, node, a)
            z0, z1 = self.o.normal_pair(w)
            for z in (z0, z1):
                t = _fma(z, sd, mu)
                if t > 0:
                    return t
                self.st["rejected"] += 1
        return mu

    def leaf(self, r, k, j, s):
        if s > self.t_max:
            self.tally[r, k, j] += 1
            self.st["recorded"] += 1
            return None
        if k == self.kcap[r]:
            self.drops[r, k] += 1
            self.st["dropped"] += 1
            return None
        return True

    def grow(self, cell, r, j, mu, sd, node, k, s):
        self.st["divisions"] += 1
        z0, z1 = self.o.normal_pair(self.o.block(self.seed, cell, node, 0))
        times = []
        for z, child in ((z0, 2 * node), (z1, 2 * node + 1)):
            t = _fma(z, sd, mu)
            if not t > 0:
                self.st["rejected"] += 1
                t = self.timer(cell, child, mu, sd)
            times.append(t)
        for t, child in zip(times, (2 * node, 2 * node + 1)):
            if self.leaf(r, k + 1, j, s + t):
                self.grow(cell, r, j, mu, sd, child, k + 1, s + t)

    def run(self):
        cell = 0
        for r, (v, n) in enumerate(self.rows):
            for _ in range(n):
                w = self.o.block(self.seed, cell, 0, 0)
                ut = self.o.u01_closed_open(w[0], 0)
                U = self.o.u01_open_closed(w[1], 0)
                js = next((i for i, c in enumerate(self.cum) if ut < c), self.T - 1)
                j = self.order[js]
                mu, sd = self.mean[j], self.sd[j]
                self.st["cells"] += 1
                if mu < 0:
                    self._root_out(r, v, j)
                else:
                    tau = None
                    for z in self.o.normal_pair([w[2], 0, w[3], 0]):
                        t = _fma(z, sd, mu)
                        if t > 0:
                            tau = t
                            break
                        self.st["rejected"] += 1
                    if tau is None:
                        tau = self.timer(cell, 1, mu, sd)
                    s0 = tau * U + tau
                    if s0 > self.t_max:
                        self._root_out(r, v, j)
                    elif self.kcap[r] == 0:
                        self.drops[r, 0] += 1
                        self.st["dropped"] += 1
                    else:
                        self.grow(cell, r, j, mu, sd, 1, 0, s0)
                cell += 1
        return self

    def _root_out(self, r, v, j):
        if v >= self.phi:
            self.tally[r, 0, j] += 1
            self.st["recorded"] += 1
        else:
            self.st["lost"] += 1


BRUTE_CASES = [
    # values, counts, proportion, mean, sd, t_max, phi, seed
    ([8.0, 3.0], [3, 2], [1.0], [10.0], [3.0], 40.0, 1.0, 1),
    ([100.0, 1.0, 37.5], [4, 0, 5], [0.53, 0.29, 0.18], [48.33, 86.3, -1.0], [21.6, 26.8, -1.0], 168.0, 0.0, 7),
    ([64.0, 5.0, 2.5], [6, 3, 2], [0.5, 0.5], [12.0, 30.0], [11.0, 40.0], 90.0, 2.0, 2**63 + 5),
    ([1000.0], [3], [0.25, 0.75], [5.0, 7.0], [6.0, 2.0], 38.0, 10.0, 123456789),
    ([4.0, 16.0], [5, 5], [0.2, 0.3, 0.5], [3.0, -1.0, 4.0], [0.0, -1.0, 9.0], 17.0, 1.0, 99),
]


@pytest.mark.parametrize("case", range(len(BRUTE_CASES)))
def test_brute_force_agreement(oracle, case):
    values, counts, prop, mean, sd, t_max, phi, seed = BRUTE_CASES[case]
    b = Brute(oracle, values, counts, prop, mean, sd, t_max, phi, seed).run()
    p = _problem(oracle, values, counts, prop, mean, sd, t_max, phi, seed)
    tally, drops, st = oracle.simulate(p, 2, want_drops=True)
    assert st == b.st
    assert tally.shape == b.tally.shape
    assert np.array_equal(tally, b.tally)
    assert np.array_equal(drops, b.drops)
    assert st["divisions"] > 0 or case == 4


class _SwappedDaughters(Brute):
    def grow(self, cell, r, j, mu, sd, node, k, s):
        self.st["divisions"] += 1
        z1, z0 = self.o.normal_pair(self.o.block(self.seed, cell, node, 0))
        for z, child in ((z0, 2 * node), (z1, 2 * node + 1)):
            t = _fma(z, sd, mu)
            if not t > 0:
                self.st["rejected"] += 1
                t = self.timer(cell, child, mu, sd)
            if self.leaf(r, k + 1, j, s + t):
                self.grow(cell, r, j, mu, sd, child, k + 1, s + t)


class _GeCap(Brute):
    """Divide while f/2 >= phi instead of f/2 > phi."""

    def __init__(self, *a):
        super().__init__(*a)
        for r, (v, _) in enumerate(self.rows):
            k, f = 0, v
            while f / 2 >= self.phi:
                f /= 2
                k += 1
            self.kcap[r] = k
        self.kdim = max(self.kcap) + 1
        self.tally = np.zeros((len(self.rows), self.kdim, self.T), dtype=np.int64)
        self.drops = np.zeros((len(self.rows), self.kdim), dtype=np.int64)


class _RedrawNoSecond(Brute):
    """Redraw uses only the first candidate of each block."""

    def timer(self, cell, node, mu, sd):
        for a in range(1, 65536):
            z0, _ = self.o.normal_pair(self.o.block(self.seed, cell, node, a))
            t = _fma(z0, sd, mu)
            if t > 0:
                return t
            self.st["rejected"] += 1
        return mu


class _UnsortedTypes(Brute):
    def __init__(self, oracle, values, counts, prop, *a):
        super().__init__(oracle, values, counts, prop, *a)
        self.order = list(range(self.T))
        self.cum, c = [], 0.0
        for j in self.order:
            c = c + prop[j]
            self.cum.append(c)


SENSITIVE = ([256.0, 40.0, 24.0], [25, 25, 20], [0.3, 0.7], [14.0, 9.0], [9.0, 7.0], 45.0, 3.0, 17)


def test_brute_force_catches_planted_mistakes(oracle):
    """Plausible mistakes planted in the scalar restatement (swapped daughter
    timers, a >= cap rule, a one-candidate redraw, unsorted types) all
    disagree with the oracle on one small input that the faithful
    restatement reproduces exactly."""
    tally, drops, st = oracle.simulate(_problem(oracle, *SENSITIVE), 2, want_drops=True)
    good = Brute(oracle, *SENSITIVE).run()
    assert np.array_equal(tally, good.tally) and st == good.st
    assert st["rejected"] > 0
    for cls in (_SwappedDaughters, _GeCap, _RedrawNoSecond, _UnsortedTypes):
        bad = cls(oracle, *SENSITIVE).run()
        same = bad.tally.shape == tally.shape and np.array_equal(bad.tally, tally) and bad.st == st
        assert not same, cls.__name__


def _rows_ge_phi(values, counts, phi):
    return {v: c for v, c in zip(values, counts) if c > 0 and v >= phi}


def test_all_quiescent_is_identity(oracle):
    values = [1.0, 2.0, 2.5, 100.0, 1e4]
    counts = [5, 0, 7, 11, 3]
    p = _problem(oracle, values, counts, [0.6, 0.4], [-1.0, -2.0], [-1.0, 3.0], 500.0, 2.0, 4)
    tally, drops, st = oracle.simulate(p, 3, want_drops=True)
    nz = [i for i, c in enumerate(counts) if c > 0]
    assert st["divisions"] == 0 and st["dropped"] == 0
    assert st["lost"] == 5  # the row of value 1 < phi
    for r, i in enumerate(nz):
        expect = counts[i] if values[i] >= 2.0 else 0
        assert tally[r, 0].sum() == expect
        assert tally[r, 1:].sum() == 0


def test_t_max_zero_is_identity(oracle):
    values = [8.0, 3.0, 0.5]
    counts = [9, 4, 6]
    p = _problem(oracle, values, counts, [0.53, 0.29, 0.18], [48.33, 86.3, -1.0], [21.6, 26.8, -1.0], 0.0, 1.0, 5)
    tally, _, st = oracle.simulate(p, 2, want_drops=True)
    assert st["divisions"] == 0
    assert [int(tally[r, 0].sum()) for r in range(3)] == [9, 4, 0]
    assert st["lost"] == 6 and st["recorded"] == 13
    # per-type columns sum to the row count and follow the type draw
    assert tally[:, 0, :].sum() == 13


def test_sd_zero_closed_form(oracle):
    """sd = 0: every timer equals the mean.  Division completions of a cell
    fall at s_1 = mean*U + mean, s_{k+1} = s_k + mean; with n = #{k: s_k <=
    t_max}, a cell yields 2^n leaves at generation n when n <= kcap,
    otherwise 2^kcap dropped cells at generation kcap."""
    mean, t_max, phi, seed = 7.0, 45.0, 1.0, 31
    values = [64.0, 8.0, 1.5, 1.0]
    counts = [40, 30, 20, 10]
    p = _problem(oracle, values, counts, [1.0], [mean], [0.0], t_max, phi, seed)
    tally, drops, st = oracle.simulate(p, 2, want_drops=True)
    kcaps = [5, 2, 0, 0]   # 64/2^5 = 2 and 2/2 = 1 is not > phi
    exp_t = np.zeros_like(tally)
    exp_d = np.zeros_like(drops)
    exp_div = 0
    cell = 0
    for r, (v, n) in enumerate(zip(values, counts)):
        for _ in range(n):
            w = oracle.block(seed, cell, 0, 0)
            U = oracle.u01_open_closed(w[1], 0)
            s = mean * U + mean
            nd = 0
            while s <= t_max:
                nd += 1
                s = s + mean
            if nd == 0:
                if v >= phi:
                    exp_t[r, 0, 0] += 1
            elif nd <= kcaps[r]:
                exp_t[r, nd, 0] += 2**nd
                exp_div += 2**nd - 1
            else:
                exp_d[r, kcaps[r]] += 2 ** kcaps[r]
                exp_div += 2 ** kcaps[r] - 1
            cell += 1
    assert np.array_equal(tally, exp_t)
    assert np.array_equal(drops, exp_d)
    assert st["divisions"] == exp_div
    # t_max/mean = 6.43: a cell of the first row completes 5 or 6 divisions
    # in time; with 6 it reaches the cap and its 32 descendants are dropped
    assert set(np.nonzero(exp_t[0, :, 0])[0]) == {5} and exp_d[0, 5] > 0


def _random_problem(rng, n_rows=40, n_types=4, cells=4000):
    values = np.exp(rng.uniform(0, np.log(5e3), n_rows))
    counts = rng.multinomial(cells, np.full(n_rows, 1.0 / n_rows)).astype(np.uint64)
    counts[rng.integers(0, n_rows, 3)] = 0
    prop = rng.dirichlet(np.ones(n_types))
    prop[-1] = 1.0 - prop[:-1].sum()
    mean = rng.uniform(8, 40, n_types)
    sd = mean * rng.uniform(0.1, 1.2, n_types)
    mean[rng.integers(0, n_types)] = -1.0
    return values, counts, prop, mean, sd


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_invariants(oracle, seed):
    rng = np.random.default_rng(seed)
    values, counts, prop, mean, sd = _random_problem(rng)
    phi = float(rng.uniform(1.0, 30.0))
    t_max = 150.0
    p = _problem(oracle, values, counts, prop, mean, sd, t_max, phi, seed)
    tally, drops, st = oracle.simulate(p, 4, want_drops=True)
    assert st["cells"] == int(counts.sum())
    assert st["recorded"] + st["dropped"] + st["lost"] == st["cells"] + st["divisions"]
    assert tally.sum() == st["recorded"] and drops.sum() == st["dropped"]
    kdim = tally.shape[1]
    w = np.array([2.0**-k for k in range(kdim)])
    nz = [i for i in range(len(values)) if counts[i] > 0]
    lost = 0
    for r, i in enumerate(nz):
        kc = 0
        f = values[i]
        while f / 2 > phi:
            f /= 2
            kc += 1
        assert tally[r, kc + 1:].sum() == 0 and drops[r, kc + 1:].sum() == 0
        mass = float((tally[r].sum(axis=1) + drops[r]) @ w)
        if values[i] >= phi:
            assert mass == float(counts[i])
        else:
            assert tally[r].sum() == 0
            lost += int(counts[i]) - int(drops[r, 0])
    assert lost == st["lost"]
    # thread-count invariance
    t1, d1, s1 = oracle.simulate(p, 1, want_drops=True)
    assert np.array_equal(t1, tally) and np.array_equal(d1, drops) and s1 == st


def test_truncation_redraw_rate(oracle):
    """Each timer is redrawn while <= 0: rejections per accepted timer follow
    a geometric law with mean q / (1 - q), q = Phi(-mean/sd)."""
    mean, sd = 10.0, 10.0
    p = _problem(oracle, [4096.0], [20000], [1.0], [mean], [sd], 60.0, 1.0, 3)
    _, _, st = oracle.simulate(p, 4)
    q = 0.5 * math.erfc(mean / sd / math.sqrt(2.0))
    timers = st["cells"] + 2 * st["divisions"]
    rate = st["rejected"] / timers
    expect = q / (1.0 - q)
    se = math.sqrt(q / (1.0 - q) ** 2 / timers) * 3   # loose: tree draws are not independent
    assert abs(rate - expect) < 5 * se, (rate, expect)


def _numpy_simulation(values, counts, prop, mean, sd, t_max, phi, rng):
    """Independent restatement of S2-S6 with numpy's generator, vectorised
    over cells generation by generation.  Returns counts[k, type] summed over
    rows (rows >= phi only)."""
    if phi <= 0:
        phi = min(v for v, c in zip(values, counts) if c > 0)
    T = len(prop)
    order = sorted(range(T), key=lambda j: -prop[j])
    cum = np.cumsum([prop[j] for j in order])
    v = np.repeat(np.asarray(values, float), np.asarray(counts, np.int64))
    n = len(v)
    js = np.minimum(np.searchsorted(cum, rng.random(n), side="right"), T - 1)
    typ = np.array(order)[js]
    mu = np.asarray(mean, float)[typ]
    sg = np.asarray(sd, float)[typ]

    def trunc(m, s):
        t = m + s * rng.standard_normal(len(m))
        bad = ~(t > 0)
        while bad.any():
            t[bad] = m[bad] + s[bad] * rng.standard_normal(int(bad.sum()))
            bad = ~(t > 0)
        return t

    out = np.zeros((64, T), dtype=np.int64)
    quiet = mu < 0
    np.add.at(out, (0, typ[quiet & (v >= phi)]), 1)
    live = ~quiet
    v, typ, mu, sg = v[live], typ[live], mu[live], sg[live]
    tau = trunc(mu, sg)
    U = 1.0 - rng.random(len(v))
    s = tau * U + tau
    k = 0
    while len(v):
        out_time = s > t_max
        rec = out_time & (v >= phi)
        np.add.at(out, (k, typ[rec]), 1)
        can = ~out_time & (v / 2 > phi)
        v, typ, mu, sg, s = v[can], typ[can], mu[can], sg[can], s[can]
        v = np.repeat(v / 2, 2)
        typ, mu, sg, s = (np.repeat(a, 2) for a in (typ, mu, sg, s))
        s = s + trunc(mu, sg)
        k += 1
    return out


def test_distribution_matches_independent_simulation(oracle):
    """Means over 10 seeds of the recorded counts per (generation, type):
    oracle vs the numpy restatement, z-scores from the per-seed spread."""
    from procell import workloads

    w = workloads.small_histogram(0)
    n_seeds = 10
    kdim = None
    a, b = [], []
    for s in range(n_seeds):
        tally, _, _ = oracle.simulate(oracle.Problem.from_workload(w, s), 4)
        kdim = tally.shape[1]
        a.append(tally.sum(axis=0))
        rng = np.random.default_rng([s, 77])
        b.append(_numpy_simulation(w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max, w.phi, rng)[:kdim])
    a, b = np.array(a, float), np.array(b, float)
    assert b[:, kdim:].sum() == 0 if b.shape[1] > kdim else True
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / n_seeds + b.var(0, ddof=1) / n_seeds)
    big = (ma + mb) > 200
    assert big.sum() >= 8
    z = np.abs(ma - mb)[big] / np.maximum(se[big], 1e-9)
    assert z.max() < 5.0, z
    # totals per type
    ta, tb = a.sum(1), b.sum(1)
    zt = np.abs(ta.mean(0) - tb.mean(0)) / np.sqrt(ta.var(0, ddof=1) / n_seeds + tb.var(0, ddof=1) / n_seeds)
    assert zt.max() < 5.0, zt
