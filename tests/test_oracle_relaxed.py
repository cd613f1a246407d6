"""CPU checks that pin the oracle's relaxed estimator and its gradient.

The relaxed estimator (oracle_relaxed) weights every type by its
proportion, uses the plain method's timers mean + sd z (z the first
candidate > 0 of the plain stream, redraw streams included), and replaces
the division test by sigmoid((t_max - s)/beta).
Pinned here by what does not depend on its code:
  * beta -> 0 with one type: the relaxed tally is the plain oracle's tally
    of the same seed, entry for entry, with sd = mean / 2 so that the redraw
    streams are exercised in most trees;
  * central finite differences of the forward with common z (the same
    seed) against the adjoint gradient, for every mean, sd and (along
    directions that keep the sum 1) proportion;
  * sd = 0, where every timer is the mean and the expected tree has a
    closed form per initial cell (2^k nodes at generation k, each with
    the product of its ancestors' sigmoids), computed here with numpy
    from the cells' phases, and its derivative by finite differences;
  * small beta with sd << mean (no candidate is ever <= 0), where the
    estimator is an unbiased Monte Carlo estimate of the Gaussian-sum
    expected histogram
    2^k (P(s_{k-1} <= t_max) - P(s_k <= t_max)), s_k | U ~ N(m (1+U+k),
    sd^2 ((1+U)^2 + k)): histogram and gradient agree within their
    sampling error.
"""
import math

import numpy as np
import pytest

from scipy import stats as sst


def _prob(oracle, values, counts, prop, mean, sd, t_max, phi, seed):
    return oracle.Problem(values, counts, prop, mean, sd, t_max, phi, seed)


def test_tiny_beta_one_type_is_the_plain_tally(oracle):
    """With beta -> 0 the sigmoid is the plain method's hard test and the
    timers are the plain method's own, so for one type the relaxed tally is
    the plain tally of the same seed."""
    values, counts = [2.0**12, 3000.0, 700.0, 40.0, 9.0], [300, 200, 500, 100, 50]
    for mean, sd, seed in ((20.0, 10.0, 3), (30.0, 15.0, 4), (25.0, 5.0, 5)):
        p = _prob(oracle, values, counts, [1.0], [mean], [sd], 150.0, 10.0, seed)
        tally, _, st = oracle.simulate(p)
        if sd / mean >= 0.5:
            assert st["rejected"] > 0  # redraw streams exercised
        h, _ = oracle.relaxed(p, 1e-9)
        assert h.shape == tally.shape
        assert np.array_equal(h, tally.astype(np.float64)), (mean, sd, np.abs(h - tally).max())


def _L(oracle, args, g, beta):
    h, _ = oracle.relaxed(_prob(oracle, *args), beta)
    return float((g * h).sum())


def test_adjoint_matches_finite_differences(oracle):
    values, counts = [64.0, 20.0, 5.0, 2.5], [60, 40, 30, 20]
    prop, mean, sd = [0.45, 0.35, 0.2], [14.0, 22.0, -1.0], [4.0, 7.0, -1.0]
    t_max, phi, seed, beta = 70.0, 1.0, 11, 2.0
    h, _ = oracle.relaxed(_prob(oracle, values, counts, prop, mean, sd, t_max, phi, seed), beta)
    g = np.random.default_rng(3).normal(size=h.shape)
    _, grad = oracle.relaxed(_prob(oracle, values, counts, prop, mean, sd, t_max, phi, seed), beta, g)
    base = [values, counts, prop, mean, sd, t_max, phi, seed]
    for which, idx in ((3, 0), (3, 1), (4, 0), (4, 1)):
        d = 1e-5 * abs(base[which][idx])
        up = [list(x) if isinstance(x, list) else x for x in base]
        dn = [list(x) if isinstance(x, list) else x for x in base]
        up[which][idx] += d
        dn[which][idx] -= d
        fd = (_L(oracle, up, g, beta) - _L(oracle, dn, g, beta)) / (2 * d)
        ad = grad[which - 3, idx]
        assert abs(fd - ad) <= 1e-6 * max(1.0, abs(ad)), (which, idx, fd, ad)
    # proportions: along p_i - p_k (the sum stays 1)
    for i, k in ((0, 1), (0, 2), (1, 2)):
        d = 1e-6
        up = [list(x) if isinstance(x, list) else x for x in base]
        dn = [list(x) if isinstance(x, list) else x for x in base]
        up[2][i] += d
        up[2][k] -= d
        dn[2][i] -= d
        dn[2][k] += d
        fd = (_L(oracle, up, g, beta) - _L(oracle, dn, g, beta)) / (2 * d)
        ad = grad[2, i] - grad[2, k]
        assert abs(fd - ad) <= 1e-6 * max(1.0, abs(ad)), (i, k, fd, ad)
    # quiescent types have no timer parameters
    assert grad[0, 2] == 0.0 and grad[1, 2] == 0.0


def _closed_form_sd0(oracle, values, counts, m, t_max, phi, seed, beta):
    """Expected relaxed tally (one type, sd = 0) from the cells' phases."""
    sig = lambda x: 1.0 / (1.0 + np.exp(-x))
    nz = [(v, c) for v, c in zip(values, counts) if c > 0]
    caps = []
    for v, _ in nz:
        k, f = 0, v
        while f / 2 > phi:
            f /= 2
            k += 1
        caps.append(k)
    kdim = max(caps) + 1
    h = np.zeros((len(nz), kdim))
    cell = 0
    for r, (v, c) in enumerate(nz):
        for _ in range(c):
            U = oracle.u01_open_closed(oracle.block(seed, cell, 0, 0)[1], 0)
            W = 1.0
            for k in range(caps[r] + 1):
                s = m * (1.0 + U) + k * m
                d = sig((t_max - s) / beta)
                if k > 0 or v >= phi:
                    h[r, k] += 2.0**k * W * (1.0 - d)
                W *= d
            cell += 1
    return h


def test_sd_zero_closed_form_and_derivative(oracle):
    values, counts = [256.0, 40.0, 3.0, 1.5], [50, 40, 30, 20]
    m, t_max, phi, seed, beta = 9.0, 60.0, 2.0, 5, 1.5
    p = _prob(oracle, values, counts, [1.0], [m], [0.0], t_max, phi, seed)
    h, _ = oracle.relaxed(p, beta)
    ref = _closed_form_sd0(oracle, values, counts, m, t_max, phi, seed, beta)
    assert h.shape[:2] == ref.shape
    assert np.allclose(h[:, :, 0], ref, rtol=1e-12, atol=1e-12)
    # the adjoint against finite differences of the closed form
    g = np.random.default_rng(1).normal(size=h.shape)
    _, grad = oracle.relaxed(p, beta, g)
    d = 1e-5
    up = _closed_form_sd0(oracle, values, counts, m + d, t_max, phi, seed, beta)
    dn = _closed_form_sd0(oracle, values, counts, m - d, t_max, phi, seed, beta)
    fd = float((g[:, :, 0] * (up - dn)).sum() / (2 * d))
    assert abs(fd - grad[0, 0]) <= 1e-6 * max(1.0, abs(fd)), (fd, grad[0, 0])


def _gauss_sum_expected(m, s, t_max, kdim, n_u=400):
    """E[h_k] per initial cell, untruncated timers, hard test (beta -> 0),
    averaged over the phase U ~ U(0, 1) by the midpoint rule."""
    us = (np.arange(n_u) + 0.5) / n_u
    out = np.zeros(kdim)
    for U in us:
        cdf = []
        for k in range(kdim):
            mu = m * (1.0 + U + k)
            sig = s * math.sqrt((1.0 + U) ** 2 + k)
            cdf.append(sst.norm.cdf((t_max - mu) / sig))
        for k in range(kdim):
            prev = 1.0 if k == 0 else cdf[k - 1]
            out[k] += 2.0**k * (prev - cdf[k]) / n_u
    return out


def test_small_beta_matches_gaussian_sum_closed_form(oracle):
    """One type, one row deep enough that no cell reaches the cap: the
    relaxed histogram and its mean / sd gradient agree with the closed form
    within 5 standard errors, estimated from 20 independent seeds."""
    m, s, t_max, beta, n, reps = 20.0, 3.0, 75.0, 0.1, 1000, 20
    v, phi = 2.0**20, 1.0
    kdim = 8  # m (1+U+k) <= 75 needs k <= 2.75; rare fast chains reach 5-7
    w = np.array([0.3, -0.7, 1.1, 0.4, -0.2, 0.1, 0.05, -0.02])
    hs, gm, gs = [], [], []
    for seed in range(reps):
        p = _prob(oracle, [v], [n], [1.0], [m], [s], t_max, phi, 100 + seed)
        g = None
        h, _ = oracle.relaxed(p, beta)
        assert h[0, kdim:, 0].sum() < 1e-6 * n
        g = np.zeros_like(h)
        g[0, :kdim, 0] = w
        _, grad = oracle.relaxed(p, beta, g)
        hs.append(h[0, :kdim, 0] / n)
        gm.append(grad[0, 0] / n)
        gs.append(grad[1, 0] / n)
    hs = np.array(hs)
    expect = _gauss_sum_expected(m, s, t_max, kdim)
    mean_h, se_h = hs.mean(0), hs.std(0, ddof=1) / math.sqrt(reps)
    for k in range(kdim):
        assert abs(mean_h[k] - expect[k]) <= 5 * se_h[k] + 1e-3, (k, mean_h[k], expect[k], se_h[k])
    for est, which in ((np.array(gm), 0), (np.array(gs), 1)):
        delta = 1e-3
        mp, sp = (m + delta, s) if which == 0 else (m, s + delta)
        mm, sm = (m - delta, s) if which == 0 else (m, s - delta)
        d_cf = (w @ _gauss_sum_expected(mp, sp, t_max, kdim) - w @ _gauss_sum_expected(mm, sm, t_max, kdim)) / (2 * delta)
        se = est.std(ddof=1) / math.sqrt(reps)
        assert abs(est.mean() - d_cf) <= 5 * se + 1e-3, (which, est.mean(), d_cf, se)
        assert se < 0.1 * abs(d_cf)  # the check has power
