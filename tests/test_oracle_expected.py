"""CPU: the oracle's plain method against its closed-form expected histogram.

For an initial cell of type j (timers Normal(mean_j, sd_j) truncated to > 0)
in a row with value v and generation cap K, division k + 1 completes at
D_{k+1} = tau_0 (1 + U) + tau_1 + ... + tau_k, and the cell's 2^k
descendants at generation k are recorded iff D_k <= t_max < D_{k+1} and
k <= K (generation 0 only when v >= phi).  Hence

    E[count(k, j)] = p_j sum_b n_b 2^k (F_k(t_max) - F_{k+1}(t_max)) [k <= K_b],

F_0 = 1, F_k the CDF of D_k: F_1 by quadrature over U, the later ones by
numerical convolution with the truncated-normal density on a fine time grid
(scipy), independent of the oracle's code and random stream.  Leaves of one
tree are correlated, so the test uses the spread over 20 seeds:
z = (mean over seeds - E) / (sd over seeds / sqrt(20)) per (generation,
type) with an expected count above 50, sum z^2 against chi^2 with that many
degrees of freedom, and every |z| < 5.
"""
import math

import numpy as np
import pytest
from scipy import signal
from scipy import stats as sst


def _cdfs(mean, sd, t_max, kmax, dt=0.004, n_u=800):
    t = np.arange(0.0, t_max + dt / 2, dt)
    a = -mean / sd
    z0 = 1.0 - sst.norm.cdf(a)
    F_tau = lambda x: np.where(x > 0, (sst.norm.cdf((x - mean) / sd) - sst.norm.cdf(a)) / z0, 0.0)
    f_tau = np.where(t > 0, sst.norm.pdf((t - mean) / sd) / sd / z0, 0.0)
    us = (np.arange(n_u) + 0.5) / n_u
    F = [np.ones_like(t)]
    F1 = np.zeros_like(t)
    for u in us:
        F1 += F_tau(t / (1.0 + u))
    F.append(F1 / n_u)
    w = f_tau * dt
    w[0] *= 0.5  # trapezoid ends (f_tau(0) = 0 anyway for mean/sd >> 0)
    for _ in range(2, kmax + 2):
        F.append(signal.fftconvolve(F[-1], w)[: len(t)])
    return np.array([Fk[-1] for Fk in F])  # F_k(t_max), k = 0 .. kmax + 1


def _expected(w, phi, kdim):
    vals = np.asarray(w.values, float)
    cnts = np.asarray(w.counts, float)
    T = len(w.proportion)
    E = np.zeros((kdim, T))
    caps = []
    for v in vals:
        k, f = 0, v
        while f / 2 > phi:
            f /= 2
            k += 1
        caps.append(k)
    caps = np.array(caps)
    for j in range(T):
        p = w.proportion[j]
        if w.mean[j] < 0:
            E[0, j] = p * cnts[vals >= phi].sum()
            continue
        Fk = _cdfs(w.mean[j], w.sd[j], w.t_max, kdim)
        for k in range(kdim):
            sel = (caps >= k) & ((vals >= phi) if k == 0 else True)
            E[k, j] = p * cnts[sel].sum() * 2.0**k * (Fk[k] - Fk[k + 1])
    return E


def _workload(which):
    from procell import workloads

    if which == "hist64":
        return workloads.small_histogram(0)
    # the 1024-bin, five-type shape (phi 50: rows below it lost at generation 0,
    # t_max 240) at 5 % of its cells
    return workloads.scaled(workloads.large_histogram(0), 0.05)


@pytest.mark.parametrize("which", ["hist64", "hist1024_5pct"])
def test_expected_histogram_closed_form(oracle, which):
    from procell import workloads

    w = _workload(which)
    n_seeds = 20
    obs = []
    for s in range(n_seeds):
        tally, _, _ = oracle.simulate(oracle.Problem.from_workload(w, s), 4)
        obs.append(tally.sum(axis=0))
    obs = np.array(obs, float)
    kdim = obs.shape[1]
    phi = oracle.resolve_phi(oracle.Problem.from_workload(w, 0))
    E = _expected(w, phi, kdim)
    m = obs.mean(0)
    se = obs.std(0, ddof=1) / math.sqrt(n_seeds)
    big = E > 50
    assert big.sum() >= 8
    # the quiescent column is deterministic given the type draws: its spread is the binomial one
    z = (m[big] - E[big]) / np.maximum(se[big], 1e-9)
    assert np.abs(z).max() < 5.0, (z, m[big], E[big])
    chi2 = float((z * z).sum())
    p = 1.0 - sst.chi2.cdf(chi2, df=int(big.sum()))
    assert p > 1e-4, (chi2, int(big.sum()), p)
    # the closed form has power: a 2 % shift of one type's mean is detected
    w2 = _workload(which)
    w2.mean = w2.mean.copy()
    w2.mean[0] *= 1.02
    E2 = _expected(w2, phi, kdim)
    z2 = (m[big] - E2[big]) / np.maximum(se[big], 1e-9)
    assert float((z2 * z2).sum()) > chi2 + 50.0
