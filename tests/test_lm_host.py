"""Levenberg-Marquardt host algebra on the CPU: the batched SPD solve, the
Gauss-Newton terms of the normalised-L2 loss against torch autograd on an
analytic stand-in for the relaxed tally, and convergence of the damped
steps on it."""
import numpy as np
import pytest
import torch

from procell import calibrate as cal


def test_spd_solve_matches_numpy():
    rng = np.random.default_rng(3)
    n, P = 7, 9
    X = rng.normal(size=(n, P, P))
    A = X @ X.transpose(0, 2, 1) + 0.1 * np.eye(P)
    b = rng.normal(size=(n, P))
    x = cal.spd_solve(torch.tensor(A), torch.tensor(b)).numpy()
    np.testing.assert_allclose(x, np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-9, atol=1e-12)


# Analytic model: per set i, bin b, type j the tally is
#   h = p_j w[i,b,j] exp(-(c_b - mean_j)^2 / (2 sd_j^2))        (proliferating)
#   h = p_j w[i,b,j]                                             (quiescent)
# so h of type j depends on type j's parameters only, as the relaxed tally.
B, T = 12, 3
CENTERS = torch.linspace(20.0, 90.0, B, dtype=torch.float64)
W = torch.tensor(np.random.default_rng(9).uniform(0.5, 2.0, size=(2, B, T)))
PROLIF = torch.tensor([[True, True, False], [True, False, False]])
REAL = torch.tensor([[True, True, True], [True, True, False]])


def _tally(mean, sd, p):
    z = (CENTERS[None, :, None] - mean[:, None, :]) / sd[:, None, :]
    f = torch.where(PROLIF[:, None, :], torch.exp(-0.5 * z * z), torch.ones_like(z)) * W
    Hj = p[:, None, :] * f
    MJ = torch.where(PROLIF[:, None, :], Hj * z / sd[:, None, :], torch.zeros_like(Hj))
    SJ = torch.where(PROLIF[:, None, :], Hj * z * z / sd[:, None, :], torch.zeros_like(Hj))
    return Hj, MJ, SJ


def _params(phi):
    mean = torch.where(PROLIF, torch.exp(phi[..., 0]), torch.ones_like(phi[..., 0]))
    sd = torch.where(PROLIF, torch.exp(phi[..., 1]), torch.ones_like(phi[..., 1]))
    logits = torch.where(REAL, phi[..., 2], torch.full_like(phi[..., 2], -torch.inf))
    return mean, sd, torch.softmax(logits, -1)


def _active():
    last = REAL.sum(1, keepdim=True) - 1
    free = REAL & (torch.arange(T)[None, :] < last)
    return torch.stack([PROLIF, PROLIF, free], -1)


def _residual(phi, q):
    Hj, _, _ = _tally(*_params(phi))
    H = Hj.sum(-1)
    return H / H.sum(-1, keepdim=True) - q


def _phi(mean, sd, p):
    return torch.stack([torch.log(mean), torch.log(sd), torch.log(p)], -1)


PHI_TRUE = _phi(torch.tensor([[40.0, 70.0, 1.0], [55.0, 1.0, 1.0]], dtype=torch.float64),
                torch.tensor([[8.0, 15.0, 1.0], [12.0, 1.0, 1.0]], dtype=torch.float64),
                torch.tensor([[0.5, 0.3, 0.2], [0.6, 0.4, 1e-300]], dtype=torch.float64))


def test_normal_equations_match_autograd():
    q = _residual(PHI_TRUE, 0.0).detach() + 0.01  # any target
    phi = PHI_TRUE + 0.1 * torch.tensor(np.random.default_rng(1).normal(size=PHI_TRUE.shape))
    act = _active()
    mean, sd, p = _params(phi)
    loss, A, g = cal.normal_equations(*_tally(mean, sd, p), mean, sd, p, q, act)
    r = _residual(phi, q)
    np.testing.assert_allclose(loss.numpy(), (r * r).sum(-1).numpy(), rtol=1e-12)
    for i in range(2):
        # J of r_i in the active parameters of set i, by autograd
        def ri(x, i=i):
            full = phi.clone()
            full[i] = x.reshape(T, 3)
            return _residual(full, q)[i]

        J = torch.autograd.functional.jacobian(ri, phi[i].reshape(-1).clone())
        mask = act[i].reshape(-1).to(torch.float64)
        J = J * mask[None, :]
        np.testing.assert_allclose(g[i].numpy(), (J.T @ r[i]).numpy(), rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(A[i].numpy(), (J.T @ J).numpy(), rtol=1e-9, atol=1e-15)


def test_damped_steps_converge_on_exact_target():
    q = _residual(PHI_TRUE, 0.0).detach()
    act = _active()
    start = PHI_TRUE + torch.tensor([[[0.12, -0.1, 0.3], [-0.1, 0.1, -0.2], [0.0, 0.0, 0.0]],
                                     [[-0.12, 0.12, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]], dtype=torch.float64)
    phi = start.clone()
    lam = torch.full((2,), 1e-3, dtype=torch.float64)

    def ev(x):
        mean, sd, p = _params(x)
        return cal.normal_equations(*_tally(mean, sd, p), mean, sd, p, q, act)

    loss, A, g = ev(phi)
    losses = [loss]
    for _ in range(40):
        trial = phi + cal.lm_step(A, g, lam, act).reshape(2, T, 3)
        l2, A2, g2 = ev(trial)
        ok = l2 < loss
        phi = torch.where(ok[:, None, None], trial, phi)
        loss = torch.where(ok, l2, loss)
        A = torch.where(ok[:, None, None], A2, A)
        g = torch.where(ok[:, None], g2, g)
        lam = torch.where(ok, lam * 0.2, lam * 4.0).clamp(1e-9, 1e9)
        losses.append(loss)
    assert all((b <= a).all() for a, b in zip(losses, losses[1:]))
    assert float(loss.max()) < 1e-20
    m, s, p = _params(phi)
    mt, st, pt = _params(PHI_TRUE)
    np.testing.assert_allclose(m[PROLIF].numpy(), mt[PROLIF].numpy(), rtol=1e-6)
    np.testing.assert_allclose(s[PROLIF].numpy(), st[PROLIF].numpy(), rtol=1e-6)
    np.testing.assert_allclose(p[REAL].numpy(), pt[REAL].numpy(), atol=1e-8)


def test_step_holds_inactive_parameters_and_limits_size():
    act = _active()
    A = torch.eye(3 * T, dtype=torch.float64).expand(2, -1, -1).clone() * 1e-6
    g = torch.ones((2, 3 * T), dtype=torch.float64)
    x = cal.lm_step(A, g, torch.full((2,), 1e-3, dtype=torch.float64), act).reshape(2, T, 3)
    assert (x[~act] == 0).all()
    assert (x[act].abs() <= cal.MAX_LOG_STEP).all() and (x[act] == -cal.MAX_LOG_STEP).all()


def test_step_holds_parameters_the_loss_ignores():
    """A zero Jacobian column (e.g. the proportions of an all-quiescent set,
    which the normalised histogram does not see) gives a zero step, not 0/0."""
    act = _active()
    A = torch.zeros((2, 3 * T, 3 * T), dtype=torch.float64)
    A[0, 0, 0] = 2.0
    g = torch.zeros((2, 3 * T), dtype=torch.float64)
    g[0, 0] = 1.0
    x = cal.lm_step(A, g, torch.full((2,), 1e-9, dtype=torch.float64), act)
    assert torch.isfinite(x).all()
    assert x[0, 0] < 0 and (x[0, 1:] == 0).all() and (x[1] == 0).all()


class _StandInPlan:
    """CPU stand-in for a batch relaxed plan: the sets' shapes and offsets
    as the library lays them out, and a 'Jacobian pass' that writes known
    values (h = p_j c_e, dh/dmean = c_e / 7, dh/dsd = -c_e / 11 for each
    entry e of type j) and records the theta it was given."""

    def __init__(self, problems, replicas=1):
        self.shapes, self.h_base = [], []
        base = 0
        for _ in range(replicas):
            for p in problems:
                r, k, t = p.tally_shape()
                self.shapes.append((int(r), int(k), int(t)))
                self.h_base.append(base)
                base += int(r) * int(k) * int(t)
        self.h_total = base
        self.c = torch.tensor(np.random.default_rng(2).uniform(0.5, 1.5, base), dtype=torch.float64)
        self.thetas = []
        self.cut_sets = set()   # plan sets whose passes report a cut walk
        self.cut_after = 0      # ... from this call on (0: every call)

    def set_cuts(self, cuts):
        cuts.zero_()
        if len(self.thetas) > self.cut_after:
            for i in self.cut_sets:
                cuts[i] = 3

    def jacobian(self, theta, h, jac):
        self.thetas.append(theta.clone())
        k = 0
        for (r, kd, t), b in zip(self.shapes, self.h_base):
            th = theta[k:k + 3 * t].reshape(3, t)
            c = self.c[b:b + r * kd * t].reshape(r, kd, t)
            h[b:b + r * kd * t] = (th[2] * c).reshape(-1)
            jac[0, b:b + r * kd * t] = (c / 7).reshape(-1)
            jac[1, b:b + r * kd * t] = (-c / 11).reshape(-1)
            k += 3 * t


@pytest.mark.parametrize("replicas", [1, 2])
def test_batch_lm_index_maps(replicas):
    """BatchLM's entry -> (set, bin, type) and parameter -> theta maps, on
    the CPU with the stand-in plan, against a loop over the output bins:
    the start theta handed to the pass, the loss and the gradient J^T r."""
    from procell import Problem, workloads

    gen, start = workloads.calibration_batch(7, 3, 800)
    probs = [Problem.from_workload(w, 10 + i) for i, w in enumerate(start)]
    targets = [(w.values[w.counts > 0], w.counts[w.counts > 0]) for w in gen]
    plan = _StandInPlan(probs, replicas)
    lm = cal.BatchLM(plan, probs, targets, device="cpu", replicas=replicas)
    lm.start()
    th = plan.thetas[0].numpy()
    want = np.concatenate([np.concatenate([w.mean, w.sd, w.proportion]) for w in start] * replicas)
    np.testing.assert_allclose(th, want, rtol=1e-12)
    c = plan.c.numpy()
    for i, (p, w, (tv, tc)) in enumerate(zip(probs, start, targets)):
        r, kd, t = plan.shapes[i]
        bins, idx = cal.output_bins(p, kd)
        H = np.zeros(len(bins))
        M = np.zeros((len(bins), t))
        Sg = np.zeros((len(bins), t))
        for rep in range(replicas):
            b0 = plan.h_base[rep * len(probs) + i]
            cc = c[b0:b0 + r * kd * t].reshape(r, kd, t)
            for row in range(r):
                for k in range(kd):
                    if idx[row, k] >= 0:
                        H[idx[row, k]] += (w.proportion * cc[row, k]).sum()
                        M[idx[row, k]] += cc[row, k] / 7
                        Sg[idx[row, k]] += -cc[row, k] / 11
        q = cal.target_histogram(bins, tv, tc)
        q = q / q.sum()
        res = H / H.sum() - q
        assert abs(float(lm.loss[i]) - (res * res).sum()) <= 1e-12 * (res * res).sum()
        # d loss / d log mean_j = 2 sum_b r_b (M_bj mean_j - H_b/S sum_b' M_b'j mean_j) / S
        S = H.sum()
        pr = w.mean >= 0
        for j in np.nonzero(pr)[0]:
            dm = (M[:, j] * w.mean[j] - H / S * (M[:, j] * w.mean[j]).sum()) / S
            ds = (Sg[:, j] * w.sd[j] - H / S * (Sg[:, j] * w.sd[j]).sum()) / S
            np.testing.assert_allclose(float(lm.g[i, 3 * j]), (res * dm).sum(), rtol=1e-9, atol=1e-18)
            np.testing.assert_allclose(float(lm.g[i, 3 * j + 1]), (res * ds).sum(), rtol=1e-9, atol=1e-18)


def _lm_instance(replicas=1, n=3):
    from procell import Problem, workloads

    gen, start = workloads.calibration_batch(7, n, 800)
    probs = [Problem.from_workload(w, 10 + i) for i, w in enumerate(start)]
    targets = [(w.values[w.counts > 0], w.counts[w.counts > 0]) for w in gen]
    return probs, targets, _StandInPlan(probs, replicas)


def test_batch_lm_rejects_trials_with_cut_walks():
    """A set whose trial passes cut a walk short keeps its point (the trial
    is rejected on the device and counted); the other sets move on."""
    probs, targets, plan = _lm_instance(replicas=2)
    plan.cut_sets = {len(probs) + 1}  # copy 1 of set 1
    plan.cut_after = 1                # the start pass is whole
    lm = cal.BatchLM(plan, probs, targets, device="cpu", replicas=2)
    lm.start()
    phi0 = lm.phi.clone()
    lm.step(3)
    assert torch.equal(lm.phi[1], phi0[1])
    assert lm.cut_trials.tolist() == [0, 3, 0]
    assert not torch.equal(lm.phi[0], phi0[0]) and not torch.equal(lm.phi[2], phi0[2])
    lm.thetas()  # the current points come from whole passes


def test_batch_lm_cut_start_raises_until_replaced():
    """A start point whose pass was cut counts as infinitely bad: the first
    whole trial replaces it; while none has, thetas() raises E_DEPTH."""
    from procell import E_DEPTH, ProcellError

    probs, targets, plan = _lm_instance()
    plan.cut_sets = {2}
    lm = cal.BatchLM(plan, probs, targets, device="cpu")
    lm.start()
    assert torch.isinf(lm.loss[2]) and bool(lm.cut[2])
    with pytest.raises(ProcellError) as e:
        lm.thetas()
    assert e.value.code == E_DEPTH
    lm.step(1)  # still cut: rejected, still raises
    with pytest.raises(ProcellError):
        lm.thetas()
    plan.cut_sets = set()
    lm.step(1)
    assert not lm.cut.any() and torch.isfinite(lm.loss).all()
    lm.thetas()


def test_batch_lm_projects_the_start():
    """A start with sd > 0.8 mean is handed to the pass at sd = 0.8 mean."""
    probs, targets, plan = _lm_instance()
    w = probs[0]
    j = int(np.nonzero(np.asarray(w.mean) >= 0)[0][0])
    sd = np.array(w.sd, dtype=float)
    sd[j] = 2.0 * w.mean[j]
    from procell import Problem

    probs[0] = Problem(w.values, w.counts, w.proportion, w.mean, sd, w.params.t_max, w.params.phi, w.params.seed)
    lm = cal.BatchLM(plan, probs, targets, device="cpu")
    lm.start()
    T = plan.shapes[0][2]
    th = plan.thetas[0][:3 * T].reshape(3, T)
    np.testing.assert_allclose(float(th[1, j]), 0.8 * w.mean[j], rtol=1e-12)
