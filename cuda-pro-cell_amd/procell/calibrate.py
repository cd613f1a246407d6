"""Calibration binding: the relaxed estimator under torch.autograd.

RelaxedTally is a torch.autograd.Function over the C ABI: its forward returns
the relaxed expected tally h[row, generation, type] of the plan's histogram
for theta = [mean; sd; proportion] (3 x n_types, float64, on the GPU).  When
theta needs a gradient and the plan walks every cell, the forward runs the
Jacobian pass (procell_relaxed_jacobian: h, dh/dmean and dh/dsd in one walk)
and the backward contracts it with the incoming cotangent dL/dh; otherwise
the forward runs procell_relaxed_forward and the backward hands dL/dh to the
library's adjoint pass (procell_relaxed_backward), which recomputes every
subtree from the random stream -- with a cell stride, on the cells the stride
selects when the backward runs.  Both give the same vector-Jacobian product
of the cells walked.  The helpers below
merge h into the plain method's output bins (value / 2^k) and form the
squared L2 distance between normalised histograms.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Problem, Relaxed


class RelaxedTally(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, plan):
        theta = theta.contiguous()
        ctx.plan = plan
        # one walk for h and its Jacobian (single-set plans without a stride):
        # every entry depends on its own type's parameters only
        ctx.jac = ctx.needs_input_grad[0] and len(plan.shape) == 3 and plan.stride == (1, 0)
        if ctx.jac:
            # the pass at unit proportions: h is linear in p_j and the pass
            # scales its records by p_j only, so it gives
            # [3, ...] = dh/dmean / p | dh/dsd / p | dh/dp exactly (also at p_j = 0)
            unit = torch.cat([theta[:2], torch.ones_like(theta[2:])])
            buf = torch.empty((3,) + tuple(plan.shape), dtype=torch.float64, device=theta.device)
            plan.jacobian(unit, buf[2], buf[:2])
            ctx.save_for_backward(theta, buf)
            return buf[2] * theta[2]
        h = torch.empty(plan.shape, dtype=torch.float64, device=theta.device)
        plan.forward(theta, h)
        ctx.save_for_backward(theta)
        return h

    @staticmethod
    def backward(ctx, g):
        if ctx.jac and ctx.plan.stride == (1, 0):
            theta, buf = ctx.saved_tensors
            grad = (buf * g).sum((1, 2))
            grad[:2] *= theta[2]
            return grad, None
        theta = ctx.saved_tensors[0]
        grad = torch.empty_like(theta)
        ctx.plan.backward(theta, g.contiguous(), grad)
        return grad, None


def relaxed_tally(plan: Relaxed, theta: torch.Tensor) -> torch.Tensor:
    """h[n_rows, kdim, n_types] of the relaxed estimator, differentiable in theta."""
    return RelaxedTally.apply(theta, plan)


def theta_of(prob: Problem) -> np.ndarray:
    """[3, n_types] = mean, sd, proportion of a problem's types (file order)."""
    return np.stack([prob.mean, prob.sd, prob.proportion]).astype(np.float64)


def output_bins(prob: Problem, kdim: int):
    """Output bin values (ascending) and, for every (nonzero row, generation)
    of the tally, the index of its bin value / 2^k, or -1 where no cell of the
    row can be recorded (beyond the row's generation cap, or generation 0 of a
    row below phi)."""
    phi = prob.resolve_phi()
    rows = [float(v) for v, c in zip(prob.values, prob.counts) if c > 0]
    vals = []
    for v in rows:
        f = v
        while f >= phi:
            vals.append(f)
            f = f / 2
    bins = np.array(sorted(set(vals)), dtype=np.float64)
    where = {float(b): i for i, b in enumerate(bins)}
    idx = np.full((len(rows), kdim), -1, dtype=np.int64)
    for r, v in enumerate(rows):
        f, k = v, 0
        while True:
            if k > 0 or f >= phi:
                idx[r, k] = where[f]
            if not (f / 2 > phi) or k + 1 >= kdim:
                break
            f = f / 2
            k += 1
    return bins, idx


class BinMap:
    """Gather/scatter indices that merge a relaxed tally into output bins,
    prepared once on the device (no host synchronisation per call)."""

    def __init__(self, prob: Problem, kdim: int, device="cuda"):
        self.bins, idx = output_bins(prob, kdim)
        flat = idx.reshape(-1)
        pos = np.nonzero(flat >= 0)[0]
        self.n_bins = len(self.bins)
        self.pos = torch.tensor(pos, dtype=torch.int64, device=device)
        self.bin = torch.tensor(flat[pos], dtype=torch.int64, device=device)

    def __call__(self, h: torch.Tensor) -> torch.Tensor:
        """Sum of h over types and over the (row, generation) entries of each bin."""
        flat = h.sum(-1).reshape(-1)
        out = torch.zeros(self.n_bins, dtype=h.dtype, device=h.device)
        return out.index_add(0, self.bin, flat.index_select(0, self.pos))


def binned(h: torch.Tensor, idx: torch.Tensor, n_bins: int) -> torch.Tensor:
    """Sum of h over types and over the (row, generation) entries of each bin
    (idx from output_bins; -1 entries are skipped)."""
    flat = h.sum(-1).reshape(-1)
    sel = idx.reshape(-1)
    keep = sel >= 0
    out = torch.zeros(n_bins, dtype=h.dtype, device=h.device)
    return out.index_add(0, sel[keep], flat[keep])


def target_histogram(bins: np.ndarray, value: np.ndarray, count: np.ndarray) -> np.ndarray:
    """A plain-method output (value, count) on the bin grid."""
    where = {float(b): i for i, b in enumerate(bins)}
    out = np.zeros(len(bins), dtype=np.float64)
    for v, c in zip(value, count):
        out[where[float(v)]] += float(c)
    return out


def normalized_l2(h_bins: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Squared L2 distance between the two histograms, each divided by its sum."""
    return ((h_bins / h_bins.sum() - target / target.sum()) ** 2).sum()


def assemble_theta(log_mean, log_sd, logits, quiet):
    """theta [3, n_types] from unconstrained parameters: mean = exp(log_mean),
    sd = exp(log_sd) for proliferating types (-1 for quiescent ones, whose
    entries the library ignores), proportions = softmax(logits)."""
    neg = torch.full_like(log_mean, -1.0)
    mean = torch.where(quiet, neg, torch.exp(log_mean))
    sd = torch.where(quiet, neg, torch.exp(log_sd))
    return torch.stack([mean, sd, torch.softmax(logits, 0)])


def fit(plan: Relaxed, prob: Problem, target: np.ndarray, theta0: np.ndarray, steps: int = 200,
        lr: float = 0.02):
    """Adam on the squared L2 distance between the normalised relaxed
    histogram and `target` (counts on the output bin grid of `prob`), over
    log mean, log sd and proportion logits, starting from theta0.  Returns
    (theta [3, n_types] numpy, per-step losses)."""
    dev = torch.device("cuda")
    bm = BinMap(prob, plan.shape[1], dev)
    tgt = torch.tensor(target, dtype=torch.float64, device=dev)
    th0 = torch.tensor(theta0, dtype=torch.float64, device=dev)
    quiet = th0[0] < 0
    one = torch.ones_like(th0[0])
    log_mean = torch.log(torch.where(quiet, one, th0[0])).clone().requires_grad_(True)
    log_sd = torch.log(torch.where(quiet, one, th0[1])).clone().requires_grad_(True)
    logits = torch.log(th0[2]).clone().requires_grad_(True)
    opt = torch.optim.Adam([log_mean, log_sd, logits], lr=lr)
    losses = []
    for _ in range(steps):
        theta = assemble_theta(log_mean, log_sd, logits, quiet)
        loss = normalized_l2(bm(relaxed_tally(plan, theta)), tgt)
        opt.zero_grad()
        loss.backward()
        opt.step()
        losses.append(loss.detach())
    with torch.no_grad():
        theta = assemble_theta(log_mean, log_sd, logits, quiet)
    return theta.cpu().numpy(), [float(x) for x in torch.stack(losses).cpu()] if losses else []


# ---------------------------------------------------------------------------
# Derivative-free baseline: Nelder-Mead over the plain method.
# ---------------------------------------------------------------------------
def pack(mean, sd, proportion):
    """Unconstrained vector of one set: log mean and log sd of each
    proliferating type, then the log proportions of all types."""
    pr = np.asarray(mean) >= 0
    return np.concatenate([np.log(np.asarray(mean)[pr]), np.log(np.asarray(sd)[pr]),
                           np.log(np.maximum(proportion, 1e-300))])


def unpack(x, prolif):
    """Inverse of pack: (mean, sd, proportion) with -1 for quiescent types."""
    prolif = np.asarray(prolif, dtype=bool)
    n = int(prolif.sum())
    mean = np.full(len(prolif), -1.0)
    sd = np.full(len(prolif), -1.0)
    mean[prolif] = np.exp(x[:n])
    sd[prolif] = np.exp(x[n:2 * n])
    z = x[2 * n:]
    p = np.exp(z - z.max())
    return mean, sd, p / p.sum()


def nelder_mead(x0, step=0.1):
    """Nelder & Mead (1965) simplex search as a coroutine: yields the next
    point to evaluate and receives its objective value through send().
    Standard coefficients: reflection 1, expansion 2, contraction 1/2,
    shrink 1/2; initial simplex x0 + step e_i."""
    d = len(x0)
    pts = [np.array(x0, dtype=np.float64)] + [np.array(x0, dtype=np.float64) + step * np.eye(d)[i]
                                               for i in range(d)]
    f = []
    for p in pts:
        f.append((yield p))
    while True:
        order = np.argsort(f)
        pts = [pts[i] for i in order]
        f = [f[i] for i in order]
        xo = np.mean(pts[:-1], axis=0)
        xr = xo + (xo - pts[-1])
        fr = yield xr
        if f[0] <= fr < f[-2]:
            pts[-1], f[-1] = xr, fr
        elif fr < f[0]:
            xe = xo + 2.0 * (xr - xo)
            fe = yield xe
            pts[-1], f[-1] = (xe, fe) if fe < fr else (xr, fr)
        else:
            xc = xo + 0.5 * ((xr if fr < f[-1] else pts[-1]) - xo)
            fc = yield xc
            if fc < min(fr, f[-1]):
                pts[-1], f[-1] = xc, fc
            else:
                for i in range(1, d + 1):
                    pts[i] = pts[0] + 0.5 * (pts[i] - pts[0])
                    f[i] = yield pts[i]


def plain_loss(bins, value, count, target_q):
    """Normalised-L2 distance of a plain-method output to a normalised target
    on the bin grid."""
    h = target_histogram(bins, value, count)
    s = h.sum()
    r = (h / s if s > 0 else h) - target_q
    return float((r * r).sum())


# ---------------------------------------------------------------------------
# Second order: Levenberg-Marquardt per set over a batched relaxed plan.
# ---------------------------------------------------------------------------
LOG_MAX_SD_RATIO = float(np.log(0.8))  # sd <= 0.8 mean keeps the relaxed trees small
MAX_LOG_STEP = 0.5                     # largest change of one parameter per iteration


def spd_solve(A: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """x with A x = b for a batch of symmetric positive definite A [n, P, P]
    and b [n, P]: Gauss-Jordan elimination without pivoting (the pivots of
    an SPD matrix are positive), P batched torch steps, no host round trip."""
    M = torch.cat([A, b.unsqueeze(-1)], dim=-1).clone()
    P = A.shape[-1]
    for k in range(P):
        row = M[:, k:k + 1, :] / M[:, k:k + 1, k:k + 1]
        M = M - M[:, :, k:k + 1] * row
        M[:, k:k + 1, :] = row
    return M[:, :, P]


def normal_equations(Hj, MJ, SJ, mean, sd, p, q, active):
    """Gauss-Newton terms of the normalised-L2 loss of every set.

    Hj, MJ, SJ [n, B, T]: per bin and type, the binned h, dh/dmean and dh/dsd
    (T padded with zeros); mean, sd, p [n, T]; q [n, B] the normalised target;
    active [n, T, 3] the free parameters (log mean, log sd, proportion logit
    of each type).  With H = sum over types, S = sum H, r = H / S - q and
    J = dr/dphi over phi = (log mean, log sd, logit) per type (proportions =
    softmax of the logits):
        loss = |r|^2,  A = J^T J,  g = J^T r (half the loss gradient).
    Returns (loss [n], A [n, 3T, 3T], g [n, 3T]) with inactive parameters'
    rows and columns of A zero and their g zero."""
    n, B, T = Hj.shape
    H = Hj.sum(-1)
    S = H.sum(-1)
    Sd = torch.where(S > 0, S, torch.ones_like(S))
    r = torch.where((S > 0)[:, None], H / Sd[:, None], torch.zeros_like(H)) - q
    d_logit = Hj - p[:, None, :] * H[:, :, None]  # dH_b / dlogit_l = Hj[b, l] - p_l H_b
    JH = torch.stack([MJ * mean[:, None, :], SJ * sd[:, None, :], d_logit], dim=-1)  # [n, B, T, 3]
    JH = JH * active[:, None, :, :].to(JH.dtype)
    JH = JH.reshape(n, B, 3 * T)
    J = (JH - (H / Sd[:, None])[:, :, None] * JH.sum(1, keepdim=True)) / Sd[:, None, None]
    loss = (r * r).sum(-1)
    Jt = J.transpose(1, 2).contiguous()
    A = torch.bmm(Jt, J)
    g = torch.bmm(Jt, r.unsqueeze(-1)).squeeze(-1)
    return loss, A, g


def lm_step(A, g, lam, active):
    """Levenberg-Marquardt step -(A + lam (diag A + tiny))^-1 g, inactive
    parameters and those the loss does not depend on (a zero Jacobian
    column) held -- their system rows become the identity -- and each
    component limited to MAX_LOG_STEP."""
    n, P, _ = A.shape
    D = torch.diagonal(A, dim1=1, dim2=2)
    act = (active.reshape(n, P) & (D > 0)).to(A.dtype)
    scale = D.amax(-1, keepdim=True).clamp_min(1e-300)
    damp = lam[:, None] * (D + 1e-9 * scale)
    M = A * act[:, :, None] * act[:, None, :] + torch.diag_embed(damp * act + (1.0 - act))
    x = -spd_solve(M, g * act)
    return x.clamp(-MAX_LOG_STEP, MAX_LOG_STEP) * act


class BatchLM:
    """Levenberg-Marquardt calibration of many parameter sets at once.

    Each iteration is one relaxed Jacobian pass over all sets
    (procell_relaxed_jacobian on a batch plan) at every set's trial point,
    then batched torch ops on the device: binning, the Gauss-Newton system
    of the normalised-L2 loss in (log mean, log sd, proportion logits), the
    accept / reject of each set's trial (damping / 5 on success, x 4
    otherwise) and the next trial.  Nothing returns to the host between
    iterations.  Keeps sd <= 0.8 mean (the start point is projected onto
    it too).  A trial whose pass cut a walk of its set short (over the
    relaxed walks' expansion budget: h and the Jacobian are then incomplete)
    is rejected on the device; thetas() raises PROCELL_E_DEPTH when a set's
    current point itself came from a cut pass (only the start can)."""

    def __init__(self, plan, problems, targets, lam0=1e-3, device="cuda", replicas=1):
        """plan: a RelaxedBatch over `problems` (their parameters are the
        start), or over `replicas` copies of them in order (copy r of set i
        is the plan's set r * len(problems) + i, same histogram and types,
        another relaxed-stream seed): the copies' tallies are summed, which
        for the normalised loss averages the estimator over seeds.
        targets: per set a plain-method output, (value, count)."""
        self.plan = plan
        self.dev = torch.device(device)
        n = len(problems)
        if len(plan.shapes) != n * replicas:
            raise ValueError(f"the plan has {len(plan.shapes)} sets, expected {n} x {replicas}")
        for r in range(1, replicas):
            for i in range(n):
                if plan.shapes[r * n + i] != plan.shapes[i]:
                    raise ValueError(f"copy {r} of set {i} has another tally shape")
        T = max(s[2] for s in plan.shapes)
        binsets = []
        for prob, s in zip(problems, plan.shapes):
            bins, idx = output_bins(prob, s[1])
            binsets.append((bins, idx))
        B = max(len(b) for b, _ in binsets)
        self.n, self.T, self.B = n, T, B
        ent, dst = [], []
        q = np.zeros((n, B))
        prolif = np.zeros((n, T), dtype=bool)
        real = np.zeros((n, T), dtype=bool)
        th_src = []
        phi0 = np.zeros((n, T, 3))
        for i, (prob, (bins, idx), s, tg) in enumerate(zip(problems, binsets, plan.shapes, targets)):
            tv, tc = (tg.value, tg.count) if hasattr(tg, "value") else tg
            R, K, Ti = s
            flat = idx.reshape(-1)
            rk = np.nonzero(flat >= 0)[0]
            for r in range(replicas):
                base = plan.h_base[r * n + i]
                for j in range(Ti):
                    ent.append(base + rk * Ti + j)
                    dst.append((i * B + flat[rk]) * T + j)
            t = target_histogram(bins, np.asarray(tv), np.asarray(tc))
            if t.sum() > 0:
                q[i, :len(bins)] = t / t.sum()
            m, sd, pr = prob.mean, prob.sd, prob.proportion
            pf = np.asarray(m) >= 0
            prolif[i, :Ti] = pf
            real[i, :Ti] = True
            phi0[i, :Ti, 0] = np.log(np.where(pf, m, 1.0))
            phi0[i, :Ti, 1] = np.log(np.where(pf, sd, 1.0))
            phi0[i, :Ti, 2] = np.log(np.maximum(pr, 1e-12))
        for r in range(replicas):  # theta of the plan's sets, in plan order
            for i, s in enumerate(plan.shapes[:n]):
                for c in range(3):
                    th_src.extend(((i * 3 + c) * T + j) for j in range(s[2]))
        d = self.dev
        self.ent = torch.tensor(np.concatenate(ent), dtype=torch.int64, device=d)
        self.dst = torch.tensor(np.concatenate(dst), dtype=torch.int64, device=d)
        self.q = torch.tensor(q, dtype=torch.float64, device=d)
        self.prolif = torch.tensor(prolif, device=d)
        self.real = torch.tensor(real, device=d)
        self.th_src = torch.tensor(th_src, dtype=torch.int64, device=d)
        # logits: all but each set's last type (softmax is unchanged by a
        # common shift, so one logit is fixed; a one-type set has none free)
        last = real.sum(1, keepdims=True) - 1
        free_logit = real & (np.arange(T)[None, :] < last)
        act = np.stack([prolif, prolif, free_logit], axis=-1)
        self.active = torch.tensor(act, device=d)
        self.phi = self._project(torch.tensor(phi0, dtype=torch.float64, device=d))
        self.lam = torch.full((n,), float(lam0), dtype=torch.float64, device=d)
        self.h = torch.empty(plan.h_total, dtype=torch.float64, device=d)
        self.jac = torch.empty((2, plan.h_total), dtype=torch.float64, device=d)
        self.replicas = replicas
        self.set_cut = torch.zeros(len(plan.shapes), dtype=torch.int64, device=d)  # last pass, per plan set
        self.cut = torch.zeros(n, dtype=torch.bool, device=d)       # current point from a cut pass
        self.cut_trials = torch.zeros(n, dtype=torch.int64, device=d)  # trials rejected for cuts
        self.iterations = 0
        self.loss = self.A = self.g = None

    def start(self):
        """Evaluates the starting point (one Jacobian pass); step() does it
        first when it has not been done."""
        if self.loss is None:
            loss, self.A, self.g, self.cut = self._evaluate(self.phi)
            self.loss = torch.where(self.cut, torch.full_like(loss, torch.inf), loss)

    def _params(self, phi):
        neg = torch.full_like(phi[..., 0], -1.0)
        mean = torch.where(self.prolif, torch.exp(phi[..., 0]), neg)
        sd = torch.where(self.prolif, torch.exp(phi[..., 1]), neg)
        logits = torch.where(self.real, phi[..., 2], torch.full_like(phi[..., 2], -torch.inf))
        p = torch.softmax(logits, -1)
        return mean, sd, p

    def theta_flat(self, phi):
        mean, sd, p = self._params(phi)
        return torch.stack([mean, sd, p], 1).reshape(-1).index_select(0, self.th_src).contiguous()

    def _evaluate(self, phi):
        mean, sd, p = self._params(phi)
        self.plan.jacobian(self.theta_flat(phi), self.h, self.jac)
        self.plan.set_cuts(self.set_cut)
        n, B, T = self.n, self.B, self.T
        cut = self.set_cut.reshape(self.replicas, n).sum(0) > 0
        vals = torch.stack([self.h.index_select(0, self.ent), self.jac[0].index_select(0, self.ent),
                            self.jac[1].index_select(0, self.ent)], -1)
        tally = torch.zeros((n * B * T, 3), dtype=torch.float64, device=self.dev).index_add_(0, self.dst, vals)
        tally = tally.reshape(n, B, T, 3)
        loss, A, g = normal_equations(tally[..., 0], tally[..., 1], tally[..., 2], mean.clamp_min(0),
                                      sd.clamp_min(0), p, self.q, self.active)
        return loss, A, g, cut

    def _project(self, phi):
        ls = torch.minimum(phi[..., 1], phi[..., 0] + LOG_MAX_SD_RATIO)
        return torch.stack([phi[..., 0], ls, phi[..., 2]], -1)

    def step(self, n_iter=1):
        """n_iter Levenberg-Marquardt iterations of every set."""
        self.start()
        for _ in range(n_iter):
            dx = lm_step(self.A, self.g, self.lam, self.active).reshape(self.n, self.T, 3)
            trial = self._project(self.phi + dx)
            loss, A, g, cut = self._evaluate(trial)
            ok = (loss < self.loss) & ~cut
            self.cut_trials += cut.to(torch.int64)
            self.cut = self.cut & ~ok
            self.phi = torch.where(ok[:, None, None], trial, self.phi)
            self.loss = torch.where(ok, loss, self.loss)
            self.A = torch.where(ok[:, None, None], A, self.A)
            self.g = torch.where(ok[:, None], g, self.g)
            self.lam = torch.where(ok, self.lam * 0.2, self.lam * 4.0).clamp(1e-9, 1e9)
            self.iterations += 1

    def thetas(self):
        """Per set [3, n_types] numpy (mean, sd, proportion; -1 for quiescent).
        Raises PROCELL_E_DEPTH when a set's current point came from a pass
        that cut a walk short."""
        if self.cut.any():
            from procell import E_DEPTH, ProcellError

            bad = torch.nonzero(self.cut).flatten().tolist()
            raise ProcellError(E_DEPTH, f"BatchLM: walks cut short at the current point of sets {bad[:8]}")
        mean, sd, p = (x.cpu().numpy() for x in self._params(self.phi))
        out = []
        for i, s in enumerate(self.plan.shapes[:self.n]):
            Ti = s[2]
            out.append(np.stack([mean[i, :Ti], sd[i, :Ti], p[i, :Ti]]))
        return out
