"""Calibration binding: the relaxed estimator under torch.autograd.

RelaxedTally is a torch.autograd.Function over the C ABI
(procell_relaxed_forward / procell_relaxed_backward): its forward returns the
relaxed expected tally h[row, generation, type] of the plan's histogram for
theta = [mean; sd; proportion] (3 x n_types, float64, on the GPU), and its
backward hands the incoming cotangent dL/dh to the library's adjoint pass,
which recomputes every subtree from the random stream.  The helpers below
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
        h = torch.empty(plan.shape, dtype=torch.float64, device=theta.device)
        plan.forward(theta, h)
        ctx.plan = plan
        ctx.save_for_backward(theta)
        return h

    @staticmethod
    def backward(ctx, g):
        (theta,) = ctx.saved_tensors
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
