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


def binned(h: torch.Tensor, idx: torch.Tensor, n_bins: int) -> torch.Tensor:
    """Sum of h over types and over the (row, generation) entries of each bin."""
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
