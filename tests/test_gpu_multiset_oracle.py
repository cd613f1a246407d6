"""GPU: plans over several parameter sets against the oracle, set by set.

A relaxed plan over heterogeneous sets (different histograms, t_max, phi,
seeds and type counts, one all-quiescent set) walks them all in one launch;
each set's h, adjoint gradient and Jacobian slice must equal oracle_relaxed
of that set alone, within the relaxed parity tolerances (FP64 atomics in
another order).  One step of a batched calibration must equal the oracle
composed in numpy: the oracle's h, the normalised-L2 loss and its cotangent
on the output bins, the oracle's adjoint with that cotangent, the chain rule
to (log mean, log sd, proportion logits) and the first Adam update.  The
walk-cut counters of a batch plan are reported per set.
"""
import numpy as np
import pytest

from test_gpu_relaxed import G_ATOL, G_RTOL, H_ATOL, H_RTOL

pytestmark = pytest.mark.gpu


def _close(x, ref, rtol, atol):
    scale = max(1.0, float(np.abs(ref).max())) if ref.size else 1.0
    return bool(np.all(np.abs(x - ref) <= rtol * np.abs(ref) + atol * scale))


def _sets():
    """Five sets: the small standard workload, a phi cut, an all-quiescent
    set, deep rows with one type, and five types (four proliferating)."""
    from procell import workloads

    w = workloads.small_histogram(4)
    rng = np.random.default_rng(21)
    vals5 = 10.0 ** np.linspace(0.0, 4.0, 40)
    return [
        dict(values=w.values, counts=w.counts, proportion=w.proportion, mean=w.mean, sd=w.sd,
             t_max=w.t_max, phi=0.0, seed=7),
        dict(values=[64.0, 20.0, 5.0, 2.5], counts=[600, 400, 300, 200], proportion=[0.45, 0.35, 0.2],
             mean=[14.0, 22.0, -1.0], sd=[4.0, 7.0, -1.0], t_max=70.0, phi=3.0, seed=11),
        dict(values=[8.0, 3.0, 0.5], counts=[30, 20, 10], proportion=[0.6, 0.4], mean=[-1.0, -1.0],
             sd=[-1.0, -1.0], t_max=40.0, phi=1.0, seed=2),
        dict(values=[2.0**20, 1000.0], counts=[50, 70], proportion=[1.0], mean=[20.0], sd=[3.0],
             t_max=75.0, phi=1.0, seed=3),
        dict(values=vals5, counts=rng.poisson(60.0, len(vals5)), proportion=[0.3, 0.25, 0.2, 0.15, 0.1],
             mean=[30.0, 45.0, 60.0, 80.0, -1.0], sd=[8.0, 12.0, 15.0, 20.0, -1.0], t_max=200.0, phi=20.0,
             seed=13),
    ]


def _args(d):
    return (d["values"], d["counts"], d["proportion"], d["mean"], d["sd"], d["t_max"], d["phi"], d["seed"])


@pytest.mark.parametrize("beta", [0.5, 2.0])
def test_batch_plan_matches_oracle_set_by_set(procell_lib, oracle, beta):
    import torch

    cases = _sets()
    probs = [procell_lib.Problem(*_args(d)) for d in cases]
    rng = np.random.default_rng(int(beta * 10))
    with procell_lib.RelaxedBatch(probs, beta) as bp:
        th = torch.tensor(np.concatenate([np.stack([d["mean"], d["sd"], d["proportion"]]).reshape(-1)
                                          for d in cases]), dtype=torch.float64, device="cuda")
        g_np = rng.normal(size=bp.h_total)
        g = torch.tensor(g_np, dtype=torch.float64, device="cuda")
        h = torch.empty(bp.h_total, dtype=torch.float64, device="cuda")
        hj = torch.empty_like(h)
        jac = torch.empty((2, bp.h_total), dtype=torch.float64, device="cuda")
        grad = torch.empty(3 * bp.types_total, dtype=torch.float64, device="cuda")
        cuts = torch.full((len(cases),), -1, dtype=torch.int64, device="cuda")
        bp.forward(th, h)
        bp.set_cuts(cuts)
        bp.jacobian(th, hj, jac)
        bp.backward(th, g, grad)
        bp.expansions()  # no walk was cut
        h, hj, jac, grad, cuts = (x.cpu().numpy() for x in (h, hj, jac, grad, cuts))
        shapes, h_base, t_base = bp.shapes, bp.h_base, bp.theta_base
    assert cuts.tolist() == [0] * len(cases)
    for i, d in enumerate(cases):
        print("set", i, "beta", beta, "shape", shapes[i], flush=True)
        op = oracle.Problem(*_args(d))
        h_ref, _ = oracle.relaxed(op, beta)
        assert tuple(h_ref.shape) == tuple(shapes[i])
        n = h_ref.size
        sl = slice(h_base[i], h_base[i] + n)
        gi = g_np[sl].reshape(h_ref.shape)
        _, grad_ref = oracle.relaxed(op, beta, gi)
        assert _close(h[sl].reshape(h_ref.shape), h_ref, H_RTOL, H_ATOL), (i, np.abs(h[sl] - h_ref.reshape(-1)).max())
        assert _close(hj[sl].reshape(h_ref.shape), h_ref, H_RTOL, H_ATOL), i
        T = shapes[i][2]
        gr = grad[t_base[i]:t_base[i] + 3 * T].reshape(3, T)
        assert _close(gr, grad_ref, G_RTOL, G_ATOL), (i, gr, grad_ref)
        # the Jacobian slice contracted with this set's cotangent
        p = np.asarray(d["proportion"], dtype=np.float64)
        jm = jac[0, sl].reshape(h_ref.shape)
        js = jac[1, sl].reshape(h_ref.shape)
        hh = hj[sl].reshape(h_ref.shape)
        gj = np.stack([(gi * jm).sum((0, 1)), (gi * js).sum((0, 1)), (gi * hh).sum((0, 1)) / p])
        assert _close(gj, grad_ref, G_RTOL, G_ATOL), (i, gj, grad_ref)
        if np.all(np.asarray(d["mean"]) < 0):
            assert np.all(jm == 0) and np.all(js == 0) and np.all(gr[:2] == 0)


def _oracle_step(oracle, w, seed, target, beta, lr, b1, b2, eps):
    """Loss at the start and theta after one Adam step, from the oracle."""
    import procell
    from procell import calibrate

    op = oracle.Problem(w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max, w.phi, seed)
    h, _ = oracle.relaxed(op, beta)
    prob = procell.Problem(w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max, w.phi, seed)
    bins, idx = calibrate.output_bins(prob, h.shape[1])
    q = calibrate.target_histogram(bins, target.value, target.count)
    q = q / q.sum()
    Hb = np.zeros(len(bins))
    rows, ks = np.nonzero(idx >= 0)
    np.add.at(Hb, idx[rows, ks], h[rows, ks].sum(-1))
    S = Hb.sum()
    r = Hb / S - q
    loss = float((r * r).sum())
    dHb = 2.0 / S * (r - (r * Hb).sum() / S)
    g = np.zeros_like(h)
    g[rows, ks, :] = dHb[idx[rows, ks]][:, None]
    _, grad = oracle.relaxed(op, beta, g)
    quiet = w.mean < 0
    one = np.ones_like(w.mean)
    x = np.stack([np.log(np.where(quiet, one, w.mean)), np.log(np.where(quiet, one, w.sd)), np.log(w.proportion)])
    p = w.proportion
    gx = np.stack([np.where(quiet, 0.0, grad[0] * w.mean), np.where(quiet, 0.0, grad[1] * w.sd),
                   p * (grad[2] - (p * grad[2]).sum())])
    m = (1.0 - b1) * gx
    v = (1.0 - b2) * gx * gx
    x1 = x - lr * (m / (1.0 - b1)) / (np.sqrt(v / (1.0 - b2)) + eps)
    z = np.exp(x1[2] - x1[2].max())
    th = np.stack([np.where(quiet, -1.0, np.exp(x1[0])), np.where(quiet, -1.0, np.exp(x1[1])), z / z.sum()])
    return loss, th


def test_calibration_step_matches_oracle(procell_lib, oracle):
    from procell import workloads

    lr, b1, b2, eps, beta = 0.05, 0.9, 0.999, 1e-8, 0.5
    gen, start = workloads.calibration_batch(23, 4, 4000.0)
    targets = [procell_lib.simulate(procell_lib.Problem.from_workload(w, 300 + i), want_raw=False)
               for i, w in enumerate(gen)]
    seeds = [900 + i for i in range(len(start))]
    probs = [procell_lib.Problem.from_workload(w, s) for w, s in zip(start, seeds)]
    with procell_lib.Calibration(probs, beta, targets, lr=lr, beta1=b1, beta2=b2, eps=eps) as cal:
        cal.step(1)
        loss, n = cal.loss()
        theta = cal.theta()
    assert n == 1
    for i, (w, s, t) in enumerate(zip(start, seeds, targets)):
        l_ref, th_ref = _oracle_step(oracle, w, s, t, beta, lr, b1, b2, eps)
        print("set", i, "loss", loss[i], l_ref, flush=True)
        assert abs(loss[i] - l_ref) <= 1e-9 * abs(l_ref), (i, loss[i], l_ref)
        assert np.allclose(theta[i], th_ref, rtol=1e-9, atol=1e-12), (i, theta[i], th_ref)


def test_walk_cuts_are_reported_per_set(procell_lib):
    """A set whose walks outgrow the expansion budget (a 2^30-fluorescence
    row, mean 10, t_max 1000: every node divides) is cut and reported in
    its own counter; the other sets' counters stay zero and their h equals
    their single-set plans'."""
    import torch

    ok = _sets()[1]
    runaway = dict(values=[2.0**30], counts=[3], proportion=[1.0], mean=[10.0], sd=[2.0], t_max=1000.0,
                   phi=1.0, seed=1)
    cases = [ok, runaway, ok]
    probs = [procell_lib.Problem(*_args(d)) for d in cases]
    with procell_lib.RelaxedBatch(probs, 0.5) as bp:
        th = torch.tensor(np.concatenate([np.stack([d["mean"], d["sd"], d["proportion"]]).reshape(-1)
                                          for d in cases]), dtype=torch.float64, device="cuda")
        h = torch.empty(bp.h_total, dtype=torch.float64, device="cuda")
        cuts = torch.full((3,), -1, dtype=torch.int64, device="cuda")
        bp.forward(th, h)
        bp.set_cuts(cuts)
        with pytest.raises(procell_lib.ProcellError) as e:
            bp.expansions()
        assert e.value.code == procell_lib.E_DEPTH
        c = cuts.cpu().numpy()
        h = h.cpu().numpy()
        base, shp = bp.h_base, bp.shapes
    print("cuts", c, flush=True)
    assert c[0] == 0 and c[2] == 0 and c[1] > 0
    with procell_lib.Relaxed(procell_lib.Problem(*_args(ok)), 0.5) as rp:
        th1 = torch.tensor(np.stack([ok["mean"], ok["sd"], ok["proportion"]]), dtype=torch.float64, device="cuda")
        h1 = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        rp.forward(th1, h1)
        h1 = h1.cpu().numpy().reshape(-1)
    for i in (0, 2):
        n = int(np.prod(shp[i]))
        assert _close(h[base[i]:base[i] + n], h1, H_RTOL, H_ATOL)
