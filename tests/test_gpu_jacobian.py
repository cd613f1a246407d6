"""GPU: the relaxed Jacobian pass, the batch plan and Levenberg-Marquardt.

procell_relaxed_jacobian tallies h and, per entry, dh/dmean_j and dh/dsd_j
of the entry's own type; contracted with a cotangent g it must give the
oracle's relaxed gradient (oracle_relaxed), with dh/dproportion_j = h /
proportion_j, within the tolerances of the relaxed parity tests (FP64
atomics in another order).  A batch plan must give every set's single-plan
forward, Jacobian and adjoint.  Levenberg-Marquardt over a batch plan must
never increase a set's loss and must reduce the parameter error of a
calibration batch.
"""
import numpy as np
import pytest

from test_gpu_relaxed import G_ATOL, G_RTOL, H_ATOL, H_RTOL, _args, _problems

pytestmark = pytest.mark.gpu


def _close(x, ref, rtol, atol):
    scale = max(1.0, float(np.abs(ref).max())) if ref.size else 1.0
    return bool(np.all(np.abs(x - ref) <= rtol * np.abs(ref) + atol * scale))


@pytest.mark.parametrize("case", range(4))
def test_jacobian_contracts_to_oracle_gradient(procell_lib, oracle, case):
    import torch

    d, beta = _problems()[case]
    print("jacobian case", case, "beta", beta, flush=True)
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), beta)
    g = np.random.default_rng(100 + case).normal(size=h_ref.shape)
    _, grad_ref = oracle.relaxed(oracle.Problem(*_args(d)), beta, g)
    prob = procell_lib.Problem(*_args(d))
    th = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, beta) as rp:
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        jac = torch.empty((2,) + tuple(rp.shape), dtype=torch.float64, device="cuda")
        rp.jacobian(th, h, jac)
        assert rp.expansions() > 0
        h, jac = h.cpu().numpy(), jac.cpu().numpy()
    assert _close(h, h_ref, H_RTOL, H_ATOL), np.abs(h - h_ref).max()
    p = np.asarray(d["proportion"], dtype=np.float64)
    grad = np.stack([(g * jac[0]).sum((0, 1)), (g * jac[1]).sum((0, 1)), (g * h).sum((0, 1)) / p])
    assert _close(grad, grad_ref, G_RTOL, G_ATOL), (grad, grad_ref)
    quiet = np.asarray(d["mean"]) < 0
    assert np.all(jac[:, :, :, quiet] == 0.0)


def test_strided_jacobian_passes_sum_to_full(procell_lib):
    """With cell stride 4 the Jacobian passes of phases 0..3 (the strided
    instantiation) sum to the full pass, h and both derivative tallies."""
    import torch

    d, beta = _problems()[0]
    prob = procell_lib.Problem(*_args(d))
    th = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, beta) as rp:
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        jac = torch.empty((2,) + tuple(rp.shape), dtype=torch.float64, device="cuda")
        rp.jacobian(th, h, jac)
        full = np.concatenate([h.cpu().numpy()[None], jac.cpu().numpy()])
        n_full = rp.expansions()
        total, n_sum = np.zeros_like(full), 0
        for phase in range(4):
            rp.set_stride(4, phase)
            rp.jacobian(th, h, jac)
            n_sum += rp.expansions()
            total += np.concatenate([h.cpu().numpy()[None], jac.cpu().numpy()])
        rp.set_stride(1)
    assert n_sum == n_full
    for a, b in zip(total, full):
        assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b) + 1e-12 * np.abs(b).max())


def test_batch_plan_equals_single_plans(procell_lib):
    import torch

    cases = [c for c, _ in _problems()]
    beta = 0.7
    probs = [procell_lib.Problem(*_args(d)) for d in cases]
    rng = np.random.default_rng(5)
    with procell_lib.RelaxedBatch(probs, beta) as bp:
        assert len(bp.shapes) == len(cases)
        th = torch.tensor(np.concatenate([np.stack([d["mean"], d["sd"], d["proportion"]]).reshape(-1)
                                          for d in cases]), dtype=torch.float64, device="cuda")
        g = torch.tensor(rng.normal(size=bp.h_total), dtype=torch.float64, device="cuda")
        h = torch.empty(bp.h_total, dtype=torch.float64, device="cuda")
        hj = torch.empty_like(h)
        jac = torch.empty((2, bp.h_total), dtype=torch.float64, device="cuda")
        grad = torch.empty(3 * bp.types_total, dtype=torch.float64, device="cuda")
        bp.forward(th, h)
        bp.jacobian(th, hj, jac)
        bp.backward(th, g, grad)
        h, hj, jac, grad, g = (x.cpu().numpy() for x in (h, hj, jac, grad, g))
        shapes, h_base, t_base = bp.shapes, bp.h_base, bp.theta_base
    for i, (d, prob) in enumerate(zip(cases, probs)):
        print("batch set", i, flush=True)
        shp = shapes[i]
        n = int(np.prod(shp))
        sl = slice(h_base[i], h_base[i] + n)
        thi = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
        with procell_lib.Relaxed(prob, beta) as rp:
            assert tuple(rp.shape) == tuple(shp)
            h1 = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
            j1 = torch.empty((2,) + tuple(rp.shape), dtype=torch.float64, device="cuda")
            gr1 = torch.empty_like(thi)
            rp.forward(thi, h1)
            h1 = h1.cpu().numpy().reshape(-1)
            rp.jacobian(thi, torch.empty(rp.shape, dtype=torch.float64, device="cuda"), j1)
            rp.backward(thi, torch.tensor(g[sl].reshape(shp), device="cuda"), gr1)
            j1, gr1 = j1.cpu().numpy().reshape(2, -1), gr1.cpu().numpy().reshape(-1)
        assert _close(h[sl], h1, H_RTOL, H_ATOL)
        assert _close(hj[sl], h1, H_RTOL, H_ATOL)
        assert _close(jac[:, sl], j1, H_RTOL, H_ATOL)
        T = shp[2]
        assert _close(grad[t_base[i]:t_base[i] + 3 * T], gr1, G_RTOL, G_ATOL)


def test_levenberg_marquardt_reduces_loss_and_error(procell_lib):
    import torch
    from procell import calibrate, workloads

    gen, start = workloads.calibration_batch(3, 12, 2e4)
    with procell_lib.Batch([procell_lib.Problem.from_workload(w, 0) for w in gen]) as b:
        b.run([1000 + i for i in range(len(gen))])
        targets = [b.result(i, want_raw=False) for i in range(len(gen))]
    probs = [procell_lib.Problem.from_workload(w, 50000 + i) for i, w in enumerate(start)]

    def err(thetas):
        e = [np.abs(th[0][w.mean >= 0] - w.mean[w.mean >= 0]) / w.mean[w.mean >= 0] for w, th in zip(gen, thetas)]
        return float(np.median(np.concatenate(e)))

    with procell_lib.RelaxedBatch(probs, 0.5) as plan:
        lm = calibrate.BatchLM(plan, probs, targets)
        lm.start()
        e0 = err(lm.thetas())
        losses = [lm.loss.clone()]
        for it in range(8):
            lm.step(1)
            losses.append(lm.loss.clone())
        torch.cuda.synchronize()
        plan.expansions()  # no walk was cut
        th = lm.thetas()
    L = torch.stack(losses).cpu().numpy()
    print("median loss", np.median(L, axis=1), "error", e0, "->", err(th), flush=True)
    assert np.all(np.isfinite(L))
    assert np.all(L[1:] <= L[:-1])
    assert np.median(L[-1]) < 0.5 * np.median(L[0])
    assert err(th) < 0.8 * e0
    for w, t in zip(start, th):
        pr = w.mean >= 0
        assert np.all(t[1][pr] <= 0.8 * t[0][pr] * (1 + 1e-12))
        assert abs(t[2].sum() - 1.0) < 1e-12


def test_levenberg_marquardt_edge_sets(procell_lib):
    """An all-quiescent set (no walks; its normalised histogram does not
    depend on its proportions, so nothing moves), a one-type set (no free
    logit) and a set with an empty target step with the others: every loss
    stays finite and never increases, the all-quiescent set keeps its
    proportions and the one-type set's proportion stays 1."""
    import torch
    from procell import calibrate, workloads

    gen, start = workloads.calibration_batch(4, 2, 5000)
    with procell_lib.Batch([procell_lib.Problem.from_workload(w, 0) for w in gen]) as b:
        b.run([1000, 1001])
        targets = [b.result(i, want_raw=False) for i in range(2)]
    w = start[0]
    quiet = procell_lib.Problem(w.values, w.counts, [0.3, 0.7], [-1.0, -1.0], [-1.0, -1.0], w.t_max, w.phi, 5)
    one = procell_lib.Problem(w.values, w.counts, [1.0], [w.mean[0]], [w.sd[0]], w.t_max, w.phi, 6)
    probs = [procell_lib.Problem.from_workload(s, 50 + i) for i, s in enumerate(start)] + [quiet, one]
    empty = (np.zeros(0), np.zeros(0, dtype=np.int64))
    tg = [targets[0], empty, targets[0], targets[0]]
    with procell_lib.RelaxedBatch(probs, 0.5) as plan:
        lm = calibrate.BatchLM(plan, probs, tg)
        lm.start()
        losses = [lm.loss.clone()]
        for _ in range(4):
            lm.step(1)
            losses.append(lm.loss.clone())
        th = lm.thetas()
    L = torch.stack(losses).cpu().numpy()
    print("losses", L, flush=True)
    assert np.all(np.isfinite(L)) and np.all(L[1:] <= L[:-1])
    assert L[-1][2] == L[0][2]
    assert np.allclose(th[2][2], [0.3, 0.7], rtol=1e-12)
    assert L[-1][3] < L[0][3] and L[-1][0] < L[0][0]
    assert th[3][2].tolist() == [1.0]
    assert np.all(th[2][0] == -1.0) and abs(th[2][2].sum() - 1.0) < 1e-12


def test_levenberg_marquardt_replicas_map_onto_their_sets(procell_lib):
    """Two copies of every set with the same seed sum to twice the tally,
    which the normalised loss does not see: the run equals the one-copy run
    (to the rounding of the sums).  A mismatched plan is rejected."""
    import torch
    from procell import calibrate, workloads

    gen, start = workloads.calibration_batch(6, 5, 1e4)
    with procell_lib.Batch([procell_lib.Problem.from_workload(w, 0) for w in gen]) as b:
        b.run([1000 + i for i in range(len(gen))])
        targets = [b.result(i, want_raw=False) for i in range(len(gen))]
    probs = [procell_lib.Problem.from_workload(w, 50 + i) for i, w in enumerate(start)]
    runs = []
    for K in (1, 2):
        with procell_lib.RelaxedBatch(probs * K, 0.5) as plan:
            lm = calibrate.BatchLM(plan, probs, targets, replicas=K)
            lm.start()
            losses = [lm.loss.clone()]
            for _ in range(3):
                lm.step(1)
                losses.append(lm.loss.clone())
            runs.append((torch.stack(losses).cpu().numpy(), lm.thetas()))
        print("replicas", K, "loss", runs[-1][0][-1], flush=True)
    np.testing.assert_allclose(runs[1][0], runs[0][0], rtol=1e-9)
    for a, b in zip(runs[0][1], runs[1][1]):
        np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-10)
    with procell_lib.RelaxedBatch(probs, 0.5) as plan:
        with pytest.raises(ValueError):
            calibrate.BatchLM(plan, probs, targets, replicas=2)
