"""GPU: batched calibration against an independent composition.

One calibration step (relaxed forward of all sets in one launch, loss and
cotangent kernel, adjoint of all sets, on-device Adam) must equal, set by
set, the single-set relaxed plan composed with torch: BinMap, the
normalised-L2 loss, torch.autograd through RelaxedTally, exp and softmax,
and torch.optim.Adam with the same hyper-parameters.  The device sums with
FP64 atomics in its own order, so agreement is to 1e-9 relative (loss and
parameters), checked after each of three steps.  Graph replays enqueued
together equal steps taken one at a time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LR, B1, B2, EPS, BETA = 0.05, 0.9, 0.999, 1e-8, 0.5


def _sets(procell_lib, n=3, cells=3000.0):
    from procell import workloads

    gen, start = workloads.calibration_batch(11, n, cells)
    targets = [procell_lib.simulate(procell_lib.Problem.from_workload(w, 100 + i), want_raw=False)
               for i, w in enumerate(gen)]
    probs = [procell_lib.Problem.from_workload(w, 7 + i) for i, w in enumerate(start)]
    return gen, start, targets, probs


def _torch_reference(procell_lib, prob, w, target, steps, stride=1):
    import torch
    from procell import calibrate

    plan = procell_lib.Relaxed(prob, BETA)
    bm = calibrate.BinMap(prob, plan.shape[1])
    tgt = torch.tensor(calibrate.target_histogram(bm.bins, target.value, target.count), dtype=torch.float64,
                       device="cuda")
    quiet = torch.tensor(w.mean < 0, device="cuda")
    one = np.ones(len(w.mean))
    lm = torch.tensor(np.log(np.where(w.mean < 0, one, w.mean)), dtype=torch.float64, device="cuda",
                      requires_grad=True)
    ls = torch.tensor(np.log(np.where(w.mean < 0, one, w.sd)), dtype=torch.float64, device="cuda",
                      requires_grad=True)
    z = torch.tensor(np.log(w.proportion), dtype=torch.float64, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([lm, ls, z], lr=LR, betas=(B1, B2), eps=EPS)
    out = []
    for step in range(steps):
        plan.set_stride(stride, step % stride)
        theta = calibrate.assemble_theta(lm, ls, z, quiet)
        loss = calibrate.normalized_l2(bm(calibrate.relaxed_tally(plan, theta)), tgt)
        opt.zero_grad()
        # with a stride the adjoint walks the next class of cells (double sampling)
        plan.set_stride(stride, (step + (1 if stride > 1 else 0)) % stride)
        loss.backward()
        opt.step()
        with torch.no_grad():
            out.append((float(loss), calibrate.assemble_theta(lm, ls, z, quiet).cpu().numpy()))
    plan.close()
    return out


def test_steps_equal_torch_composition(procell_lib):
    gen, start, targets, probs = _sets(procell_lib)
    refs = [_torch_reference(procell_lib, p, w, t, 3) for p, w, t in zip(probs, start, targets)]
    with procell_lib.Calibration(probs, BETA, targets, lr=LR, beta1=B1, beta2=B2, eps=EPS) as cal:
        for step in range(3):
            cal.step(1)
            loss, n = cal.loss()
            theta = cal.theta()
            assert n == step + 1
            for i, ref in enumerate(refs):
                l_ref, th_ref = ref[step]
                print("step", step, "set", i, loss[i], l_ref, flush=True)
                assert abs(loss[i] - l_ref) <= 1e-9 * abs(l_ref), (step, i, loss[i], l_ref)
                assert np.allclose(theta[i], th_ref, rtol=1e-9, atol=1e-12), (step, i, theta[i], th_ref)


def test_graph_replays_equal_single_steps(procell_lib):
    gen, start, targets, probs = _sets(procell_lib, 4, 2000.0)
    with procell_lib.Calibration(probs, BETA, targets, lr=LR) as a, \
            procell_lib.Calibration(probs, BETA, targets, lr=LR) as b:
        for _ in range(5):
            a.step(1)
            a.theta()
        b.step(5)
        la, na = a.loss()
        lb, nb = b.loss()
        assert na == nb == 5
        assert np.allclose(la, lb, rtol=1e-9, atol=0)
        for x, y in zip(a.theta(), b.theta()):
            assert np.allclose(x, y, rtol=1e-8, atol=1e-12)


def test_invalid_targets_and_parameters(procell_lib):
    gen, start, targets, probs = _sets(procell_lib, 2, 1000.0)
    bad = [(np.array([12345.678]), np.array([3]))] + targets[1:]
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.Calibration(probs, BETA, bad)
    assert e.value.code == procell_lib.E_INVALID
    with pytest.raises(procell_lib.ProcellError):
        procell_lib.Calibration(probs, 0.0, targets)


def test_edge_sets_equal_torch_composition(procell_lib):
    """One proliferating type alone, an all-quiescent set (no walks at all;
    its normalised histogram does not depend on the proportions) and four
    types with a quiescent one, in one calibration."""
    from procell import workloads

    vals = [10.0 ** (4.0 * i / 63.0) for i in range(64)]
    rng = np.random.default_rng(2)
    counts = rng.poisson(30.0, 64)
    W = workloads.Workload
    gen = [
        W("one", np.array(vals), counts.astype(np.uint64), np.array([1.0]), np.array([45.0]), np.array([12.0]),
          150.0, 0.0),
        W("quiet", np.array(vals), counts.astype(np.uint64), np.array([0.3, 0.7]), np.array([-1.0, -1.0]),
          np.array([-1.0, -1.0]), 150.0, 0.0),
        W("four", np.array(vals), counts.astype(np.uint64), np.array([0.2, 0.3, 0.4, 0.1]),
          np.array([40.0, 70.0, 55.0, -1.0]), np.array([10.0, 20.0, 15.0, -1.0]), 180.0, 0.0),
    ]
    start = []
    for w in gen:
        m = np.where(w.mean >= 0, w.mean * 1.1, -1.0)
        s = np.where(w.mean >= 0, w.sd * 0.9, -1.0)
        p = 0.5 * w.proportion + 0.5 / len(w.proportion)
        start.append(W(w.name + "0", w.values, w.counts, p / p.sum(), m, s, w.t_max, w.phi))
    targets = [procell_lib.simulate(procell_lib.Problem.from_workload(w, 300 + i), want_raw=False)
               for i, w in enumerate(gen)]
    probs = [procell_lib.Problem.from_workload(w, 9 + i) for i, w in enumerate(start)]
    refs = [_torch_reference(procell_lib, p, w, t, 2) for p, w, t in zip(probs, start, targets)]
    with procell_lib.Calibration(probs, BETA, targets, lr=LR, beta1=B1, beta2=B2, eps=EPS) as cal:
        for step in range(2):
            cal.step(1)
            loss, _ = cal.loss()
            theta = cal.theta()
            for i, ref in enumerate(refs):
                l_ref, th_ref = ref[step]
                assert abs(loss[i] - l_ref) <= 1e-9 * abs(l_ref) + 1e-15, (step, i, loss[i], l_ref)
                assert np.allclose(theta[i], th_ref, rtol=1e-9, atol=1e-12), (step, i, theta[i], th_ref)
        assert np.allclose(cal.theta()[1][2], start[1].proportion, rtol=1e-12)


def test_strided_steps_equal_torch_composition(procell_lib):
    """Cell stride 2: step s's forward pass walks the cells of phase s mod 2
    and its adjoint those of phase s + 1 mod 2, as a single-set plan composed
    with torch does with set_stride before each pass."""
    gen, start, targets, probs = _sets(procell_lib, 3, 4000.0)
    refs = [_torch_reference(procell_lib, p, w, t, 3, stride=2) for p, w, t in zip(probs, start, targets)]
    with procell_lib.Calibration(probs, BETA, targets, lr=LR, beta1=B1, beta2=B2, eps=EPS, cell_stride=2) as cal:
        for step in range(3):
            cal.step(1)
            loss, _ = cal.loss()
            theta = cal.theta()
            for i, ref in enumerate(refs):
                l_ref, th_ref = ref[step]
                assert abs(loss[i] - l_ref) <= 1e-9 * abs(l_ref), (step, i, loss[i], l_ref)
                assert np.allclose(theta[i], th_ref, rtol=1e-9, atol=1e-12), (step, i, theta[i], th_ref)
    with pytest.raises(procell_lib.ProcellError):
        procell_lib.Calibration(probs, BETA, targets, cell_stride=3)


def test_sd_kept_below_bound(procell_lib):
    """A target generated with sd = 0.95 mean pulls the sd up; every update
    keeps sd <= 0.8 mean, where the relaxed trees stay finite."""
    from procell import workloads

    vals = [10.0 ** (4.0 * i / 63.0) for i in range(64)]
    counts = np.random.default_rng(3).poisson(40.0, 64).astype(np.uint64)
    W = workloads.Workload
    gen = W("wide", np.array(vals), counts, np.array([1.0]), np.array([40.0]), np.array([38.0]), 150.0, 0.0)
    start = W("wide0", np.array(vals), counts, np.array([1.0]), np.array([40.0]), np.array([31.0]), 150.0, 0.0)
    target = procell_lib.simulate(procell_lib.Problem.from_workload(gen, 5), want_raw=False)
    prob = procell_lib.Problem.from_workload(start, 6)
    with procell_lib.Calibration([prob], BETA, [target], lr=0.05) as cal:
        ratios = []
        for _ in range(8):
            cal.step(5)
            th = cal.theta()[0]
            ratios.append(th[1, 0] / th[0, 0])
        assert max(ratios) <= 0.8 * (1 + 1e-12), ratios


def test_full_size_calibration_sample_equals_torch_composition(procell_lib):
    """The 256-set calibration at full size (2e5 cells per set): after one
    step, three sets spread over the batch equal the single-set composition
    with torch."""
    from procell import workloads

    gen, start = workloads.calibration_batch(0, 256, 2e5)
    with procell_lib.Batch([procell_lib.Problem.from_workload(w, 0) for w in gen]) as b:
        b.run([1000 + i for i in range(len(gen))])
        targets = [b.result(i, want_raw=False) for i in range(len(gen))]
    probs = [procell_lib.Problem.from_workload(w, 50000 + i) for i, w in enumerate(start)]
    pick = [0, 137, 255]
    refs = {i: _torch_reference(procell_lib, probs[i], start[i], targets[i], 1) for i in pick}
    with procell_lib.Calibration(probs, BETA, targets, lr=LR, beta1=B1, beta2=B2, eps=EPS) as cal:
        cal.step(1)
        loss, _ = cal.loss()
        theta = cal.theta()
    for i in pick:
        l_ref, th_ref = refs[i][0]
        assert abs(loss[i] - l_ref) <= 1e-9 * abs(l_ref), (i, loss[i], l_ref)
        assert np.allclose(theta[i], th_ref, rtol=1e-9, atol=1e-12), (i, theta[i], th_ref)


def test_empty_target_and_invalid_start(procell_lib):
    """A set whose target histogram is empty (all-zero normalised target: the
    loss is the squared norm of the normalised relaxed histogram) steps like
    the others and stays finite; a proliferating start type with sd <= 0
    (its log parameter is undefined) is rejected."""
    gen, start, targets, probs = _sets(procell_lib, 2, 1000.0)
    empty = (np.zeros(0), np.zeros(0, dtype=np.int64))
    with procell_lib.Calibration(probs, BETA, [targets[0], empty], lr=LR) as cal:
        cal.step(3)
        loss, n = cal.loss()
        assert n == 3 and np.all(np.isfinite(loss)) and loss[1] > 0
        for th in cal.theta():
            assert np.all(np.isfinite(th))
    w = start[0]
    bad_sd = np.where(w.mean >= 0, 0.0, -1.0)
    bad = procell_lib.Problem(w.values, w.counts, w.proportion, w.mean, bad_sd, w.t_max, w.phi, 1)
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.Calibration([bad], BETA, [targets[0]])
    assert e.value.code == procell_lib.E_INVALID
