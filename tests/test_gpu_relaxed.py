"""GPU: the relaxed estimator and its adjoint against the oracle.

The device kernels sum with FP64 atomics in an order of their own and call
the device exp, so they agree with the oracle to rounding, not bit for bit:
every relaxed tally entry within 1e-10 relative (plus 1e-12 of the largest
entry), every gradient entry within 1e-9 relative plus 1e-10 of the largest
gradient entry.  The torch.autograd.Function must give the library's
gradient, and the gradient of a normalised-histogram loss must agree with
FP64 central differences of the device forward pass (common random numbers).
The plain method is not changed by relaxed calls.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H_RTOL, H_ATOL = 1e-10, 1e-12
G_RTOL, G_ATOL = 1e-9, 1e-10


def _problems():
    from procell import workloads

    w = workloads.small_histogram(4)
    return [
        # the small standard workload (3 types incl. quiescent), 1e5 cells
        (dict(values=w.values, counts=w.counts, proportion=w.proportion, mean=w.mean, sd=w.sd,
              t_max=w.t_max, phi=0.0, seed=7), 0.5),
        # a generation-0 cut (phi above some values), sharp beta
        (dict(values=[64.0, 20.0, 5.0, 2.5], counts=[600, 400, 300, 200], proportion=[0.45, 0.35, 0.2],
              mean=[14.0, 22.0, -1.0], sd=[4.0, 7.0, -1.0], t_max=70.0, phi=3.0, seed=11), 0.05),
        # deep rows (cap 20), wide beta, one type
        (dict(values=[2.0**20, 1000.0], counts=[50, 70], proportion=[1.0], mean=[20.0], sd=[3.0],
              t_max=75.0, phi=1.0, seed=3), 4.0),
        # t_max 0: everything records at generation 0 with weight 1 - sigmoid(-s/beta)
        (dict(values=[100.0, 37.5], counts=[40, 50], proportion=[0.5, 0.5], mean=[10.0, 20.0], sd=[3.0, 4.0],
              t_max=0.0, phi=1.0, seed=5), 1.0),
    ]


def _args(d):
    return (d["values"], d["counts"], d["proportion"], d["mean"], d["sd"], d["t_max"], d["phi"], d["seed"])


def _device_pass(procell_lib, d, beta, g=None, theta=None):
    import torch

    prob = procell_lib.Problem(*_args(d))
    th = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]) if theta is None else theta,
                      dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, beta) as rp:
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        rp.forward(th, h)
        grad = None
        if g is not None:
            gt = torch.tensor(g, dtype=torch.float64, device="cuda")
            grad_t = torch.empty_like(th)
            rp.backward(th, gt, grad_t)
            grad = grad_t.cpu().numpy()
        torch.cuda.synchronize()
        return h.cpu().numpy(), grad


@pytest.mark.parametrize("case", range(4))
def test_forward_and_gradient_match_oracle(procell_lib, oracle, case):
    d, beta = _problems()[case]
    print("relaxed case", case, "beta", beta, flush=True)
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), beta)
    g = np.random.default_rng(case).normal(size=h_ref.shape)
    _, grad_ref = oracle.relaxed(oracle.Problem(*_args(d)), beta, g)
    h, grad = _device_pass(procell_lib, d, beta, g)
    assert h.shape == h_ref.shape
    assert np.all(np.abs(h - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max())), \
        np.abs(h - h_ref).max()
    assert h.sum() > 0
    scale = max(1.0, np.abs(grad_ref).max())
    assert np.all(np.abs(grad - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale), (grad, grad_ref)


def test_all_quiescent_and_no_division(procell_lib, oracle):
    d = dict(values=[8.0, 3.0, 0.5], counts=[30, 20, 10], proportion=[0.6, 0.4], mean=[-1.0, -1.0],
             sd=[-1.0, -1.0], t_max=40.0, phi=1.0, seed=2)
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), 0.5)
    g = np.random.default_rng(9).normal(size=h_ref.shape)
    _, grad_ref = oracle.relaxed(oracle.Problem(*_args(d)), 0.5, g)
    h, grad = _device_pass(procell_lib, d, 0.5, g)
    assert np.allclose(h, h_ref, rtol=1e-14, atol=0)
    assert np.allclose(grad, grad_ref, rtol=1e-12, atol=1e-12)
    assert grad[0].tolist() == [0.0, 0.0] and grad[1].tolist() == [0.0, 0.0]


def test_autograd_function_and_finite_differences(procell_lib):
    import torch
    from procell import calibrate

    d, beta = _problems()[1]
    prob = procell_lib.Problem(*_args(d))
    with procell_lib.Relaxed(prob, beta * 20) as rp:
        bins, idx = calibrate.output_bins(prob, rp.shape[1])
        idx_t = torch.tensor(idx, device="cuda")
        target = torch.tensor(np.linspace(1.0, 2.0, len(bins)), dtype=torch.float64, device="cuda")

        def loss(th):
            h = calibrate.relaxed_tally(rp, th)
            return calibrate.normalized_l2(calibrate.binned(h, idx_t, len(bins)), target)

        th0 = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
        th = th0.clone().requires_grad_(True)
        L = loss(th)
        L.backward()
        grad = th.grad.cpu().numpy()
        # the Function's gradient is the library's backward with the same cotangent
        with torch.no_grad():
            w = torch.randn(rp.shape, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(1))
        th2 = th0.clone().requires_grad_(True)
        (calibrate.relaxed_tally(rp, th2) * w).sum().backward()
        direct = torch.empty_like(th0)
        rp.backward(th0, w, direct)
        torch.cuda.synchronize()
        assert torch.allclose(th2.grad, direct, rtol=1e-12, atol=1e-12 * float(direct.abs().max()))
        # central differences of the device forward
        for i, j in ((0, 0), (0, 1), (1, 0), (1, 1)):
            step = 1e-5 * abs(float(th0[i, j]))
            up, dn = th0.clone(), th0.clone()
            up[i, j] += step
            dn[i, j] -= step
            with torch.no_grad():
                fd = (loss(up) - loss(dn)).item() / (2 * step)
            assert abs(fd - grad[i, j]) <= 1e-5 * max(abs(fd), 1e-12) + 1e-12, (i, j, fd, grad[i, j])
        # proportions along p_a - p_b (the loss is not invariant to their sum)
        step = 1e-6
        up, dn = th0.clone(), th0.clone()
        up[2, 0] += step
        up[2, 1] -= step
        dn[2, 0] -= step
        dn[2, 1] += step
        with torch.no_grad():
            fd = (loss(up) - loss(dn)).item() / (2 * step)
        ad = grad[2, 0] - grad[2, 1]
        assert abs(fd - ad) <= 1e-5 * max(abs(fd), 1e-12) + 1e-12, (fd, ad)


def test_autograd_routes(procell_lib):
    """RelaxedTally walks once (Jacobian pass, contracted in backward) when
    theta needs a gradient and the plan has no stride; without a gradient it
    runs the forward pass only, and a stride set between forward and backward
    sends the backward to the adjoint pass on the cells that stride selects.
    Each route equals the library call it stands for."""
    import torch
    from procell import calibrate

    d, beta = _problems()[1]
    prob = procell_lib.Problem(*_args(d))
    print("autograd routes, beta", beta, flush=True)
    with procell_lib.Relaxed(prob, beta) as rp:
        th0 = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
        w = torch.randn(rp.shape, dtype=torch.float64, device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(3))
        h_ref = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        rp.forward(th0, h_ref)
        g_ref = torch.empty_like(th0)
        rp.backward(th0, w, g_ref)
        with torch.no_grad():
            h_ng = calibrate.relaxed_tally(rp, th0)
        th = th0.clone().requires_grad_(True)
        h_jac = calibrate.relaxed_tally(rp, th)
        (h_jac * w).sum().backward()
        tol = dict(rtol=1e-12, atol=1e-12 * float(h_ref.abs().max()))
        assert torch.allclose(h_ng, h_ref, **tol)
        assert torch.allclose(h_jac.detach(), h_ref, **tol)
        assert torch.allclose(th.grad, g_ref, rtol=1e-12, atol=1e-12 * float(g_ref.abs().max()))
        # a proportion of 0: dL/dp of that type is still the adjoint's
        thz = th0.clone()
        thz[2, 0] = 0.0
        g_z = torch.empty_like(th0)
        rp.backward(thz, w, g_z)
        xz = thz.clone().requires_grad_(True)
        (calibrate.relaxed_tally(rp, xz) * w).sum().backward()
        assert float(g_z[2, 0]) != 0.0
        assert torch.allclose(xz.grad, g_z, rtol=1e-12, atol=1e-12 * float(g_z.abs().max()))
        # a stride set after the forward pass: the adjoint on phase 1 of 2
        th2 = th0.clone().requires_grad_(True)
        h2 = calibrate.relaxed_tally(rp, th2)
        rp.set_stride(2, 1)
        (h2 * w).sum().backward()
        g_half = torch.empty_like(th0)
        rp.backward(th0, w, g_half)
        rp.set_stride(1)
        torch.cuda.synchronize()
        assert torch.allclose(th2.grad, g_half, rtol=1e-12, atol=1e-12 * float(g_half.abs().max()))
        assert not torch.allclose(g_half, g_ref, rtol=1e-3)


def test_plain_method_unchanged(procell_lib, oracle):
    import torch
    from procell import workloads

    w = workloads.small_histogram(1)
    prob = procell_lib.Problem.from_workload(w, 0)
    with procell_lib.Plan(prob) as pl:
        pl.run(123)
        before = pl.result()
    th = torch.tensor(np.stack([w.mean, w.sd, w.proportion]), dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, 0.5) as rp:
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        rp.forward(th, h)
        g = torch.ones_like(h)
        grad = torch.empty_like(th)
        rp.backward(th, g, grad)
        torch.cuda.synchronize()
    with procell_lib.Plan(prob) as pl:
        pl.run(123)
        after = pl.result()
    assert np.array_equal(before.raw_tally, after.raw_tally)
    tally, _, _ = oracle.simulate(oracle.Problem.from_workload(w, 123))
    assert np.array_equal(after.raw_tally, tally)


def test_stream_order_with_torch_work(procell_lib, oracle):
    """theta written by a torch kernel just before the call (no
    synchronisation) on the default stream (NULL) and on a side stream: the
    library's kernels run after it, on the same stream."""
    import torch

    d, beta = _problems()[3]
    prob = procell_lib.Problem(*_args(d))
    ref = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), beta)
    side = torch.cuda.Stream()
    with procell_lib.Relaxed(prob, beta) as rp:
        for st in (None, side):
            torch.cuda.synchronize()
            with torch.cuda.stream(st if st is not None else torch.cuda.default_stream()):
                th = torch.zeros_like(ref)
                for _ in range(20):  # queue work ahead of the write that matters
                    th.mul_(1.0)
                th.add_(ref)
                h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
                rp.forward(th, h, stream=0 if st is None else st.cuda_stream)
                got = h.cpu().numpy() if st is None else None
            if st is not None:
                side.synchronize()
                got = h.cpu().numpy()
            assert np.all(np.abs(got - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max()))


def test_global_fallbacks_match_oracle(procell_lib, oracle):
    """One cell per row and few wavefronts (blocks_per_cu = 1): every
    wavefront's range spans more rows than the LDS row stage holds (64) and a
    tally window larger than the LDS window (1024 entries), so the kernel
    searches rows in device memory and records with global FP64 atomics."""
    import torch

    n = 40000
    values = [10.0 ** (5.0 * i / (n - 1)) for i in range(n)]
    d = dict(values=values, counts=[1] * n, proportion=[0.7, 0.3], mean=[30.0, -1.0], sd=[8.0, -1.0],
             t_max=100.0, phi=1.0, seed=21)
    beta = 0.5
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), beta)
    g = np.random.default_rng(5).normal(size=h_ref.shape)
    _, grad_ref = oracle.relaxed(oracle.Problem(*_args(d)), beta, g)
    prob = procell_lib.Problem(*_args(d), blocks_per_cu=1)
    th = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, beta) as rp:
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        rp.forward(th, h)
        gt = torch.tensor(g, dtype=torch.float64, device="cuda")
        grad = torch.empty_like(th)
        rp.backward(th, gt, grad)
        # the Jacobian pass records its three tallies with global atomics too
        hj = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        jac = torch.empty((2,) + tuple(rp.shape), dtype=torch.float64, device="cuda")
        rp.jacobian(th, hj, jac)
        h, grad, hj, jac = h.cpu().numpy(), grad.cpu().numpy(), hj.cpu().numpy(), jac.cpu().numpy()
    assert np.all(np.abs(h - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max()))
    assert np.all(np.abs(hj - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max()))
    scale = max(1.0, np.abs(grad_ref).max())
    assert np.all(np.abs(grad - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale), (grad, grad_ref)
    p = np.asarray(d["proportion"])
    gj = np.stack([(g * jac[0]).sum((0, 1)), (g * jac[1]).sum((0, 1)), (g * hj).sum((0, 1)) / p])
    assert np.all(np.abs(gj - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale), (gj, grad_ref)


def test_global_fallbacks_of_a_later_batch_set(procell_lib, oracle):
    """A batch plan whose second set takes the global fallbacks (rows and
    tallies beyond the LDS stage and window, blocks_per_cu = 1): its h,
    Jacobian and gradient sit at an offset into the concatenated arrays."""
    import torch

    small = _problems()[1][0]
    n = 30000
    values = [10.0 ** (5.0 * i / (n - 1)) for i in range(n)]
    big = dict(values=values, counts=[1] * n, proportion=[0.6, 0.4], mean=[30.0, 45.0], sd=[8.0, 9.0],
               t_max=100.0, phi=1.0, seed=22)
    cases = [small, big]
    beta = 0.5
    probs = [procell_lib.Problem(*_args(d), blocks_per_cu=1) for d in cases]
    rng = np.random.default_rng(12)
    with procell_lib.RelaxedBatch(probs, beta) as bp:
        th = torch.tensor(np.concatenate([np.stack([d["mean"], d["sd"], d["proportion"]]).reshape(-1)
                                          for d in cases]), dtype=torch.float64, device="cuda")
        g_np = rng.normal(size=bp.h_total)
        h = torch.empty(bp.h_total, dtype=torch.float64, device="cuda")
        jac = torch.empty((2, bp.h_total), dtype=torch.float64, device="cuda")
        grad = torch.empty(3 * bp.types_total, dtype=torch.float64, device="cuda")
        bp.jacobian(th, h, jac)
        bp.backward(th, torch.tensor(g_np, dtype=torch.float64, device="cuda"), grad)
        h, jac, grad = h.cpu().numpy(), jac.cpu().numpy(), grad.cpu().numpy()
        shapes, h_base, t_base = bp.shapes, bp.h_base, bp.theta_base
    for i, d in enumerate(cases):
        op = oracle.Problem(*_args(d))
        h_ref, _ = oracle.relaxed(op, beta)
        sl = slice(h_base[i], h_base[i] + h_ref.size)
        gi = g_np[sl].reshape(h_ref.shape)
        _, grad_ref = oracle.relaxed(op, beta, gi)
        hi = h[sl].reshape(h_ref.shape)
        assert np.all(np.abs(hi - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max())), i
        T = shapes[i][2]
        scale = max(1.0, np.abs(grad_ref).max())
        gr = grad[t_base[i]:t_base[i] + 3 * T].reshape(3, T)
        assert np.all(np.abs(gr - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale), (i, gr, grad_ref)
        p = np.asarray(d["proportion"])
        gj = np.stack([(gi * jac[0, sl].reshape(h_ref.shape)).sum((0, 1)),
                       (gi * jac[1, sl].reshape(h_ref.shape)).sum((0, 1)), (gi * hi).sum((0, 1)) / p])
        assert np.all(np.abs(gj - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale), (i, gj, grad_ref)


def test_many_types_match_oracle(procell_lib, oracle):
    """70 types (69 proliferating, one quiescent): gradient sums and the
    staged type list cover more than one wavefront's width."""
    rng = np.random.default_rng(8)
    T = 70
    means = list(rng.uniform(25.0, 60.0, T - 1)) + [-1.0]
    sds = [0.25 * m for m in means[:-1]] + [-1.0]
    p = rng.dirichlet(np.ones(T))
    d = dict(values=[300.0, 90.0, 12.0], counts=[40, 30, 20], proportion=list(p), mean=means, sd=sds,
             t_max=110.0, phi=2.0, seed=4)
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), 1.0)
    g = np.random.default_rng(6).normal(size=h_ref.shape)
    _, grad_ref = oracle.relaxed(oracle.Problem(*_args(d)), 1.0, g)
    h, grad = _device_pass(procell_lib, d, 1.0, g)
    assert np.all(np.abs(h - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max()))
    scale = max(1.0, np.abs(grad_ref).max())
    assert np.all(np.abs(grad - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale)


def test_strided_passes_sum_to_full(procell_lib):
    """With cell stride 4 the passes of phases 0..3 walk disjoint sets of
    initial cells covering all of them: their tallies and gradients sum to
    the full pass's (to rounding: different atomic orders)."""
    import torch

    d, beta = _problems()[0]
    prob = procell_lib.Problem(*_args(d))
    th = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, beta) as rp:
        g = torch.tensor(np.random.default_rng(4).normal(size=rp.shape), dtype=torch.float64, device="cuda")
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        gr = torch.empty_like(th)
        rp.forward(th, h)
        rp.backward(th, g, gr)
        full_h, full_g = h.cpu().numpy(), gr.cpu().numpy()
        n_full = rp.expansions()
        sum_h, sum_g, n_sum = np.zeros_like(full_h), np.zeros_like(full_g), 0
        for phase in range(4):
            rp.set_stride(4, phase)
            rp.forward(th, h)
            n_sum += rp.expansions()
            rp.backward(th, g, gr)
            part = h.cpu().numpy()
            assert 0.15 * full_h.sum() < part.sum() < 0.35 * full_h.sum()
            sum_h += part
            sum_g += gr.cpu().numpy()
        rp.set_stride(1)
    assert n_sum == n_full
    assert np.all(np.abs(sum_h - full_h) <= 1e-12 * np.abs(full_h) + 1e-12 * np.abs(full_h).max())
    assert np.all(np.abs(sum_g - full_g) <= 1e-10 * np.abs(full_g) + 1e-10 * np.abs(full_g).max())


def test_full_size_calibration_instance_matches_oracle(procell_lib, oracle):
    """The single-instance calibration workload at full size (256 bins,
    ~1e6 cells, two proliferating types and a quiescent one, beta 0.5):
    forward and adjoint against the OpenMP oracle."""
    from procell import workloads

    w = workloads.calibration_instance(0)
    d = dict(values=w.values, counts=w.counts, proportion=w.proportion, mean=w.mean, sd=w.sd,
             t_max=w.t_max, phi=w.phi, seed=1)
    beta = w.meta["beta"]
    print("full-size relaxed parity:", w.n_cells, "cells", flush=True)
    h_ref, _ = oracle.relaxed(oracle.Problem(*_args(d)), beta)
    g = np.random.default_rng(12).normal(size=h_ref.shape)
    _, grad_ref = oracle.relaxed(oracle.Problem(*_args(d)), beta, g)
    h, grad = _device_pass(procell_lib, d, beta, g)
    assert np.all(np.abs(h - h_ref) <= H_RTOL * np.abs(h_ref) + H_ATOL * max(1.0, np.abs(h_ref).max()))
    scale = max(1.0, np.abs(grad_ref).max())
    assert np.all(np.abs(grad - grad_ref) <= G_RTOL * np.abs(grad_ref) + G_ATOL * scale)


def test_runaway_walk_is_cut_and_reported(procell_lib):
    """sd = 3 x mean with 29 generations allowed and a far t_max: every
    walk would expand a full 2^29-node tree.  Walks stop at 2^22 expansions
    and the call reports PROCELL_E_DEPTH instead of running for hours."""
    import torch

    d = dict(values=[2.0**30], counts=[4], proportion=[1.0], mean=[10.0], sd=[30.0], t_max=1000.0, phi=1.0, seed=1)
    prob = procell_lib.Problem(*_args(d))
    th = torch.tensor(np.stack([d["mean"], d["sd"], d["proportion"]]), dtype=torch.float64, device="cuda")
    with procell_lib.Relaxed(prob, 1.0) as rp:
        h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
        rp.forward(th, h)
        with pytest.raises(procell_lib.ProcellError) as e:
            rp.expansions()
        assert e.value.code == procell_lib.E_DEPTH
        # a later call within the domain is clean again
        th2 = torch.tensor([[100.0], [10.0], [1.0]], dtype=torch.float64, device="cuda")
        rp.forward(th2, h)
        assert rp.expansions() > 0


def test_tiny_beta_one_type_is_the_plain_tally(procell_lib):
    """beta -> 0, one type: the relaxed walks take the plain method's timers
    (redraw streams included) and its hard test, so the device's relaxed
    tally is the device's plain tally of the same seed (integers summed in
    FP64, exact in any order)."""
    import torch

    values = [2.0**12, 3000.0, 700.0, 40.0, 9.0]
    counts = [3000, 2000, 5000, 1000, 500]
    for mean, sd, seed in ((20.0, 10.0, 3), (30.0, 15.0, 4)):
        print("tiny beta", mean, sd, seed, flush=True)
        prob = procell_lib.Problem(values, counts, [1.0], [mean], [sd], 150.0, 10.0, seed)
        plain = procell_lib.simulate(prob)
        assert plain.stats["rejected"] > 0  # redraw streams exercised
        th = torch.tensor([[mean], [sd], [1.0]], dtype=torch.float64, device="cuda")
        with procell_lib.Relaxed(prob, 1e-9) as rp:
            h = torch.empty(rp.shape, dtype=torch.float64, device="cuda")
            rp.forward(th, h)
            torch.cuda.synchronize()
        assert np.array_equal(h.cpu().numpy(), plain.raw_tally.astype(np.float64))
