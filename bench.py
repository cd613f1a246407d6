"""Benchmark: cell divisions per second of the plain proliferation method on
one MI355X.

A step is one complete simulation (procell_plan_run: clear the tallies, the
root kernel, the persistent division-tree kernel) of the 1024-bin, ~1e7-cell histogram
with five cell types (four proliferating, one quiescent), t_max 240, phi 50
and per-type tracking.  Inputs are uploaded once before timing (resident in
HBM); every step uses a different seed.  value = divisions summed over the
timed steps / wall time between the two synchronisations.

Prints one JSON line on stdout; everything else goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "cuda-pro-cell_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "cell divisions/s on 1 GPU (plain) + calib grad-steps/s batched vs plain"
# Nominal VALU work per division (lane operations): Philox4x32-10 block,
# two 53-bit uniforms, FP64 Box-Muller pair, timer update and compares.
LANE_OPS_PER_DIVISION = 180
# MI355X VALU issue peak: 256 CUs x 64 lanes/clk x 2.4 GHz (lane-ops/s).
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
TRAFFIC_FILE = os.path.join(REPO, "profiles", "simulation_kernels_hbm_traffic.txt")
COUNTER_FILE = os.path.join(REPO, "profiles", "simulation_kernels_valu_counters.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profile():
    """HBM bytes per run (root + tree kernel) from the committed rocprofv3
    --pmc summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r"hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def issue_frac_from_profile():
    """VALU issue utilisation (SQ_INSTS_VALU / (CUs x cycles)) from the
    committed rocprofv3 counter summary."""
    try:
        with open(COUNTER_FILE) as f:
            m = re.search(r"VALU issue utilisation.*=\s*([0-9.]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, budget_s: float):
    """The CPU oracle (OpenMP, all host cores of this box) on the same
    workload, seeds 0, 1, ... until about budget_s of wall time."""
    from oracle import binding as ob

    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    divs, seeds, t0 = 0, 0, time.perf_counter()
    while True:
        _, _, st = ob.simulate(ob.Problem.from_workload(workload, seeds), threads)
        divs += st["divisions"]
        seeds += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    # one core on a 1/20 sample of the cells
    from procell import workloads

    small = workloads.scaled(workload, 0.05)
    t1 = time.perf_counter()
    _, _, st1 = ob.simulate(ob.Problem.from_workload(small, 0), 1)
    dt1 = time.perf_counter() - t1
    return {
        "value": divs / dt,
        "unit": "divisions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full workload, seeds 0..{seeds - 1} ({divs} divisions, {dt:.2f} s wall), OpenMP oracle",
        "cpu": cpu_model(),
        "one_core_value": st1["divisions"] / dt1,
        "one_core_sample": f"5% of the cells, seed 0 ({st1['divisions']} divisions, {dt1:.2f} s)",
    }


def batched_measurement(procell, workloads, np, stream, runs=20):
    """The batched plain method on 512 heterogeneous parameter sets (one
    launch of each kernel per batch run, replayed from a hipGraph; knobs from
    the library's autotuner), beside the same sets run one plan at a time
    (a 64-set sample)."""
    ws = workloads.parameter_sets(0, 512)
    probs = [procell.Problem.from_workload(w, 0, flags=procell.FLAG_TIMING if i == 0 else 0)
             for i, w in enumerate(ws)]
    b = procell.Batch(probs, stream=stream)
    seeds = np.arange(len(ws), dtype=np.uint64) + 5000
    chosen = b.autotune(seeds, reps=1)
    for r in range(3):
        b.run(seeds + np.uint64(r))
    b.kernel_times()
    for r in range(runs):
        b.run(seeds + np.uint64(10 + r))
    kt = b.kernel_times()
    divs = 0
    for r in range(runs):
        b.run(seeds + np.uint64(10 + r))
        divs += sum(b.result(i, want_raw=False).stats["divisions"] for i in range(len(ws)))
    b.close()
    ms = float(np.mean(kt))
    value = divs / (runs * ms * 1e-3)
    n = 64
    plans = [procell.Plan(p) for p in probs[:n]]
    for i, p in enumerate(plans):
        p.run(int(seeds[i]))
    for p in plans:
        p.result(want_raw=False)
    t0 = time.perf_counter()
    for i, p in enumerate(plans):
        p.run(int(seeds[i]) + 1)
    d2 = sum(p.result(want_raw=False).stats["divisions"] for p in plans)
    el = time.perf_counter() - t0
    for p in plans:
        p.close()
    log(f"batched: {len(ws)} sets, {ms:.3f} ms per batch run, {value:.4g} div/s; sequential plans {d2 / el:.4g} div/s")
    return {
        "value": value,
        "unit": "divisions/s",
        "workload": "512 parameter sets: 128-1024 bins, 1e3-1e6 cells (lognormal), 1-5 types, "
                    "t_max U(72,336), phi U(1,100) (workloads.parameter_sets(0, 512))",
        "cells": int(sum(w.n_cells for w in ws)),
        "divisions_per_run": divs / runs,
        "ms_per_batch_run": ms,
        "runs": runs,
        "autotuned": chosen,
        "sequential_plans_value": d2 / el,
        "sequential_plans_sample": f"first {n} sets, one plan each, run back to back ({el * 1e3:.2f} ms wall)",
    }


def relaxed_measurement(procell, workloads, np, torch, with_cpu=True, reps=50, fit_steps=300):
    """The relaxed estimator on the single-instance calibration workload
    (256 bins, ~1e6 cells, two proliferating types and one quiescent, beta
    0.5): device time of one forward and one adjoint pass, gradient steps
    per second of an Adam fit through the torch binding, and the parameters
    that fit recovers from a perturbed start against a plain-method target
    at the hidden parameters."""
    from procell import calibrate

    w = workloads.calibration_instance(0)
    beta = w.meta["beta"]
    hidden = np.stack([w.mean, w.sd, w.proportion])
    prolif = w.mean >= 0
    target_res = procell.simulate(procell.Problem.from_workload(w, 0), want_raw=False)
    prob = procell.Problem.from_workload(w, 1)
    plan = procell.Relaxed(prob, beta)
    bins, _ = calibrate.output_bins(prob, plan.shape[1])
    target = calibrate.target_histogram(bins, target_res.value, target_res.count)
    start = hidden.copy()
    start[0, prolif] *= np.array([1.15, 0.87])[: prolif.sum()]
    start[1, prolif] *= np.array([1.25, 0.8])[: prolif.sum()]
    start[2] = 0.5 * start[2] + 0.5 / len(start[2])
    th = torch.tensor(start, dtype=torch.float64, device="cuda")
    h = torch.empty(plan.shape, dtype=torch.float64, device="cuda")
    g = torch.randn(plan.shape, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(0))
    grad = torch.empty_like(th)
    times = {}
    for name, fn in (("forward", lambda: plan.forward(th, h)), ("backward", lambda: plan.backward(th, g, grad))):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1) / reps
        times[name + "_expansions"] = plan.expansions()
    # loss at the hidden parameters and between two plain samples (noise floor)
    bm = calibrate.BinMap(prob, plan.shape[1])
    tgt = torch.tensor(target, dtype=torch.float64, device="cuda")
    with torch.no_grad():
        loss_hidden = float(calibrate.normalized_l2(
            bm(calibrate.relaxed_tally(plan, torch.tensor(hidden, dtype=torch.float64, device="cuda"))), tgt))
    other = procell.simulate(procell.Problem.from_workload(w, 7), want_raw=False)
    t2 = calibrate.target_histogram(bins, other.value, other.count)
    noise_floor = float(((t2 / t2.sum() - target / target.sum()) ** 2).sum())
    calibrate.fit(plan, prob, target, start, steps=10)  # warm-up (torch kernels, allocator)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    theta, losses = calibrate.fit(plan, prob, target, start, steps=fit_steps)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    plan.close()
    rel = np.abs(theta - hidden) / np.abs(hidden)
    err = {"mean_rel": rel[0, prolif].tolist(), "sd_rel": rel[1, prolif].tolist(),
           "proportion_abs": np.abs(theta[2] - hidden[2]).tolist()}
    exp_pass = times["forward_expansions"]
    pass_ms = times["forward"] + times["backward"]
    log(f"relaxed: forward {times['forward']:.4f} ms, backward {times['backward']:.4f} ms, "
        f"{exp_pass} expansions per pass; fit {fit_steps} steps {fit_s:.2f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    out = {
        "workload": "hist256_1e6_types3_calibration: 256 log-spaced bins 1..1e4, lognormal Poisson counts "
                    "(~1e6 cells), 2 proliferating types + 1 quiescent, t_max 168, phi auto, beta 0.5 "
                    "(workloads.calibration_instance(0))",
        "cells": w.n_cells,
        "beta": beta,
        "forward_ms": times["forward"],
        "backward_ms": times["backward"],
        "expansions_per_pass": exp_pass,
        "expansions_per_s": 2 * exp_pass / (pass_ms * 1e-3),
        "grad_steps_per_s_kernels": 1e3 / pass_ms,
        "grad_steps_per_s_fit": fit_steps / fit_s,
        "fit": {"steps": fit_steps, "optimizer": "Adam lr 0.02 on log mean, log sd, proportion logits",
                "loss_start": losses[0], "loss_end": losses[-1], "loss_at_hidden": loss_hidden,
                "plain_noise_floor": noise_floor,
                "note": "loss at the hidden parameters equals the distance between two plain samples "
                        "(seeds 0 and 7): the relaxation's bias is below the sampling noise; the fit "
                        "reaches that floor along a flat valley, so one 1e6-cell histogram identifies "
                        "(mean, sd, proportion) only loosely",
                "target": "plain method at the hidden parameters, seed 0; relaxed stream seed 1",
                "start": start.tolist(), "hidden": hidden.tolist(), "recovered": theta.tolist(),
                "relative_error": err},
        "dtype": "f64",
    }
    if with_cpu:
        from oracle import binding as ob

        threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        op = ob.Problem(w.values, w.counts, start[2], start[0], start[1], w.t_max, w.phi, 1)
        gh = g.cpu().numpy()
        t1 = time.perf_counter()
        ob.relaxed(op, beta, None, threads)
        ob.relaxed(op, beta, gh, threads)
        dt = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "grad-steps/s (forward + adjoint)", "cores": threads,
                               "kind": "port", "sample": f"full workload, one forward and one adjoint pass "
                                                         f"({dt:.2f} s wall), OpenMP oracle",
                               "cpu": cpu_model()}
        log(f"relaxed cpu baseline: {dt:.3f} s per forward + adjoint on {threads} threads")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=1.5, help="seconds of CPU-baseline wall time")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-method measurement")
    ap.add_argument("--no-relaxed", action="store_true", help="skip the relaxed-estimator measurement")
    ap.add_argument("--fit-steps", type=int, default=300, help="Adam steps of the relaxed-estimator fit")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import procell
    from procell import workloads

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    if args.copy_calib:
        n = args.copy_calib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {n} bytes read, {n} bytes written")
        del src, dst
    w = workloads.large_histogram(0)
    log(f"workload {w.name}: {w.n_cells} cells, {len(w.values)} bins, {len(w.proportion)} types")
    stream = torch.cuda.Stream()
    prob = procell.Problem.from_workload(w, 0, flags=procell.FLAG_TIMING)
    plan = procell.Plan(prob, stream=stream.cuda_stream)
    K, W = args.steps, args.warmup
    seeds = [1000 + i for i in range(K)]
    for i in range(W):
        plan.run(i)
    torch.cuda.synchronize()
    plan.kernel_times()  # drop warm-up timings
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for s in seeds:
        plan.run(s)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = start.elapsed_time(stop)
    kt = plan.kernel_times()
    # divisions of each timed seed (re-run untimed; results are a pure
    # function of the seed)
    total_div = 0
    stats = None
    for s in seeds:
        plan.run(s)
        stats = plan.result(want_raw=False).stats
        total_div += stats["divisions"]
    plan.close()
    # end-to-end C-ABI call (host rows in, host histogram out: plan set-up,
    # uploads, run, download, merge), reported beside the metric
    e2e = []
    for s in seeds[:5]:
        t1 = time.perf_counter()
        procell.simulate(procell.Problem.from_workload(w, s), want_raw=False)
        e2e.append((time.perf_counter() - t1) * 1e3)
    elapsed = ev_ms * 1e-3
    value = total_div / elapsed
    kernel_ms = float(np.mean(kt)) if len(kt) else float("nan")
    div_per_run = total_div / K
    achieved = div_per_run * LANE_OPS_PER_DIVISION / (kernel_ms * 1e-3)
    log(f"timed {K} steps: {ev_ms:.2f} ms events, {wall * 1e3:.2f} ms wall, kernel {kernel_ms:.4f} ms avg, "
        f"{total_div} divisions, {value:.4g} div/s")
    traffic = traffic_from_profile()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "divisions/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ev_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "hist1024_1e7_types5: 1024 log-spaced bins 1..1e5, lognormal Poisson counts "
                        "(~1e7 cells), 4 proliferating types + 1 quiescent, Dirichlet ratios, t_max 240, "
                        "phi 50, per-type tracking",
            "cells": w.n_cells,
            "bins": int(len(w.values)),
            "types": int(len(w.proportion)),
            "divisions_per_step": div_per_run,
            "seeds": f"{seeds[0]}..{seeds[-1]}",
            "workgroups": stats["workgroups"],
            "block_size": stats["block_size"],
            "lds_bytes": stats["lds_bytes"],
            "abi_end_to_end_ms": float(np.median(e2e)),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": VALU_PEAK_LANE_OPS / 1e12,
            "unit": "Tlane-op/s",
            "frac": achieved / VALU_PEAK_LANE_OPS,
            "traffic": traffic,
            "kernel": "zero_kernel + root_kernel + tree_kernel (one run, back to back, timed together)",
            "kernel_ms": kernel_ms,
            "work_per_unit": f"{LANE_OPS_PER_DIVISION} lane-ops per division",
            "peak_source": "MI355X FP64 vector peak 78.6 TFLOP/s / 2 = one wave64 VALU instruction "
                           "per CU per clock (256 CUs x 64 lanes x 2.4 GHz)",
            "issue_frac_counters": issue_frac_from_profile(),
        },
    }
    if not args.no_batch:
        out["batched"] = batched_measurement(procell, workloads, np, stream.cuda_stream)
    if not args.no_relaxed:
        out["relaxed"] = relaxed_measurement(procell, workloads, np, torch, with_cpu=not args.no_cpu,
                                             fit_steps=args.fit_steps)
    if not args.no_cpu:
        cb = cpu_baseline(w, args.cpu_budget)
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.4g} div/s on {cb['cores']} cores; 1 core {cb['one_core_value']:.4g}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
