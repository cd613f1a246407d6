"""CPU checks of bench.py's printed line: headline fields first, every
trajectory left to the detail file, and the line short enough for a reader
that keeps only the tail of stdout."""
import json

import bench


def _recovery(i):
    return {"mean_rel_median": 0.07 / (i + 1), "mean_rel_p90": 0.2, "mean_within_2pct": 0.4123456789,
            "sd_rel_median": 0.08, "sd_rel_p90": 0.39, "sd_within_2pct": 0.25,
            "proportion_abs_median": 0.03, "proportion_within_0.02": 0.44}


def _traj(n, key="steps"):
    return [{key: 25 * i, "gpu_s": 0.1234567 * i, "loss_median": 4.7e-6, **_recovery(i)} for i in range(n)]


def _full_record():
    """A record shaped like a full default run (all legs, long trajectories)."""
    return {
        "metric": bench.METRIC, "value": 7.123456789e10, "unit": "divisions/s", "n_gpus": 1, "steps": 200,
        "warmup": 20, "ms_per_step": 0.3412345, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "hist1024_1e7_types5: " + "x" * 200, "cells": 9999325, "bins": 1024, "types": 5,
                   "divisions_per_step": 25840000.0, "seeds": "1000..1199", "abi_end_to_end_ms": 310.2,
                   "knobs": {"workgroups": 768, "block_size": 256, "lds_bytes": 40000, "window_rows": 30}},
        "roofline": {"bound": "valu", "achieved": 13.6, "peak": 39.32, "unit": "Tlane-op/s", "frac": 0.3461234,
                     "traffic": 155069128.0, "kernel": "k" * 100, "kernel_ms": 0.342, "work_per_unit": "180 lane-ops",
                     "peak_source": "p" * 150, "valu_issue_all": 0.746, "valu_frac_int32_f64_cvt": 0.31,
                     "hbm_gbs": 453.1},
        "batched": {"value": 8.4e10, "unit": "divisions/s", "workload": "w" * 200, "cells": 52000000,
                    "divisions_per_run": 2.58e8, "ms_per_batch_run": 3.05, "runs": 20,
                    "autotuned": {"segments": 8, "idle_lanes": 16, "blocks_per_cu": 3},
                    "sequential_plans_value": 2.3e9, "sequential_plans_sample": "s" * 100},
        "relaxed": {"workload": "w" * 250, "cells": 999517, "beta": 0.5, "forward_ms": 0.149, "backward_ms": 0.163,
                    "jacobian_ms": 0.16, "expansions_per_pass": 8200000, "expansions_per_s": 5.3e10,
                    "roofline": {"bound": "valu", "forward_frac": 0.33, "adjoint_frac": 0.35},
                    "grad_steps_per_s_kernels": 3200.0, "grad_steps_per_s_fit": 1550.0,
                    "fit": {"steps": 300, "loss_start": 1e-4, "loss_end": 1.3e-6, "plain_noise_floor": 1.32e-6,
                            "start": [[1.0] * 3] * 3, "hidden": [[1.0] * 3] * 3, "recovered": [[1.0] * 3] * 3,
                            "note": "n" * 300},
                    "cpu_baseline": {"value": 2.1, "unit": "grad-steps/s", "cores": 16, "kind": "port",
                                     "sample": "s" * 100}},
        "calibration": {"workload": "w" * 300, "cells": 51200000, "steps": 200, "ms_per_step": 23.4,
                        "grad_steps_per_s": 42.7, "roofline_frac_nominal": 0.47, "trajectory": _traj(9),
                        "strided": {"cell_stride": 4, "ms_per_step": 10.2, "trajectory": _traj(9)},
                        "levenberg_marquardt": {"iterations": 15, "ms_per_iteration": 16.0, "walks_cut": False,
                                                "gpu_time_ratio_adam_over_lm_at_loss": 27.0,
                                                "trajectory": _traj(16)},
                        "derivative_free": {"gpu_s": 0.95, "wall_s": 38.0, "trajectory": _traj(12, "rounds"),
                                            "gpu_time_ratio_at_common": 0.3, "gpu_time_ratio_strided_at_common": 0.93,
                                            "gpu_time_ratio_levenberg_marquardt_at_common": 1.3, "note": "n" * 200},
                        "rich_target": {"summary": {"gpu_time_ratio_at_2pct": 11.0, "lm_mean_rel_median": 0.01,
                                                    "nelder_mead_mean_rel_median": 0.02},
                                        "trajectory": _traj(20)},
                        "cpu_baseline": {"value": 0.015, "unit": "grad-steps/s", "cores": 16}},
        "cpu_baseline": {"value": 3.06e8, "unit": "divisions/s", "cores": 16, "kind": "port",
                         "sample": "median of 5 runs of the full workload (seeds 0..4, 129000000 divisions, 1.52 s wall)",
                         "cpu": "AMD EPYC", "one_core_value": 2.0e7, "one_core_sample": "5% of the cells"},
        "detail_file": "gpurun_out/bench_detail.json",
    }


def test_line_is_short_and_headline_first():
    line = json.dumps(bench.compact_line(_full_record()))
    assert len(line) <= bench.LINE_LIMIT, len(line)
    obj = json.loads(line)
    keys = list(obj)
    assert keys[:3] == ["metric", "value", "unit"]
    for k in ("ms_per_step", "steps", "warmup", "dtype", "config", "roofline", "cpu_baseline"):
        assert k in obj
    assert keys.index("roofline") < keys.index("batched") and keys.index("cpu_baseline") < keys.index("batched")
    assert obj["roofline"]["frac"] == 0.34612 and obj["roofline"]["valu_frac_int32_f64_cvt"] == 0.31
    assert obj["cpu_baseline"]["value"] == 3.06e8 and obj["cpu_baseline"]["cores"] == 16
    assert "trajectory" not in line


def test_line_keeps_the_summaries():
    obj = bench.compact_line(_full_record())
    cal = obj["calibration"]
    assert cal["adam_final"]["mean_rel_median"] == float(f"{0.07 / 9:.5g}")
    assert cal["lm"]["gpu_time_ratio_adam_over_lm_at_loss"] == 27.0
    assert cal["nelder_mead"]["gpu_time_ratio_at_common"] == 0.3
    assert cal["rich_target"]["gpu_time_ratio_at_2pct"] == 11.0
    assert obj["batched"]["autotuned"]["segments"] == 8
    assert obj["relaxed"]["forward_frac"] == 0.33


def test_minimal_record():
    """A run with every extra leg switched off still gives a valid line."""
    full = _full_record()
    for k in ("batched", "relaxed", "calibration", "cpu_baseline", "detail_file"):
        full.pop(k)
    obj = json.loads(json.dumps(bench.compact_line(full)))
    assert obj["value"] == 7.1235e10 and obj["cpu_baseline"] == {}
