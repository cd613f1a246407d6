"""GPU end-to-end: the procell CLI (file in, file out) against the oracle's
own file pipeline, byte for byte, and size-independent properties of full
simulations (event balance, per-row mass) at the 1e7-cell workload."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
HIST = os.path.join(GOLDEN, "readme_histogram.txt")
TYPES = os.path.join(GOLDEN, "readme_types.txt")


def _run_cli(procell_lib, args):
    print("cli", args, flush=True)
    return subprocess.run([procell_lib.CLI_PATH] + args, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("t_max,ratio,seed", [(168.0, True, 5), (1000.0, True, 6), (400.0, False, 2**64 - 1)])
def test_cli_matches_oracle_files(procell_lib, oracle, tmp_path, t_max, ratio, seed):
    out = str(tmp_path / "gpu.txt")
    ref = str(tmp_path / "oracle.txt")
    args = ["-h", HIST, "-c", TYPES, "-t", repr(t_max), "-o", out, "--seed", str(seed)]
    if ratio:
        args.append("-r")
    r = _run_cli(procell_lib, args)
    assert r.returncode == 0, r.stdout + r.stderr
    assert oracle.run_files(HIST, TYPES, t_max, 0.0, ratio, seed, ref, 0) == 0
    got, want = open(out).read(), open(ref).read()
    assert got == want
    assert len(want.splitlines()) >= 2
    bins = set(json.load(open(os.path.join(GOLDEN, "readme_bins.json")))["bins"])
    for line in got.splitlines():
        v = float(line.split("\t")[0])
        assert any(v == b or line.split("\t")[0] == "%.10g" % b for b in bins)


def test_cli_stdout_and_phi_alias(procell_lib, oracle, tmp_path):
    ref = str(tmp_path / "oracle.txt")
    r = _run_cli(procell_lib, ["--histogram", HIST, "--cell-types", TYPES, "--time-max", "200", "--phi", "20",
                               "-s", "9"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert oracle.run_files(HIST, TYPES, 200.0, 20.0, False, 9, ref, 0) == 0
    assert r.stdout == open(ref).read()


def test_full_size_event_balance_and_mass(procell_lib):
    """1e7 cells: every division adds one cell, every cell ends recorded,
    dropped or lost; the recorded mass of a row at generation k is 2^-k per
    cell, so sum_k tally 2^-k <= count per row, with equality for rows whose
    cells never reach their cap in time."""
    from procell import workloads

    w = workloads.large_histogram(1)
    res = procell_lib.simulate(procell_lib.Problem.from_workload(w, 77))
    st = res.stats
    assert st["cells"] == w.n_cells
    assert st["recorded"] + st["dropped"] + st["lost"] == st["cells"] + st["divisions"]
    assert int(res.raw_tally.sum()) == st["recorded"]
    assert int(res.count.sum()) == st["recorded"]
    kdim = res.raw_tally.shape[1]
    wts = np.array([2.0**-k for k in range(kdim)])
    mass = res.raw_tally.sum(axis=2) @ wts
    counts = w.counts[w.counts > 0].astype(np.float64)
    assert np.all(mass <= counts + 1e-9)
    dropped_mass = counts.sum() - mass.sum()
    assert dropped_mass >= st["lost"]
    assert st["divisions"] > 2e7
