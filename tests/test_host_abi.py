"""CPU checks of libprocell.so and the procell CLI (no GPU needed).

  * the library loads and exports every function include/procell.h declares,
  * host-side entry points (sizing, phi, validation codes, formats) agree
    with the oracle and with the hand-written fixtures of tests/golden,
  * the CLI reproduces the reference's argument messages and exit statuses
    (src/io/cmdargs.cpp:11-375) and fails loudly without a GPU,
  * the product path has no fallback: a missing library or device raises.
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _header_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(procell_[a-z_]+)\s*\(", text))


def _has_gpu():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


def test_library_exports_every_declared_function(procell_lib):
    declared = _header_functions(procell_lib.HEADER_PATH)
    assert len(declared) == 40
    L = procell_lib.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert set(procell_lib.EXPORTS) == declared
    out = subprocess.run(["nm", "-D", "--defined-only", procell_lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (procell_[a-z_]+)$", out, flags=re.M))
    assert declared <= exported


def test_abi_version_and_strerror(procell_lib):
    text = open(procell_lib.HEADER_PATH).read()
    ver = int(re.search(r"#define PROCELL_ABI_VERSION (\d+)", text).group(1))
    assert procell_lib.lib().procell_abi_version() == ver
    codes = {int(v): k for k, v in re.findall(r"#define (PROCELL_E_[A-Z_]+) \((-\d+)\)", text)}
    assert sorted(codes) == list(range(-13, 0))
    seen = set()
    for c in codes:
        msg = procell_lib.strerror(c)
        assert msg and msg != "unknown error"
        seen.add(msg)
    assert len(seen) >= 12
    assert procell_lib.strerror(0) == "success"
    assert procell_lib.strerror(-99) == "unknown error"
    assert procell_lib.strerror(-2) == "proportion distribution of cell types does not sum to 1"


def _readme_problem(procell_lib, **kw):
    v, c = procell_lib.read_histogram(os.path.join(GOLDEN, "readme_histogram.txt"))
    p, m, s = procell_lib.read_types(os.path.join(GOLDEN, "readme_types.txt"))
    return procell_lib.Problem(v, c, p, m, s, kw.pop("t_max", 168.0), **kw)


def test_read_readme_fixtures(procell_lib):
    v, c = procell_lib.read_histogram(os.path.join(GOLDEN, "readme_histogram.txt"))
    assert v.tolist() == [1.0, 8.144, 9823.85] and c.tolist() == [0, 53, 274]
    p, m, s = procell_lib.read_types(os.path.join(GOLDEN, "readme_types.txt"))
    assert p.tolist() == [0.53, 0.29, 0.18]
    assert m.tolist() == [48.33, 86.3, -1.0] and s.tolist() == [21.6, 26.8, -1.0]


def test_output_bin_set_of_readme(procell_lib):
    with open(os.path.join(GOLDEN, "readme_bins.json")) as f:
        g = json.load(f)
    prob = _readme_problem(procell_lib)
    assert prob.resolve_phi() == g["phi"]
    assert prob.output_capacity() == len(g["bins"])
    rows, kdim, T = prob.tally_shape()
    # 9823.85 halves 10 times while staying above 8.144 (the last one is 9.59)
    assert (rows, kdim, T) == (2, 11, 3)


def test_duplicate_bins_merge(procell_lib):
    prob = procell_lib.Problem([8.0, 16.0, 3.0], [1, 1, 1], [1.0], [10.0], [1.0], 10.0, phi=1.0)
    # {16, 8, 4, 2, 1} u {8, 4, 2, 1} u {3, 1.5}
    assert prob.output_capacity() == 7
    prob = procell_lib.Problem([8.0, 16.0, 3.0], [1, 0, 1], [1.0], [10.0], [1.0], 10.0, phi=1.0)
    assert prob.output_capacity() == 6


def _oracle_code(oracle, *args):
    return oracle.max_generation(oracle.Problem(*args))


VALIDATION_CASES = [
    # values, counts, proportion, mean, sd, t_max, phi, expected code
    ([1.0, 2.0], [1, 1], [0.5, 0.49], [1.0, 1.0], [1.0, 1.0], 5.0, 1.0, -2),
    ([1.0, 2.0], [1, 1], [0.5, 0.5 - 2e-8], [1.0, 1.0], [1.0, 1.0], 5.0, 1.0, -2),
    ([1.0, 2.0], [1, 1], [0.5, 0.5 - 5e-9], [1.0, 1.0], [1.0, 1.0], 5.0, 1.0, 1),
    ([1.0, 2.0], [1, 1], [0.5, 0.5], [1.0, 0.0], [1.0, 1.0], 5.0, 1.0, -3),
    ([1.0, float("nan")], [1, 1], [1.0], [1.0], [1.0], 5.0, 1.0, -4),
    ([1.0, 2.0], [1, 1], [1.0], [float("inf")], [1.0], 5.0, 1.0, -4),
    ([1.0, 2.0], [1, 1], [1.0], [1.0], [1.0], float("nan"), 1.0, -1),
    ([1.0, 2.0], [1, 1], [1.0], [1.0], [1.0], -1.0, 1.0, -1),
    ([1.0, 2.0], [1, 1], [1.0], [1.0], [1.0], float("-inf"), 1.0, -1),
    ([1.0, 2.0], [1, 1], [1.0], [1.0], [1.0], float("inf"), 1.0, 1),
    ([1.0, 2.0], [1, 1], [], [], [], 5.0, 1.0, -2),
    ([1.0, 2.0], [0, 0], [], [], [], 5.0, 1.0, 0),
    ([-1.0, 2.0], [1, 1], [1.0], [1.0], [1.0], 5.0, 0.0, -5),
    ([1e30, 2.0], [1, 1], [1.0], [1.0], [1.0], 5.0, 1e-3, -6),
    ([1.0, 2.0], [2**32, 1], [1.0], [1.0], [1.0], 5.0, 1.0, -7),
    ([1.0, 2.0], [0, 0], [1.0], [1.0], [1.0], 5.0, 0.0, 0),
]


@pytest.mark.parametrize("case", range(len(VALIDATION_CASES)))
def test_validation_codes_match_oracle(procell_lib, oracle, case):
    *args, expect = VALIDATION_CASES[case]
    values, counts, prop, mean, sd, t_max, phi = args
    prob = procell_lib.Problem(values, counts, prop, mean, sd, t_max, phi=phi)
    rc = procell_lib.lib().procell_output_capacity(ctypes.byref(prob.params), ctypes.byref(ctypes.c_uint64()))
    orc = _oracle_code(oracle, values, counts, prop, mean, sd, t_max, phi, 0)
    if expect == -7:
        # the product packs a cell id into 32 bits; the oracle takes up to 2^48
        assert rc == expect and orc >= 0
    elif expect < 0:
        assert rc == expect and orc == expect
    else:
        assert rc == 0 and orc >= 0


def test_shapes_and_phi_match_oracle(procell_lib, oracle):
    from procell import workloads

    for w in (workloads.small_histogram(0), workloads.large_histogram(0), workloads.small_histogram(3)):
        prob = procell_lib.Problem.from_workload(w)
        op = oracle.Problem.from_workload(w, 0)
        assert prob.resolve_phi() == oracle.resolve_phi(op)
        rows, kdim, T = prob.tally_shape()
        assert rows == int(np.count_nonzero(w.counts))
        assert kdim == oracle.max_generation(op) + 1
        assert T == len(w.proportion)


def test_write_histogram_format(procell_lib, tmp_path):
    with open(os.path.join(GOLDEN, "format_cases.json")) as f:
        g = json.load(f)
    rows = np.array(g["rows"], dtype=object)
    value = np.array([float(r[0]) for r in rows])
    count = np.array([int(r[1]) for r in rows], dtype=np.int64)
    tc = np.array(g["type_counts"], dtype=np.int64)
    path = str(tmp_path / "out.txt")
    procell_lib.write_histogram(path, value, count)
    assert open(path).read() == g["text"]
    procell_lib.write_histogram(path, value, count, tc, track_ratio=True)
    assert open(path).read() == g["text_ratio"]
    # rows with a zero count are not printed
    count0 = count.copy()
    count0[1] = 0
    procell_lib.write_histogram(path, value, count0)
    assert open(path).read() == "".join(l + "\n" for i, l in enumerate(g["text"].splitlines()) if i != 1)
    # halving prints as in the reference's README example
    procell_lib.write_histogram(path, [9823.85 / 2, 9823.85], [1, 2])
    assert open(path).read() == "4911.925\t1\n9823.85\t2\n"


def test_oracle_files_agree_with_readme_bins(oracle, tmp_path):
    with open(os.path.join(GOLDEN, "readme_bins.json")) as f:
        bins = set(json.load(f)["bins"])
    out = str(tmp_path / "o.txt")
    rc = oracle.run_files(os.path.join(GOLDEN, "readme_histogram.txt"), os.path.join(GOLDEN, "readme_types.txt"),
                          1000.0, 0.0, True, 5, out, 2)
    assert rc == 0
    lines = open(out).read().splitlines()
    vals = [float(l.split("\t")[0]) for l in lines]
    assert vals == sorted(vals) and len(lines) >= 6
    for l in lines:
        parts = l.split("\t")
        assert len(parts) == 5
        assert int(parts[1]) == sum(int(x) for x in parts[2:])
        assert any(float(parts[0]) == b or parts[0] == "%.10g" % b for b in bins)


def _cli(procell_lib, args, cwd=None):
    return subprocess.run([procell_lib.CLI_PATH] + args, capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_messages_and_status(procell_lib):
    with open(os.path.join(GOLDEN, "cli_messages.json")) as f:
        cases = json.load(f)
    for c in cases:
        r = _cli(procell_lib, c["args"])
        assert (r.returncode, r.stdout) == (c["status"], c["stdout"]), c["args"]


def test_cli_help(procell_lib):
    r = _cli(procell_lib, ["-t", "3", "--help"])
    assert r.returncode == 0
    assert r.stdout.splitlines()[0] == "Questo è l'help"


def test_cli_rejects_bad_proportions(procell_lib, tmp_path):
    types = tmp_path / "t.txt"
    types.write_text("0.5 10 2\n0.4 20 2\n")
    r = _cli(procell_lib, ["-h", os.path.join(GOLDEN, "readme_histogram.txt"), "-c", str(types), "-t", "10"])
    assert r.returncode == 1
    assert r.stdout == "ERROR: proportion distribution of cell types does not sum to 1, aborting.\n"


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a device")
def test_cli_fails_loudly_without_device(procell_lib):
    r = _cli(procell_lib, ["-h", os.path.join(GOLDEN, "readme_histogram.txt"), "-c",
                           os.path.join(GOLDEN, "readme_types.txt"), "--phi", "8.144", "-t", "168", "-r"])
    assert r.returncode == 1
    assert r.stdout == "ERROR: no usable gfx950 device, aborting.\n"


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a device")
def test_simulate_fails_loudly_without_device(procell_lib):
    prob = _readme_problem(procell_lib)
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.simulate(prob)
    assert e.value.code == procell_lib.E_NO_DEVICE
    with pytest.raises(procell_lib.ProcellError):
        procell_lib.Plan(prob)


def test_batch_validation(procell_lib):
    """Batch creation validates every set before it touches a device."""
    good = _readme_problem(procell_lib)
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.Batch([])
    assert e.value.code == procell_lib.E_INVALID
    bad = procell_lib.Problem([8.0], [3], [0.5, 0.4], [10.0, 20.0], [2.0, 2.0], 10.0, 1.0)
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.Batch([good, bad])
    assert e.value.code == procell_lib.E_PROPORTIONS
    zero_mean = procell_lib.Problem([8.0], [3], [1.0], [0.0], [2.0], 10.0, 1.0)
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.Batch([zero_mean, good])
    assert e.value.code == procell_lib.E_ZERO_MEAN


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a device")
def test_batch_fails_loudly_without_device(procell_lib):
    prob = _readme_problem(procell_lib)
    with pytest.raises(procell_lib.ProcellError) as e:
        procell_lib.Batch([prob, prob])
    assert e.value.code == procell_lib.E_NO_DEVICE


def test_relaxed_validation_and_no_device(procell_lib):
    import ctypes

    prob = _readme_problem(procell_lib)
    h = ctypes.c_void_p()
    L = procell_lib.lib()
    for beta in (0.0, -1.0, float("nan"), float("inf")):
        assert L.procell_relaxed_create(ctypes.byref(prob.params), beta, ctypes.byref(h)) == procell_lib.E_INVALID
    assert L.procell_relaxed_create(None, 0.5, ctypes.byref(h)) == procell_lib.E_INVALID
    for null_call in (L.procell_relaxed_forward(None, None, None, None),
                      L.procell_relaxed_backward(None, None, None, None, None),
                      L.procell_relaxed_shape(None, None, None, None),
                      L.procell_relaxed_jacobian(None, None, None, None, None),
                      L.procell_relaxed_set_shape(None, 0, None, None, None, None)):
        assert null_call == procell_lib.E_INVALID
    # batch plans: no sets, no array, bad beta
    assert L.procell_relaxed_create_batch(None, 1, 0.5, ctypes.byref(h)) == procell_lib.E_INVALID
    assert L.procell_relaxed_create_batch(ctypes.byref(prob.params), 0, 0.5, ctypes.byref(h)) == procell_lib.E_INVALID
    for beta in (0.0, float("nan")):
        assert L.procell_relaxed_create_batch(ctypes.byref(prob.params), 1, beta, ctypes.byref(h)) == \
            procell_lib.E_INVALID
    if not _has_gpu():
        with pytest.raises(procell_lib.ProcellError) as e:
            procell_lib.Relaxed(prob, 0.5)
        assert e.value.code == procell_lib.E_NO_DEVICE
        with pytest.raises(procell_lib.ProcellError) as e:
            procell_lib.RelaxedBatch([prob, prob], 0.5)
        assert e.value.code == procell_lib.E_NO_DEVICE


def test_missing_library_raises(procell_lib, monkeypatch):
    monkeypatch.setattr(procell_lib, "_lib", None)
    monkeypatch.setattr(procell_lib, "LIB_PATH", "/nonexistent/libprocell.so")
    with pytest.raises(RuntimeError, match="missing"):
        procell_lib.lib()
