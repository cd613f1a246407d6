"""The command line against the reference's own argument parser (CPU).

oracle/_ref/cmdargs_ref is src/io/cmdargs.cpp:11-375 of cuda-pro-cell,
compiled from the reference's sources by oracle/build.py with a main() that
prints the parsed fields when the parser returns.  These tests

  * replay tests/golden/cli_messages.json against it, so the fixture is pinned
    by the reference itself and not only by reading it,
  * run the same argument lines through this build's `procell` command and
    assert that the two agree except for the documented decisions:
      - -p/--phi-min is optional here (the reference lists it among the
        missing arguments, cmdargs.cpp:48-49,69-72),
      - --output and --phi are accepted as aliases of --output-histogram and
        --phi-min, and --seed/-s and --device are additional options (the
        reference answers "Invalid option", cmdargs.cpp:40-44),
      - --help prints the reference's line followed by a usage summary
        (cmdargs.cpp:33-38).

Without the reference's sources (the binary cannot be built) they skip.
"""
import json
import os
import subprocess

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MISSING = "The following missing arguments are required:\n"
PHI_MISSING = "--phi-min (-p)\n"
HELP_LINE = "Questo è l'help"
# Argument lines name the files "x" (a histogram) and "y" (cell types whose
# proportions do not sum to 1): an accepted line then stops at the proportion
# check, before any device is touched, on every machine.
BAD_TYPES = "ERROR: proportion distribution of cell types does not sum to 1, aborting.\n"


@pytest.fixture(scope="module")
def ref_cmdargs():
    from oracle import build as ob

    path = ob.REF_CMDARGS
    if not os.path.isfile(path):
        try:
            path = ob.build_ref()
        except (OSError, subprocess.CalledProcessError) as e:
            pytest.skip(f"reference argument parser does not build: {e}")
    if not path:
        pytest.skip("reference sources absent: oracle/_ref/cmdargs_ref cannot be built")
    return path


@pytest.fixture
def workdir(tmp_path):
    with open(os.path.join(GOLDEN, "readme_histogram.txt")) as f:
        (tmp_path / "x").write_text(f.read())
    (tmp_path / "y").write_text("0.5 10 2\n0.4 20 2\n")
    return tmp_path


def _run(path, args, cwd):
    r = subprocess.run([path] + args, capture_output=True, text=True, timeout=60, cwd=cwd)
    return r.returncode, r.stdout


def _ours_expected(ref):
    """What this build's command prints for an argument line, from the reference's answer."""
    status, out = ref
    if out.startswith("parsed "):
        return "accepted"
    if out.startswith(MISSING):
        mine = out.replace(PHI_MISSING, "")
        return "accepted" if mine == MISSING else (status, mine)
    if out.startswith(HELP_LINE):
        return "help"
    return ref


def _check_pair(ref_path, cli_path, args, cwd):
    ref = _run(ref_path, args, cwd)
    ours = _run(cli_path, args, cwd)
    exp = _ours_expected(ref)
    if exp == "accepted":
        # The parse succeeded; the command stops at the cell-type file.
        assert ours == (1, BAD_TYPES), (args, ref, ours)
    elif exp == "help":
        assert ref == (0, HELP_LINE + "\n"), args
        assert ours[0] == 0 and ours[1].splitlines()[0] == HELP_LINE, (args, ours)
        assert len(ours[1].splitlines()) > 1, "the usage summary follows the reference's line"
    else:
        assert ours == exp, (args, ref, ours)
    return ref, ours


def test_golden_messages_match_the_reference_parser(ref_cmdargs, tmp_path):
    with open(os.path.join(GOLDEN, "cli_messages.json")) as f:
        cases = json.load(f)
    for c in cases:
        got = _run(ref_cmdargs, c["args"], tmp_path)
        assert got == (c["status"], c.get("reference_stdout", c["stdout"])), c["args"]
        if "reference_stdout" in c:
            # The only documented difference in these cases: -p is not required here.
            assert c["reference_stdout"].replace(PHI_MISSING, "") == c["stdout"]


EXTRA_LINES = [
    ["-h", "--help"],                      # -h takes the next word as its file name
    ["-h", "-c", "b"],                     # ... so "b" is an invalid option
    ["--help", "--bogus"],
    ["--bogus", "--help"],
    ["-h", "a", "--help"],
    [""],
    ["-r", "-r"],
    ["-h", "a", "-c"],
    ["-t", "3", "-t"],
    ["-t", "nan"],
    ["-t", "-0.5"],
    ["-p", "-1"],
    ["-p", "1e-400"],                      # atof underflows to 0
    ["-d", "abc"],
    ["-d", "-5"],
    ["-d", "0"],
    ["-h", "x", "-c", "y", "-t", "1"],     # only -p missing: accepted here
    ["-h", "x", "-c", "y", "-t", "abc", "-p", "1"],   # atof("abc") = 0 >= 0
    ["-h", "x", "-c", "y", "-t", "1e400", "-p", "2"],
    ["-h", "x", "-c", "y", "-t", "168", "-p", "0.5", "-d", "7x", "-r"],
    ["-p", "1", "-h", "x", "-c", "y", "-t", "2", "-o", "out", "-r", "-d", "23"],
    ["--histogram", "x", "--cell-types", "y", "--time-max", "5", "--phi-min", "3", "--output-histogram", "o",
     "--tree-depth", "1", "--track-ratio"],
]


def test_argument_lines_agree_with_the_reference_parser(ref_cmdargs, procell_lib, workdir):
    for args in EXTRA_LINES:
        _check_pair(ref_cmdargs, procell_lib.CLI_PATH, args, workdir)


def test_reference_parsed_values(ref_cmdargs, tmp_path):
    """The harness reports what the reference parsed: atof/sscanf semantics and the -d default."""
    status, out = _run(ref_cmdargs, ["-h", "x", "-c", "y", "-t", "abc", "-p", "1"], tmp_path)
    assert status == 0
    assert out == "parsed h0=x cell_types=y output=- t_max=0 phi_min=1 tree_depth=23 track_ratio=0\n"
    status, out = _run(ref_cmdargs, ["-h", "x", "-c", "y", "-t", "168", "-p", "0.5", "-d", "7x", "-r",
                                     "-o", "z"], tmp_path)
    assert status == 0
    assert out == "parsed h0=x cell_types=y output=z t_max=168 phi_min=0.5 tree_depth=7 track_ratio=1\n"


@pytest.mark.parametrize("args", [["--output", "o"], ["--phi", "1"], ["--seed", "1"], ["-s", "1"],
                                  ["--device", "0"]])
def test_documented_extra_spellings(ref_cmdargs, procell_lib, workdir, args):
    """Spellings this build adds; the reference rejects each as an invalid option."""
    assert _run(ref_cmdargs, args, workdir) == (1, f"Invalid option {args[0]}\n")
    status, out = _run(procell_lib.CLI_PATH, args, workdir)
    assert status == 1 and out.startswith(MISSING) and PHI_MISSING not in out
    full = ["-h", "x", "-c", "y", "-t", "1"] + args
    assert _run(procell_lib.CLI_PATH, full, workdir) == (1, BAD_TYPES)
