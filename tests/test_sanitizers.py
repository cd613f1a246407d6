"""Host code and the oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

* libprocell's host-only sources (cuda-pro-cell_amd/csrc/procell_setup.cpp:
  validation, set-up, segment split, merge; procell_formats.cpp: the text
  formats) are built with g++ -fsanitize=address,undefined together with the
  C oracle and tests/sanitize/host_checks.cpp, which drives them against the
  oracle on the golden and generated inputs, every documented error case and
  malformed files.  Leak detection is on.
* The oracle is built the same way as a shared library and the brute-force,
  invariant, closed-form and relaxed-estimator tests of the oracle run against
  it in a child pytest with the ASan runtime preloaded.

Any sanitizer report ends the program with a nonzero status
(-fno-sanitize-recover=all, halt_on_error), which fails the test.
"""
import os
import shutil
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "cuda-pro-cell_amd", "csrc")
ORACLE_C = os.path.join(REPO, "oracle", "procell_oracle.c")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ORACLE_FLAGS = ["-std=c11", "-ffp-contract=off", "-fno-fast-math", "-mfma", "-fopenmp"]
ORACLE_TESTS = ["test_oracle_primitives.py", "test_oracle_simulation.py", "test_oracle_relaxed.py",
                "test_oracle_expected.py"]


def _runtime(name):
    try:
        out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True,
                             timeout=60).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.fixture(scope="module")
def san_dir(tmp_path_factory):
    if not (shutil.which("gcc") and shutil.which("g++") and _runtime("libasan.so")):
        pytest.skip("gcc/g++ with the AddressSanitizer runtime are not available")
    return tmp_path_factory.mktemp("sanitize")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(map(str, cmd))}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r


def test_host_code_under_sanitizers(san_dir):
    obj = san_dir / "oracle.o"
    exe = san_dir / "host_checks"
    _run(["gcc"] + SAN + ORACLE_FLAGS + ["-c", ORACLE_C, "-o", str(obj)], timeout=300)
    _run(["g++"] + SAN + ["-std=c++17", "-D_GLIBCXX_ASSERTIONS", "-Wall", "-Wextra", "-Werror",
                          "-I", os.path.join(REPO, "include"), "-I", CSRC, "-I", os.path.join(REPO, "oracle"),
                          os.path.join(REPO, "tests", "sanitize", "host_checks.cpp"),
                          os.path.join(CSRC, "procell_setup.cpp"), os.path.join(CSRC, "procell_formats.cpp"),
                          str(obj), "-fopenmp", "-lm", "-o", str(exe)], timeout=300)
    work = san_dir / "work"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    r = _run([str(exe), os.path.join(REPO, "tests", "golden"), str(work)], env=env, timeout=600)
    assert " 0 failed" in r.stdout and "runs against the oracle" in r.stdout, r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]


def test_oracle_tests_under_sanitizers(san_dir):
    lib = san_dir / "libprocell_oracle_san.so"
    _run(["gcc"] + SAN + ORACLE_FLAGS + ["-fPIC", "-shared", "-o", str(lib), ORACLE_C, "-lm"], timeout=300)
    env = dict(os.environ, LD_PRELOAD=_runtime("libasan.so"), PROCELL_ORACLE_LIB=str(lib),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               PYTHONPATH=REPO)
    # the child really runs the sanitized build, with the ASan runtime loaded
    probe = ("import ctypes; from oracle import binding as b; b.lib(); "
             "ctypes.CDLL(None).__asan_init; print(b.LIB_PATH)")
    r = _run([sys.executable, "-c", probe], env=env, cwd=REPO, timeout=300)
    assert r.stdout.strip() == str(lib)
    tests = [os.path.join(REPO, "tests", t) for t in ORACLE_TESTS]
    r = _run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + tests,
             env=env, cwd=REPO, timeout=1200)
    assert " passed" in r.stdout and " failed" not in r.stdout, r.stdout[-4000:]
    assert "runtime error" not in r.stderr + r.stdout
