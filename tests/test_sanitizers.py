"""The oracle and the host library built with AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/Makefile, target asan), and the CPU tests
that exercise them run against those builds in a child process: an
out-of-bounds access, a use after free or undefined behaviour in either
library fails the child."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _runtime(name):
    p = subprocess.run(["g++", "-print-file-name=" + name], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.slow
def test_cpu_tests_under_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("sanitizer runtimes not installed")
    subprocess.check_call(["make", "-s", "-C", ORACLE, "asan"])
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + ":" + ubsan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["SELM_ORACLE_LIB"] = os.path.join(ORACLE, "_asan", "liboracle.so")
    env["SELM_HOST_LIB"] = os.path.join(ORACLE, "_asan", "libselm_host.so")
    env["OMP_NUM_THREADS"] = "4"
    tests = [os.path.join(ROOT, "tests", t) for t in ("test_oracle.py", "test_host.py", "test_msd.py")]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + tests,
                       env=env, capture_output=True, text=True, timeout=900)
    print(p.stdout[-2000:], p.stderr[-3000:])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr
