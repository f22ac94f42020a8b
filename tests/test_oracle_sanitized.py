"""The oracle's host code under AddressSanitizer and UndefinedBehaviorSanitizer:
the oracle tests run again in a child process against a sanitized build of
the same sources (oracle/Makefile target `sanitize`). Any out-of-bounds
access, use after free or undefined behaviour ends the child with an error."""
import os
import shutil
import subprocess
import sys

import pytest

from tests.test_abi import ROOT

FILES = ["tests/test_oracle_math.py", "tests/test_oracle_paths.py", "tests/test_oracle_guides.py",
         "tests/test_golden.py"]


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_tests_pass_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    assert os.path.isabs(asan), "gcc has no AddressSanitizer runtime"
    env = dict(os.environ, WR_ORACLE_LIB=str(ROOT / "oracle" / "build" / "liboracle_san.so"), LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *FILES],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
