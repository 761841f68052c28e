"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(`make -C oracle asan`) runs a driver that calls every entry point of
oracle/ppvm_oracle.h on random programs (every op code, thresholds, weight
caps, preserve sets, clones, overlap, trace, patterns, export, fingerprint,
marked and sharded runs, lossy sums).  Any invalid access or undefined
operation aborts the driver."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.timeout(600)
def test_oracle_under_address_and_ub_sanitizers():
    if shutil.which(os.environ.get("CXX", "g++")) is None:
        pytest.skip("no host C++ compiler")
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "_asan", "asan_check")], capture_output=True, text=True,
                       env=env, timeout=540)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "asan ok" in r.stdout
