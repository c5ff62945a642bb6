"""Sanitizer builds of the host code (CPU).

The engine's host utilities (ymc3d_amd/csrc/host_util.cpp: multiplier search, safe radius and safe
box on std::thread pools, seeding, chains, voxeliser, file I/O) are built with g++ under
AddressSanitizer + UndefinedBehaviorSanitizer and, separately, ThreadSanitizer, and run through
tests/sanitize/host_driver.cpp; the CPU oracle (oracle/ymc_oracle.cpp, OpenMP over chains) is built
under AddressSanitizer + UndefinedBehaviorSanitizer and run through tests/sanitize/oracle_driver.cpp
(every tally, records and the absorption grid, two tables, both illuminations).  Any sanitizer
report fails the test (reports abort with a non-zero status).  The oracle is not run under
ThreadSanitizer: libgomp's barriers are not instrumented, so every read after a parallel region of
a value written atomically inside it is reported as a race.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
CSRC = os.path.join(ROOT, "ymc3d_amd", "csrc")
GXX = shutil.which("g++")

pytestmark = pytest.mark.skipif(GXX is None, reason="g++ not available")


def _build_and_run(tmp_path, name, sources, flags, args=()):
    exe = str(tmp_path / name)
    cmd = [GXX, "-O1", "-g", "-std=c++17", *flags, "-I" + CSRC, *sources, "-o", exe, "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-5000:]
    for marker in ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "LeakSanitizer"):
        assert marker not in r.stderr, r.stderr[-5000:]
    return r.stdout


def test_host_utilities_under_address_and_undefined_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, "host_asan", [os.path.join(SAN, "host_driver.cpp"), os.path.join(CSRC, "host_util.cpp")],
                         ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"], [str(tmp_path)])
    assert "host driver ok" in out


def test_host_utilities_under_thread_sanitizer(tmp_path):
    out = _build_and_run(tmp_path, "host_tsan", [os.path.join(SAN, "host_driver.cpp"), os.path.join(CSRC, "host_util.cpp")],
                         ["-fsanitize=thread"], [str(tmp_path)])
    assert "host driver ok" in out


def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, "oracle_asan",
                         [os.path.join(SAN, "oracle_driver.cpp"), os.path.join(ROOT, "oracle", "ymc_oracle.cpp")],
                         ["-fopenmp", "-ffp-contract=off", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"])
    assert "oracle driver ok" in out
