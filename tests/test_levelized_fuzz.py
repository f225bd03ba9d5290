"""Levelized-graph validation under damage, in C++ with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpp/levelized_fuzz.cpp): random graphs round-trip through
files, plans with changed entries or sizes are rejected or harmless and never read out of
bounds while being checked, and damaged or truncated files are refused."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "grgl_amd", "csrc")


def test_levelized_validation_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "lvlz_fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-pthread", f"-I{CSRC}", os.path.join(HERE, "cpp", "levelized_fuzz.cpp"),
                    os.path.join(CSRC, "levelize.cpp"), os.path.join(CSRC, "host_graph.cpp"), "-o", str(exe)],
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1")
    proc = subprocess.run([str(exe), str(tmp_path), "60", "200"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=600, env=env)
    assert proc.returncode == 0, proc.stdout[-4000:]
    assert "levelized fuzz ok" in proc.stdout
