"""Grouped-level plans under random graphs and limits, in C++ with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpp/grouping_fuzz.cpp): every plan validates, fits its LDS
limit, and executed the grouped kernel's way gives the level plan's float bits."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "grgl_amd", "csrc")


def test_grouping_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "grouping_fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-pthread", f"-I{CSRC}", os.path.join(HERE, "cpp", "grouping_fuzz.cpp"),
                    os.path.join(CSRC, "grouping.cpp"), os.path.join(CSRC, "levelize.cpp"),
                    os.path.join(CSRC, "host_graph.cpp"), "-o", str(exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1")
    proc = subprocess.run([str(exe), "120"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600,
                          env=env)
    assert proc.returncode == 0, proc.stdout[-4000:]
    assert "grouping fuzz ok" in proc.stdout
