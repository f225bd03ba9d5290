"""The haplotype-clustering constructor's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpp/fast_build_fuzz.cpp): random tables with missing data,
window lengths, index rebuild regimes, tree splits and thread counts; carriers exact."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "grgl_amd", "csrc")


def test_fast_build_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "fast_build_fuzz"
    srcs = [os.path.join(HERE, "cpp", "fast_build_fuzz.cpp")] + [
        os.path.join(CSRC, f) for f in ("fast_build.cpp", "construct.cpp", "grg_io.cpp", "host_graph.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-pthread", f"-I{CSRC}", *srcs, "-o", str(exe)], check=True)
    env = dict(os.environ, UBSAN_OPTIONS="halt_on_error=1")
    proc = subprocess.run([str(exe), "150"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600,
                          env=env)
    assert proc.returncode == 0, proc.stdout[-4000:]
    assert "fast build fuzz ok" in proc.stdout
