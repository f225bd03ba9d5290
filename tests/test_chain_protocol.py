"""The chained traversal's claim / wait protocol, emulated on the CPU with threads standing for
thread blocks (tests/cpp/chain_protocol_test.cpp): results equal a sequential level-by-level
evaluation for any number of blocks, waits are bounded, the abort word ends every block and is
counted once, and the state words are left zero.  Built once plain and once with
ThreadSanitizer, which checks that every row read is ordered after its write by the counters'
release / acquire pairs."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "chain_protocol_test.cpp")


@pytest.mark.parametrize("sanitize", [False, True])
def test_chain_protocol(tmp_path, sanitize):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / ("chain_tsan" if sanitize else "chain")
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", SRC, "-o", str(exe)]
    if sanitize:
        cmd[1:1] = ["-fsanitize=thread", "-g"]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    proc = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600, env=env)
    assert proc.returncode == 0, proc.stdout[-4000:]
    assert "chain protocol ok" in proc.stdout
