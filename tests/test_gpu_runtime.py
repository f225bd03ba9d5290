"""One HIP runtime per process: `import grgl_amd` does not import torch, the numpy matmul path
runs on the GPU without it, and torch imported afterwards shares the runtime, so its tensors'
device pointers go straight into the C ABI.  With GRGL_AMD_NO_TORCH=1 the numpy path uses the
system runtime and works the same.  Each case runs in a fresh process (one GPU process at a
time)."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

NUMPY_PATH = """
import sys
import numpy as np
import grgl_amd
import oracle
assert 'torch' not in sys.modules
g = grgl_amd.grg_from_coalescent(1000, 10000, 1, seed=5)
og = oracle.OracleGraph(g.arrays())
x = np.ones((2, g.num_samples), np.int32)
assert np.array_equal(grgl_amd.matmul(g, x, grgl_amd.UP), og.matmul(x, oracle.UP))
assert 'torch' not in sys.modules
"""

TORCH_AFTER = """
import torch
from grgl_amd import capi
cg = capi.Graph(g.arrays())
xd = torch.from_numpy(x).cuda()
yd = torch.zeros((2, cg.num_mutations), dtype=torch.int32, device="cuda")
cg.matmul_ptr(capi.I32, xd.data_ptr(), cg.num_samples, 2, capi.UP, yd.data_ptr(), yd.numel(),
              stream=torch.cuda.current_stream().cuda_stream)
torch.cuda.synchronize()
assert np.array_equal(yd.cpu().numpy(), og.matmul(x, oracle.UP))
"""


@pytest.mark.parametrize("case", ["torch_after", "no_torch"])
def test_runtime_sharing(gpu, case):
    code = NUMPY_PATH + (TORCH_AFTER if case == "torch_after" else "") + "print('ok')\n"
    env = dict(os.environ)
    if case == "no_torch":
        env["GRGL_AMD_NO_TORCH"] = "1"
    proc = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=150)
    assert proc.returncode == 0 and proc.stdout.strip().endswith("ok"), proc.stdout[-3000:]
