"""A graph of more than 2^32 gather entries: 2^20 samples and 4,200 internal nodes that each
have every sample as a child (4.4e9 edges, 17.6 GB of child indices).  List offsets wrap 32
bits and the largest levels run as several launches.  Results have closed forms: UP of ones
is the sample count at every internal node; DOWN of ones is one per mutation above a sample.

It needs about 90 GB of host memory (86-87 GB peak measured) and 40 GB of device memory and
takes about a minute on an MI355X box.  It is skipped on hosts with less than 120 GB of
memory, or when GRGL_HIP_SKIP_LARGE_TESTS=1."""

import gc
import os
import resource
import sys

import numpy as np
import pytest

import grgl_amd

def _host_memory_bytes():
    try:
        return os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
    except (ValueError, OSError, AttributeError):
        return 0


pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("GRGL_HIP_SKIP_LARGE_TESTS") == "1",
                                 reason="GRGL_HIP_SKIP_LARGE_TESTS=1"),
              pytest.mark.skipif(_host_memory_bytes() < 120e9, reason="needs ~90 GB of host memory")]

UP, DOWN = grgl_amd.UP, grgl_amd.DOWN
N = 1 << 20
P = 4200
EXTRA = 1000  # mutations on samples


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_more_than_2_32_gather_entries(gpu):
    assert N * P > 1 << 32
    _say(f"building arrays: {N} samples, {P} internal nodes, {N * P} edges")
    off = np.zeros(N + P + 1, np.uint64)
    off[N + 1:] = np.arange(1, P + 1, dtype=np.uint64) * N
    children = np.tile(np.arange(N, dtype=np.uint32), P)
    rng = np.random.default_rng(1)
    sample_muts = rng.integers(0, N, size=EXTRA).astype(np.uint32)
    muts = np.concatenate([np.arange(N, N + P, dtype=np.uint32), sample_muts])
    g = grgl_amd.grg_from_arrays(N, 2, off, children, muts)
    del children
    gc.collect()
    M = P + EXTRA
    assert g.num_edges == N * P and g.num_mutations == M

    _say("UP, int32 ones (device graph build included)")
    up = grgl_amd.matmul(g, np.ones((1, N), np.int32), UP)
    assert (up[0, :P] == N).all() and (up[0, P:] == 1).all()

    _say("DOWN, int32 ones")
    down = grgl_amd.matmul(g, np.ones((1, M), np.int32), DOWN)
    want = np.full(N, P, np.int64) + np.bincount(sample_muts, minlength=N)
    assert (down[0] == want).all()

    _say("UP and DOWN, float32")
    x = rng.standard_normal((2, N)).astype(np.float32)
    upf = grgl_amd.matmul(g, x, UP)
    ref = x.astype(np.float64).sum(axis=1)
    for k in range(2):
        # every internal node sums the same list: identical bits, close to the FP64 sum
        assert (upf[k, :P] == upf[k, 0]).all()
        assert abs(upf[k, 0] - ref[k]) <= 1e-6 * np.abs(x[k]).sum()
        assert (upf[k, P:] == x[k, sample_muts]).all()
    w = rng.standard_normal((2, M)).astype(np.float32)
    downf = grgl_amd.matmul(g, w, DOWN)
    base = w[:, :P].astype(np.float64).sum(axis=1)
    extra = np.zeros((2, N))
    for k in range(2):
        np.add.at(extra[k], sample_muts, w[k, P:].astype(np.float64))
    wantf = base[:, None] + extra
    scale = np.abs(w).sum(axis=1)[:, None]
    assert (np.abs(downf - wantf) <= 1e-6 * scale).all()
    _timing(g, N, M)
    _say(f"done; peak host memory {resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6:.1f} GB")


def _timing(g, N, M):
    """Device-resident traversals on the same graph (reported, not asserted): at this size a
    traversal streams 4 bytes of index per edge from HBM."""
    import time

    import torch

    from grgl_amd import capi

    L = capi.lib()
    handle = g.device_handle()
    stream = torch.cuda.current_stream().cuda_stream
    E = g.num_edges
    for direction, name, cin, cout in ((UP, "UP", N, M), (DOWN, "DOWN", M, N)):
        for K in (1, 8):
            x = torch.randn((K, cin), device="cuda")
            y = torch.zeros((K, cout), device="cuda")

            def call():
                capi.check(L.grgl_hip_matmul(handle, capi.F32, x.data_ptr(), cin, K, int(direction), y.data_ptr(),
                                             y.numel(), 0, 0, None, capi.INIT_ZERO, None, 0, stream))

            call()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(3):
                call()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) / 3 * 1e3
            entries = E + (M if direction == DOWN else 0)
            _say(f"timing {name} K={K}: {ms:.2f} ms, {E * K / ms / 1e9:.3f}e12 edge*RHS/s, "
                 f"index stream {4 * entries / ms / 1e9:.2f} TB/s")
