"""Concurrent matmul calls from Python threads.  The binding releases the GIL for the whole
call, as the reference's does (src/python/_grgl.cpp:169-183), so calls overlap on the device.

* test_parallel_split_down restates the reference's test_parallel
  (test/endtoend/test_matmul.py:264-319): DOWN on the pieces of a graph split by genome range,
  run on a thread pool and summed, equals DOWN on the whole graph.  The pieces here keep the
  whole node set and only their range's mutations (the reference's `grg split` also prunes
  nodes; pruning unreachable nodes does not change a DOWN result).
* test_threads_same_graph: many threads on ONE graph, mixed directions, dtypes and K; every
  result equals the same call made alone (bit for bit: the per-column arithmetic does not
  depend on what else runs).
"""

import concurrent.futures as cf

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import bits_equal, make_input

pytestmark = pytest.mark.gpu

UP, DOWN = grgl_amd.UP, grgl_amd.DOWN
_G = {}


def _graph():
    if not _G:
        g = grgl_amd.grg_from_coalescent(3000, 30000, 6, seed=17)
        _G["g"] = g
        _G["arrays"] = g.arrays()
    return _G["g"], _G["arrays"]


def _pieces(arrays, n):
    """Graphs over the same nodes holding consecutive ranges of mutation IDs (mutation IDs
    follow genome position), like the reference's split by base-pair range."""
    M = len(arrays["mutation_nodes"])
    bounds = np.linspace(0, M, n + 1).astype(np.int64)
    out = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        out.append((int(a), int(b), grgl_amd.grg_from_arrays(arrays["num_samples"], arrays["ploidy"],
                                                             arrays["down_offsets"], arrays["down_children"],
                                                             np.ascontiguousarray(arrays["mutation_nodes"][a:b]))))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.int32])
def test_parallel_split_down(gpu, dtype):
    g, arrays = _graph()
    rng = np.random.default_rng(3)
    K = 10
    w = make_input(rng, dtype, (K, g.num_mutations))
    full = grgl_amd.matmul(g, w, DOWN)
    pieces = _pieces(arrays, 7)
    assert sum(b - a for a, b, _ in pieces) == g.num_mutations
    with cf.ThreadPoolExecutor(8) as ex:
        futs = [ex.submit(grgl_amd.matmul, pg, np.ascontiguousarray(w[:, a:b]), DOWN) for a, b, pg in pieces]
        parts = [f.result() for f in futs]
    total = parts[0].copy()
    for p in parts[1:]:
        total += p
    if np.dtype(dtype).kind == "i":
        assert bits_equal(total, full)
        assert bits_equal(full, oracle.OracleGraph(arrays).matmul(w, oracle.DOWN))
    else:
        np.testing.assert_allclose(total, full, rtol=1e-12, atol=1e-12 * np.abs(full).max())


def test_threads_same_graph(gpu):
    g, arrays = _graph()
    rng = np.random.default_rng(5)
    calls = []
    for i in range(48):
        d = UP if i % 2 == 0 else DOWN
        dt = [np.float32, np.int32, np.float64, np.int8][i % 4]
        K = [1, 3, 8, 17, 64, 130][i % 6]
        cols = g.num_samples if d == UP else g.num_mutations
        calls.append((make_input(rng, dt, (K, cols)), d))
    alone = [grgl_amd.matmul(g, x, d) for x, d in calls]
    with cf.ThreadPoolExecutor(12) as ex:
        futs = [ex.submit(grgl_amd.matmul, g, x, d) for x, d in calls]
        together = [f.result() for f in futs]
    for i, (a, b) in enumerate(zip(alone, together)):
        assert bits_equal(a, b), i
    og = oracle.OracleGraph(arrays)
    for (x, d), out in zip(calls, together):
        if x.dtype.kind == "i":
            assert bits_equal(out, og.matmul(x, oracle.UP if d == UP else oracle.DOWN))


def test_threads_device_pointers_on_own_streams(gpu):
    """Device-pointer calls on one graph from several threads, each on its own HIP stream
    (the library keeps a workspace per stream), equal the same calls made one by one."""
    import torch

    from grgl_amd import capi

    _, arrays = _graph()
    cg = capi.Graph(arrays)
    rng = np.random.default_rng(9)
    jobs = []
    for i in range(24):
        direction = capi.UP if i % 2 == 0 else capi.DOWN
        K = [1, 4, 16, 100][i % 4]
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        x = torch.from_numpy(make_input(rng, np.float32, (K, cin))).cuda()
        jobs.append((direction, K, cin, cout, x))
    torch.cuda.synchronize()

    def run(job, stream):
        direction, K, cin, cout, x = job
        with torch.cuda.stream(stream):  # the zero fill is ordered before the traversal
            y = torch.zeros((K, cout), dtype=torch.float32, device="cuda")
        cg.matmul_ptr(capi.F32, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(),
                      stream=stream.cuda_stream)
        stream.synchronize()
        return y

    default = torch.cuda.current_stream()
    alone = [run(j, default).cpu().numpy() for j in jobs]
    streams = [torch.cuda.Stream() for _ in range(6)]
    with cf.ThreadPoolExecutor(6) as ex:
        futs = [ex.submit(run, j, streams[i % 6]) for i, j in enumerate(jobs)]
        together = [f.result().cpu().numpy() for f in futs]
    for i, (a, b) in enumerate(zip(alone, together)):
        assert bits_equal(a, b), i
    cg.close()
