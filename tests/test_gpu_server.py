"""Serving loop on the GPU: continuous batching must not change any result, admission
follows the HBM budget, cancellation and back-pressure behave as documented."""

import numpy as np
import pytest

import oracle
from grgl_amd import capi
from helpers import bits_equal, make_input, random_grg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs(gpu):
    import grgl_amd

    a1 = grgl_amd.grg_from_coalescent(2000, 20000, 4, seed=11).arrays()
    a2 = random_grg(5, num_samples=300, num_internal=3000, heavy=(500, 2000), num_mutations=4000)
    return [(a1, capi.Graph(a1), oracle.OracleGraph(a1)), (a2, capi.Graph(a2), oracle.OracleGraph(a2))]


def _solo(cg, x, direction):
    return cg.matmul_host(x, direction)


def test_mixed_stream_equals_alone(graphs):
    srv = capi.Server(num_streams=2, max_batch_rows=256, batch_window_us=20000)
    ids = [srv.add_graph(cg) for _, cg, _ in graphs]
    rng = np.random.default_rng(0)
    subs = []
    for i in range(60):
        gi = int(rng.integers(0, 2))
        _, cg, og = graphs[gi]
        d = capi.UP if rng.random() < 0.5 else capi.DOWN
        K = int(rng.choice([1, 8, 32, 128], p=[0.4, 0.3, 0.2, 0.1]))
        dt = np.float32 if rng.random() < 0.7 else np.int32
        x = make_input(rng, dt, (K, cg.num_samples if d == capi.UP else cg.num_mutations))
        rid, out = srv.submit(ids[gi], x, d)
        subs.append((rid, out, x, gi, d))
    for rid, out, x, gi, d in subs:
        assert srv.wait(rid) >= 0.0
        _, cg, og = graphs[gi]
        assert bits_equal(out, _solo(cg, x, d)), "batched result differs from the same request alone"
        if x.dtype == np.int32:
            assert bits_equal(out, og.matmul(x, oracle.UP if d == capi.UP else oracle.DOWN))
    st = srv.stats()
    assert st["completed"] == 60 and st["batches"] >= 1 and st["batches"] < 60
    assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0.0
    assert st["hbm_in_use"] == 0 and st["queue_depth"] == 0
    srv.close()


def test_admission_rejects_over_budget(graphs):
    _, cg, _ = graphs[0]
    srv = capi.Server(hbm_budget_bytes=4096)  # 4 KiB: smaller than any traversal of this graph
    gid = srv.add_graph(cg)
    rid, _ = srv.submit(gid, np.ones((4, cg.num_samples), np.float32), capi.UP)
    with pytest.raises(capi.GrglError) as ei:
        srv.wait(rid)
    assert ei.value.status == capi.ERR_NOMEM
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_budget_limits_concurrency_not_results(graphs):
    _, cg, og = graphs[0]
    one = 3 * 1024 * 1024 + (cg.num_nodes + cg.num_mutations) * 64 * 4
    srv = capi.Server(num_streams=4, max_batch_rows=16, hbm_budget_bytes=one)
    gid = srv.add_graph(cg)
    rng = np.random.default_rng(3)
    subs = []
    for _ in range(12):
        x = make_input(rng, np.int32, (8, cg.num_mutations))
        subs.append((x,) + srv.submit(gid, x, capi.DOWN))
    for x, rid, out in subs:
        srv.wait(rid)
        assert bits_equal(out, og.matmul(x, oracle.DOWN))
    st = srv.stats()
    assert st["hbm_peak"] <= one
    srv.close()


def test_cancel_and_queue_full(graphs):
    _, cg, _ = graphs[1]
    srv = capi.Server(num_streams=1, max_batch_rows=1, max_queue=8)
    gid = srv.add_graph(cg)
    x = np.ones((1, cg.num_samples), np.float64)
    subs = []
    busy = 0
    for _ in range(40):
        try:
            subs.append(srv.submit(gid, x, capi.UP))
        except capi.GrglError as e:
            assert e.status == capi.ERR_BUSY
            busy += 1
    assert busy > 0
    cancelled = 0
    for rid, _ in reversed(subs):
        try:
            srv.cancel(rid)
            cancelled += 1
        except capi.GrglError as e:
            assert e.status in (capi.ERR_NOT_READY, capi.ERR_MISUSE)
    for rid, out in subs:
        try:
            srv.wait(rid)
            assert np.allclose(out, cg.matmul_host(x, capi.UP))
        except capi.GrglError as e:
            assert e.status in (capi.ERR_CANCELLED, capi.ERR_MISUSE)
    st = srv.stats()
    assert st["cancelled"] == cancelled and st["rejected"] == busy
    srv.close()


def test_device_submit_zero_copy_equals_alone(graphs):
    import torch

    srv = capi.Server(num_streams=2, max_batch_rows=512, batch_window_us=20000)
    ids = [srv.add_graph(cg) for _, cg, _ in graphs]
    rng = np.random.default_rng(9)
    subs = []
    for _ in range(40):
        gi = int(rng.integers(0, 2))
        _, cg, og = graphs[gi]
        d = capi.UP if rng.random() < 0.5 else capi.DOWN
        K = int(rng.choice([1, 3, 8, 32, 100]))
        dt = np.float32 if rng.random() < 0.5 else np.int32
        cin, cout = (cg.num_samples, cg.num_mutations) if d == capi.UP else (cg.num_mutations, cg.num_samples)
        x = make_input(rng, dt, (K, cin))
        xd = torch.from_numpy(x).cuda()
        yd = torch.zeros((K, cout), dtype=xd.dtype, device="cuda")
        torch.cuda.synchronize()
        rid = srv.submit_device(ids[gi], capi.NUMPY_TO_DTYPE[np.dtype(dt)], d, K, xd.data_ptr(), yd.data_ptr())
        subs.append((rid, x, xd, yd, gi, d))
    for rid, x, xd, yd, gi, d in subs:
        srv.wait(rid)
        _, cg, og = graphs[gi]
        assert bits_equal(yd.cpu().numpy(), cg.matmul_host(x, d))
    assert srv.stats()["batches"] < 40
    srv.close()


def test_submit_misuse_and_close_with_work_queued(graphs):
    """Submissions the server cannot serve are misuse errors at submit time (unknown graph id,
    more rows than max_batch_rows, invalid direction, zero rows) as are unknown request ids,
    and closing a server with requests still queued or running returns cleanly."""
    import ctypes

    _, cg, _ = graphs[0]
    srv = capi.Server(num_streams=1, max_batch_rows=16)
    gid = srv.add_graph(cg)
    L = capi.lib()
    x = np.ones((17, cg.num_samples), np.float32)
    out = np.zeros((17, cg.num_mutations), np.float32)
    rid = ctypes.c_uint64()
    for graph_id, direction, rows in ((gid + 7, capi.UP, 4), (gid, capi.UP, 17), (gid, 3, 4), (gid, capi.UP, 0)):
        st = L.grgl_hip_server_submit(srv.handle, graph_id, capi.F32, direction, rows, x.ctypes.data, out.ctypes.data,
                                      ctypes.byref(rid))
        assert st == capi.ERR_MISUSE, (graph_id, direction, rows)
    with pytest.raises(capi.GrglError) as ei:
        srv.wait(123456789)
    assert ei.value.status == capi.ERR_MISUSE
    y = np.ones((4, cg.num_samples), np.float32)
    keep = [srv.submit(gid, y, capi.UP) for _ in range(30)]
    assert len(keep) == 30
    srv.close()  # cancels what is queued, waits for running batches
