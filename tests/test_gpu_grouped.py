"""Grouped levels (grgl_tuning::chain = 3: the later levels of a direction cut into connected
groups that one block runs by itself, values held in LDS): bit-identical to the level-by-level
launches for every dtype and row class (4-64 bytes), wider rows as 64-byte column slices, every
option that changes how rows start or are read out, stacked batches, host buffers and hipGraph
replay; integer results equal the oracle."""

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import _grgl_amd as M
from grgl_amd import capi
from helpers import bits_equal, csr_from_lists, make_input, random_grg

pytestmark = pytest.mark.gpu

_G = {}


def _graphs():
    if not _G:
        _G["tree"] = grgl_amd.grg_from_coalescent(3000, 30000, 3, seed=5).arrays()
        _G["heavy"] = random_grg(31, num_samples=300, num_internal=2500, max_children=9, heavy=(100, 900, 2500),
                                 num_mutations=2000, with_miss=True, with_coal=True)
        rng = np.random.default_rng(41)
        N = 600
        kids = [[] for _ in range(N)]
        for i in range(900):
            kids.append(sorted(rng.choice(N, size=70 + i % 31, replace=False).tolist()))
        kids.append(list(range(N, N + 900, 7)))
        muts = rng.integers(N, len(kids), size=1500)
        _G["wide"] = csr_from_lists(N, kids, muts)
        src = grgl_amd.grg_from_coalescent(2000, 20000, 4, seed=9)
        off, car = src.mutation_carriers()
        _G["built"] = grgl_amd.fast_grg_from_samples(2000, 2, carrier_offsets=off, carriers=car).arrays()
    return _G


def _grouped_classes(arrays, direction):
    g = grgl_amd.grg_from_arrays(arrays["num_samples"], arrays["ploidy"], arrays["down_offsets"],
                                 arrays["down_children"], arrays["mutation_nodes"])
    d = grgl_amd.UP if direction == capi.UP else grgl_amd.DOWN
    return [p["slot_bytes"] for p in M._group_plans(g, d) if p["level_begin"] < p["level_end"]]


def _run(cg, dt, x, K, direction, cin, cout, stream, **kw):
    import torch

    y = torch.zeros((K, cout), dtype=x.dtype, device="cuda")
    cg.matmul_ptr(dt, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), stream=stream, **kw)
    torch.cuda.synchronize()
    return y.cpu().numpy()


DTYPES = [np.float32, np.int32, np.float64, np.int64, np.uint32, np.int16, np.uint8, np.bool_]


@pytest.mark.parametrize("name", ["tree", "heavy", "wide", "built"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_grouped_equals_level_launches(gpu, name, dtype):
    import torch

    arrays = _graphs()[name]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    es = np.dtype(dtype).itemsize
    rng = np.random.default_rng(3)
    sp = torch.cuda.current_stream().cuda_stream
    for direction in (capi.UP, capi.DOWN):
        assert _grouped_classes(arrays, direction), (name, direction)
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        # Rows of 4-64 bytes (one slot class each), then two wider ones (64-byte column slices,
        # the last one partly padding).
        wide = [k for k in (9, 17, 40, 100, 130) if 64 < k * es <= 640][:2]
        for K in [k for k in (1, 2, 3, 4, 8, 12, 16, 24, 32, 64) if k * es <= 64] + wide:
            print(f"grouped: {name} {np.dtype(dtype).name} dir={direction} K={K}", flush=True)
            x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()
            cg.set_tuning(direction, dt, K, chain=1)
            ref = _run(cg, dt, x, K, direction, cin, cout, sp)
            if np.dtype(dtype).kind in "iub":
                assert bits_equal(ref, og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN))
            for tw, use_graph, blk in ((1, 0, 0), (4, 1, 0), (4, 0, 256)):
                cg.set_tuning(direction, dt, K, team_lanes=tw, chain=3, use_graph=use_graph, block_threads=blk)
                for rep in range(2):
                    got = _run(cg, dt, x, K, direction, cin, cout, sp)
                    assert bits_equal(got, ref), (name, direction, K, tw, use_graph, blk, rep)
            cg.set_tuning(direction, dt, K)
    cg.close()


@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64])
def test_grouped_options(gpu, dtype):
    """Node initialisation (vector, matrix, xtx), emit_all_nodes, by_individual, miss and
    accumulation through grouped levels equal the level-by-level launches (K = 20: column
    slices)."""
    import torch

    arrays = _graphs()["heavy"]
    cg = capi.Graph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    rng = np.random.default_rng(5)
    sp = torch.cuda.current_stream().cuda_stream
    V, N, Mu = cg.num_nodes, cg.num_samples, cg.num_mutations
    for direction in (capi.UP, capi.DOWN):
        for K in (1, 3, 20):
            cin = N if direction == capi.UP else Mu
            x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()
            vec = torch.from_numpy(make_input(rng, dtype, (K,))).cuda()
            mat = torch.from_numpy(make_input(rng, dtype, (K, V))).cuda()
            miss = torch.from_numpy(make_input(rng, dtype, (K, N if direction == capi.DOWN else Mu))).cuda()
            cases = [
                ("vector", {}, dict(init_ptr=vec.data_ptr(), init_mode=capi.INIT_VECTOR)),
                ("matrix", {}, dict(init_ptr=mat.data_ptr(), init_mode=capi.INIT_MATRIX)),
                ("xtx", {}, dict(init_mode=capi.INIT_XTX)),
                ("emit_all", dict(emit_all=True), {}),
                ("accumulate", {}, dict(flags=capi.FLAG_ACCUMULATE)),
            ]
            if direction == capi.UP:
                cases.append(("miss", {}, dict(miss_ptr=miss.data_ptr())))
            else:
                cases.append(("by_individual", dict(by_individual=True), {}))
            for label, shape_kw, kw in cases:
                if label == "emit_all":
                    out_cols = V
                elif label == "by_individual":
                    out_cols = N // 2
                else:
                    out_cols = Mu if direction == capi.UP else N
                outs = []
                for chain in (1, 3):
                    print(f"grouped options: {np.dtype(dtype).name} dir={direction} K={K} {label} chain={chain}",
                          flush=True)
                    cg.set_tuning(direction, dt, K, chain=chain)
                    y = torch.ones((K, out_cols), dtype=x.dtype, device="cuda")
                    miss.zero_()
                    cg.matmul_ptr(dt, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), stream=sp,
                                  emit_all=shape_kw.get("emit_all", False),
                                  by_individual=shape_kw.get("by_individual", False), **kw)
                    torch.cuda.synchronize()
                    outs.append((y.cpu().numpy(), miss.cpu().numpy().copy()))
                assert bits_equal(outs[0][0], outs[1][0]), (direction, K, label)
                if label == "miss":
                    assert bits_equal(outs[0][1], outs[1][1]), (direction, K, label)
                cg.set_tuning(direction, dt, K)
    cg.close()


def test_grouped_batch_host_buffers_and_order_preserving(gpu):
    """Stacked requests through one grouped traversal equal the oracle; host buffers and the
    order-preserving mode (which keeps per-level launches) are unchanged."""
    import torch

    arrays = _graphs()["built"]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    rng = np.random.default_rng(8)
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        Ks = [1, 2, 5, 3]
        xs = [torch.from_numpy(make_input(rng, np.int32, (k, cin))).cuda() for k in Ks]
        for k in range(1, 65):
            cg.set_tuning(direction, capi.I32, k, chain=3)
        outs = [torch.zeros((k, cout), dtype=torch.int32, device="cuda") for k in Ks]
        cg.matmul_batch_ptr(capi.I32, direction, [(k, x.data_ptr(), o.data_ptr()) for k, x, o in zip(Ks, xs, outs)],
                            stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for x, o in zip(xs, outs):
            want = og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN)
            assert bits_equal(o.cpu().numpy(), want), direction
        xh = make_input(rng, np.float32, (4, cin))
        cg.set_tuning(direction, capi.F32, 4, chain=1)
        a = cg.matmul_host(xh, direction)
        b = cg.matmul_host(xh, direction, flags=capi.FLAG_ORDER_PRESERVING)
        cg.set_tuning(direction, capi.F32, 4, chain=3)
        assert bits_equal(cg.matmul_host(xh, direction), a)
        assert bits_equal(cg.matmul_host(xh, direction, flags=capi.FLAG_ORDER_PRESERVING), b)
    cg.close()


def test_grouped_tuning_report(gpu):
    """get_tuning names the form a call takes: grouped levels (and their block size) where the
    direction has them for the row size, one launch per level otherwise."""
    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    for direction in (capi.UP, capi.DOWN):
        t = cg.get_tuning(direction, capi.F32, 1)
        assert t["chain"] == 3 and t["block_threads"] == 1024, t
        assert cg.get_tuning(direction, capi.F32, 16)["chain"] == 3
        assert cg.get_tuning(direction, capi.F32, 16)["block_threads"] == 256
        # Rows over 64 bytes: 64-byte slices by default up to 512 bytes UP, 256 bytes DOWN (over
        # DOWN's small levels only).
        assert cg.get_tuning(direction, capi.F32, 32)["chain"] == 3
        assert cg.get_tuning(direction, capi.F32, 64)["chain"] == 3
        assert cg.get_tuning(direction, capi.F32, 128)["chain"] == (3 if direction == capi.UP else 1)
        assert cg.get_tuning(direction, capi.F32, 129)["chain"] == 1
        cg.set_tuning(direction, capi.F32, 1, block_threads=256)
        assert cg.get_tuning(direction, capi.F32, 1)["block_threads"] == 256
        cg.set_tuning(direction, capi.F32, 32, chain=3)
        assert cg.get_tuning(direction, capi.F32, 32)["chain"] == 3
        cg.set_tuning(direction, capi.F32, 511, chain=3)
        assert cg.get_tuning(direction, capi.F32, 511)["chain"] == 3
        cg.set_tuning(direction, capi.F32, 1, chain=1)
        assert cg.get_tuning(direction, capi.F32, 1)["chain"] == 1
    cg.close()
