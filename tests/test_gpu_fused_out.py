"""DOWN traversals of rows of at most 16 bytes write their last level (every sample's row)
straight to the output columns instead of through a read-out launch: bit-identical to the
traversal with a read-out (hipGraph replay and accumulation never take the fused form), for
every dtype and K up to 16 bytes a row (and just past it), integers equal to the oracle, rows
of -0.0 read out as +0.0 as the read-out does, heavy sample rows (block and wavefront forms)
included, and childless nodes that are not samples left in the row buffer."""

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import capi
from helpers import bits_equal, csr_from_lists, make_input

pytestmark = pytest.mark.gpu

_G = {}


def _graphs():
    if not _G:
        _G["tree"] = grgl_amd.grg_from_coalescent(3000, 30000, 3, seed=7).arrays()
        # 600 samples, each the child of ~130 of 900 wide nodes: every sample row is heavy (the
        # per-wavefront heavy form); a few childless internal nodes carry mutations too.
        rng = np.random.default_rng(23)
        N = 600
        kids = [[] for _ in range(N)]
        for i in range(900):
            kids.append(sorted(rng.choice(N, size=70 + i % 31, replace=False).tolist()))
            if i % 150 == 0:
                kids.append([])
        muts = rng.integers(N, len(kids), size=2000)
        _G["wide"] = csr_from_lists(N, kids, muts)
        # One sample with 100 parents (a heavy row in a level of light ones: the block form).
        N = 300
        kids = [[] for _ in range(N)]
        for i in range(400):
            pool = list(range(1, len(kids)))
            c = sorted(rng.choice(pool, size=3, replace=False).tolist())
            if i < 100:
                c = sorted(set(c) | {0})
            kids.append(c)
        muts = rng.integers(0, len(kids), size=900)
        _G["one_heavy"] = csr_from_lists(N, kids, muts)
    return _G


def _down(cg, dt, x, stream, **kw):
    import torch

    K = x.shape[0]
    y = torch.full((K, cg.num_samples), 7, dtype=x.dtype, device="cuda")
    cg.matmul_ptr(dt, x.data_ptr(), cg.num_mutations, K, capi.DOWN, y.data_ptr(), y.numel(), stream=stream, **kw)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("name", ["tree", "wide", "one_heavy"])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int64, np.uint32, np.int16, np.uint8,
                                   np.bool_])
def test_fused_output_equals_readout(gpu, name, dtype):
    import torch

    arrays = _graphs()[name]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    es = np.dtype(dtype).itemsize
    sp = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(31)
    for K in [k for k in (1, 2, 3, 4, 5, 8, 16, 17) if k * es <= 20]:
        x = torch.from_numpy(make_input(rng, dtype, (K, cg.num_mutations))).cuda()
        vec = torch.from_numpy(make_input(rng, dtype, (K,))).cuda()
        for chain in (0, 1, 3):
            print(f"fused out: {name} {np.dtype(dtype).name} K={K} chain={chain}", flush=True)
            cg.set_tuning(capi.DOWN, dt, K, chain=chain)
            init = {} if dtype == np.bool_ else dict(init_ptr=vec.data_ptr(), init_mode=capi.INIT_VECTOR)
            fused = _down(cg, dt, x, sp)
            fused_init = _down(cg, dt, x, sp, **init)
            cg.set_tuning(capi.DOWN, dt, K, chain=chain, use_graph=1)  # replayed graphs keep the read-out
            for rep in range(2):
                assert bits_equal(fused, _down(cg, dt, x, sp)), (name, K, chain, rep)
                assert bits_equal(fused_init, _down(cg, dt, x, sp, **init)), (name, K, chain, rep)
            if np.dtype(dtype).kind in "iub":
                assert bits_equal(fused, og.matmul(x.cpu().numpy(), oracle.DOWN)), (name, K, chain)
        cg.set_tuning(capi.DOWN, dt, K)
        if dtype == np.bool_:
            continue
        # Accumulation keeps the read-out (it adds to what the output holds).
        base = torch.from_numpy(make_input(rng, dtype, (K, cg.num_samples))).cuda()
        y = base.clone()
        cg.matmul_ptr(dt, x.data_ptr(), cg.num_mutations, K, capi.DOWN, y.data_ptr(), y.numel(), stream=sp,
                      flags=capi.FLAG_ACCUMULATE)
        torch.cuda.synchronize()
        plain = _down(cg, dt, x, sp)
        if np.dtype(dtype).kind in "iu":
            assert bits_equal(y.cpu().numpy(), (base.cpu().numpy() + plain).astype(dtype)), K
        else:
            assert np.array_equal(y.cpu().numpy(), base.cpu().numpy() + plain), K
    cg.close()


def test_fused_output_negative_zero(gpu):
    """Rows of -0.0 (every node starting at -0.0, every input -0.0) come out as +0.0
    (0 + value), as through the read-out."""
    import torch

    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    sp = torch.cuda.current_stream().cuda_stream
    x = torch.full((1, cg.num_mutations), -0.0, dtype=torch.float32, device="cuda")
    vec = torch.full((1,), -0.0, dtype=torch.float32, device="cuda")
    kw = dict(init_ptr=vec.data_ptr(), init_mode=capi.INIT_VECTOR)
    cg.set_tuning(capi.DOWN, capi.F32, 1)
    fused = _down(cg, capi.F32, x, sp, **kw)
    assert not np.signbit(fused).any() and not fused.any()
    cg.set_tuning(capi.DOWN, capi.F32, 1, use_graph=1)
    assert bits_equal(fused, _down(cg, capi.F32, x, sp, **kw))
    cg.close()
