"""grgl_hip_matmul_noalloc (no device allocation; row buffer in the library's fixed scratch
array) against grgl_hip_matmul with device pointers, through the C ABI: bit for bit, for
several dtypes, both directions and the options, on two calls in a row with the second
smaller; inputs larger than the array run as row blocks; misuse is rejected."""

import numpy as np
import pytest

import grgl_amd
from grgl_amd import capi
from helpers import make_input, random_grg

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _both(cg, dtype, x, direction, cout, **kw):
    """(grgl_hip_matmul result, grgl_hip_matmul_noalloc result) as numpy, on one stream."""
    import torch

    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    xd = _dev(x)
    K, cin = x.shape
    sp = torch.cuda.current_stream().cuda_stream
    outs = []
    for fn in (cg.matmul_ptr, cg.matmul_noalloc_ptr):
        y = torch.zeros((K, cout), dtype=xd.dtype, device="cuda")
        fn(dt, xd.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), stream=sp, **kw)
        outs.append(y)
    torch.cuda.synchronize()
    return outs[0].cpu().numpy(), outs[1].cpu().numpy()


def _same(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                                 np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int8, np.bool_])
def test_noalloc_equals_matmul_two_calls(gpu, dtype):
    graphs = [grgl_amd.grg_from_coalescent(1000, 10000, 1, seed=20261203).arrays(),
              random_grg(7, num_samples=300, num_internal=2000, max_children=8, heavy=(90, 700, 3000),
                         num_mutations=3000)]
    rng = np.random.default_rng(5)
    for a in graphs:
        cg = capi.Graph(a)
        for direction in (capi.UP, capi.DOWN):
            cin = cg.num_samples if direction == capi.UP else cg.num_mutations
            cout = cg.num_mutations if direction == capi.UP else cg.num_samples
            for K in (40, 3):  # the second call smaller: leftovers of the first stay in the array
                x = make_input(rng, dtype, (K, cin))
                ref, got = _both(cg, dtype, x, direction, cout)
                assert _same(ref, got), (direction, K)
        cg.close()


def test_noalloc_options(gpu):
    """by_individual, emit_all_nodes, init (vector / matrix), miss, accumulate, order_preserving."""
    import torch

    a = random_grg(13, num_samples=200, num_internal=900, heavy=(150, 400), num_mutations=1200, with_miss=True,
                   with_coal=True)
    cg = capi.Graph(a)
    rng = np.random.default_rng(3)
    sp = torch.cuda.current_stream().cuda_stream
    K = 4
    for dtype in (np.float32, np.int64):
        dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
        for direction in (capi.UP, capi.DOWN):
            cin = cg.num_samples if direction == capi.UP else cg.num_mutations
            cout = cg.num_mutations if direction == capi.UP else cg.num_samples
            nind = cg.num_samples // 2
            cases = [
                dict(x=make_input(rng, dtype, (K, nind if direction == capi.UP else cin)),
                     cout=cout if direction == capi.UP else nind, kw=dict(by_individual=True)),
                dict(x=make_input(rng, dtype, (K, cin)), cout=cg.num_nodes, kw=dict(emit_all=True)),
                dict(x=make_input(rng, dtype, (K, cin)), cout=cout, kw=dict(flags=capi.FLAG_ORDER_PRESERVING)),
            ]
            for c in cases:
                ref, got = _both(cg, dtype, c["x"], direction, c["cout"], **c["kw"])
                assert _same(ref, got), (dtype, direction, c["kw"])
            # init vector / matrix, miss, accumulate (device buffers)
            x = _dev(make_input(rng, dtype, (K, cin)))
            vec = _dev(make_input(rng, dtype, (K,)))
            mat = _dev(make_input(rng, dtype, (K, cg.num_nodes)))
            miss = make_input(rng, dtype, (K, cg.num_mutations))
            prior = make_input(rng, dtype, (K, cout))
            for init_ptr, mode in ((vec.data_ptr(), capi.INIT_VECTOR), (mat.data_ptr(), capi.INIT_MATRIX)):
                res = []
                for fn in (cg.matmul_ptr, cg.matmul_noalloc_ptr):
                    y = _dev(prior)
                    m = _dev(miss)
                    fn(dt, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), init_ptr=init_ptr,
                       init_mode=mode, miss_ptr=m.data_ptr(), flags=capi.FLAG_ACCUMULATE, stream=sp)
                    torch.cuda.synchronize()
                    res.append((y.cpu().numpy(), m.cpu().numpy()))
                assert _same(res[0][0], res[1][0]) and _same(res[0][1], res[1][1]), (dtype, direction, mode)
    cg.close()


def test_noalloc_row_blocks_and_misuse(gpu):
    """100,000 x 500,000 at K = 64 f32 needs a 190 MB row buffer for DOWN: the call runs row
    blocks that fit the 128 MiB array and still equals grgl_hip_matmul (a single traversal);
    then a smaller call.  Host pointers and a graph whose one-row buffer exceeds the array are
    misuse errors."""
    import torch

    a = grgl_amd.grg_from_coalescent(100_000, 500_000, 20, seed=20261203).arrays()
    cg = capi.Graph(a)
    rng = np.random.default_rng(9)
    for direction, K in ((capi.DOWN, 64), (capi.UP, 64), (capi.DOWN, 5)):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        x = rng.standard_normal((K, cin)).astype(np.float32)
        ref, got = _both(cg, np.float32, x, direction, cout)
        assert _same(ref, got), (direction, K)
    x = np.zeros((1, cg.num_mutations), np.float32)
    y = np.zeros((1, cg.num_samples), np.float32)
    with pytest.raises(capi.GrglError, match="device pointers only"):
        cg.matmul_noalloc_ptr(capi.F32, x.ctypes.data, cg.num_mutations, 1, capi.DOWN, y.ctypes.data, y.size,
                              flags=capi.FLAG_HOST_PTRS)
    cg.close()
    # 17,000,000 samples under one root: an f64 UP row buffer of 136 MB for one input row.
    N = 17_000_000
    off = np.array([0] * (N + 1) + [N], dtype=np.uint64)
    big = {"num_samples": N, "ploidy": 2, "down_offsets": off, "down_children": np.arange(N, dtype=np.uint32),
           "mutation_nodes": np.array([N], np.uint32), "mutation_miss_nodes": None, "coal_counts": None}
    cg = capi.Graph(big)
    xs = torch.ones((1, N), dtype=torch.float64, device="cuda")
    ys = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(capi.GrglError, match="graph too large"):
        cg.matmul_noalloc_ptr(capi.F64, xs.data_ptr(), N, 1, capi.UP, ys.data_ptr(), 1)
    # f32 fits (68 MB): the root sums all samples.
    x32 = torch.ones((1, N), dtype=torch.float32, device="cuda")
    y32 = torch.zeros((1, 1), dtype=torch.float32, device="cuda")
    cg.matmul_noalloc_ptr(capi.F32, x32.data_ptr(), N, 1, capi.UP, y32.data_ptr(), 1)
    torch.cuda.synchronize()
    assert y32.item() == float(N)
    cg.close()
