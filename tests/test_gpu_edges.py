"""Edge cases the reference's matmul accepts or rejects (src/python/_grgl.cpp:100-241,
include/grgl/internal_mult.h:130-158): graphs without mutations or internal nodes, unmapped
mutations, a single haploid sample, the largest and one-too-large row counts, empty inputs."""

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import bits_equal, csr_from_lists, make_input, random_grg

pytestmark = pytest.mark.gpu
INVALID = grgl_amd.INVALID_NODE
UP, DOWN = grgl_amd.UP, grgl_amd.DOWN


def _pg(a):
    return grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                    a["mutation_nodes"])


def _both(a, x, direction, **kw):
    got = grgl_amd.matmul(_pg(a), x, direction, **kw)
    want = oracle.OracleGraph(a).matmul(x, oracle.UP if direction == UP else oracle.DOWN, **kw)
    return got, want


def test_no_mutations(gpu):
    a = csr_from_lists(6, [[] for _ in range(6)] + [[0, 1], [2, 3, 6]], np.zeros(0, np.int64))
    g = _pg(a)
    out = grgl_amd.matmul(g, np.ones((3, 6), np.float32), UP)
    assert out.shape == (3, 0)
    out = grgl_amd.matmul(g, np.ones((2, 6), np.int32), UP, emit_all_nodes=True)
    assert out[:, 7].tolist() == [4, 4]
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(g, np.ones((1, 0), np.float64), DOWN)


def test_samples_only_graph(gpu):
    kids = [[] for _ in range(10)]
    muts = np.array([0, 3, 3, 9, INVALID, 5], np.int64)
    a = csr_from_lists(10, kids, muts)
    rng = np.random.default_rng(1)
    for dt in (np.float64, np.int32, np.bool_):
        x = make_input(rng, dt, (4, 10))
        got, want = _both(a, x, UP)
        assert bits_equal(got, want)
        w = make_input(rng, dt, (4, 6))
        got, want = _both(a, w, DOWN)
        assert bits_equal(got, want)


def test_all_mutations_unmapped(gpu):
    a = csr_from_lists(4, [[], [], [], [], [0, 1], [2, 3, 4]], np.full(5, INVALID, np.int64))
    up = grgl_amd.matmul(_pg(a), np.ones((2, 4), np.float32), UP)
    assert up.shape == (2, 5) and not up.any()
    down = grgl_amd.matmul(_pg(a), np.ones((2, 5), np.float32), DOWN)
    assert down.shape == (2, 4) and not down.any()


def test_single_haploid_sample(gpu):
    a = csr_from_lists(1, [[], [0], [1]], np.array([0, 1, 2, INVALID], np.int64), ploidy=1)
    got, want = _both(a, np.array([[2.5]]), UP)
    assert bits_equal(got, want) and got.tolist() == [[2.5, 2.5, 2.5, 0.0]]
    got, want = _both(a, np.array([[1.0, 2.0, 4.0, 8.0]]), DOWN)
    assert bits_equal(got, want) and got.tolist() == [[7.0]]


def test_row_limits(gpu):
    """Any number of input rows, as on the CPU path: more than 65536 rows run as several
    traversals of independent row blocks.  Zero rows is misuse (_grgl.cpp:100-110)."""
    a = csr_from_lists(4, [[], [], [], [], [0, 1], [2, 3], [4, 5]], np.array([4, 6, 1], np.int64))
    rng = np.random.default_rng(9)
    for rows in (65536, 65537, 140001):
        x = rng.integers(-100, 100, size=(rows, 4)).astype(np.int8)
        got, want = _both(a, x, UP)
        assert bits_equal(got, want), rows
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(_pg(a), np.ones((0, 4), np.int8), UP)


def test_many_rows_with_init_and_miss(gpu):
    """Row blocks slice the init vector / matrix and the miss matrix with the input."""
    from helpers import random_grg

    a = random_grg(51, num_samples=20, num_internal=30, num_mutations=25, with_miss=True, with_coal=True)
    g = _pg_miss(a)
    og = oracle.OracleGraph(a)
    rng = np.random.default_rng(10)
    K = 70000
    for direction, od in ((UP, oracle.UP), (DOWN, oracle.DOWN)):
        cols = g.num_samples if direction == UP else g.num_mutations
        x = make_input(rng, np.int32, (K, cols))
        for init in (make_input(rng, np.int32, (K,)), make_input(rng, np.int32, (K, g.num_nodes)), "xtx"):
            assert bits_equal(grgl_amd.matmul(g, x, direction, init=init), og.matmul(x, od, init=init))
        m1 = make_input(rng, np.int32, (K, g.num_mutations))
        m2 = m1.copy()
        assert bits_equal(grgl_amd.matmul(g, x, direction, miss=m1), og.matmul(x, od, miss=m2))
        assert bits_equal(m1, m2)


def test_many_rows_device_pointers(gpu):
    import torch

    from grgl_amd import capi
    from helpers import random_grg

    a = random_grg(52, num_samples=30, num_internal=60, heavy=(70,), num_mutations=40)
    cg = capi.Graph(a)
    og = oracle.OracleGraph(a)
    rng = np.random.default_rng(11)
    K = 66000
    for direction, od, cin, cout in ((capi.UP, oracle.UP, cg.num_samples, cg.num_mutations),
                                     (capi.DOWN, oracle.DOWN, cg.num_mutations, cg.num_samples)):
        xh = make_input(rng, np.float64, (K, cin))
        x = torch.from_numpy(xh).cuda()
        y = torch.zeros((K, cout), dtype=torch.float64, device="cuda")
        cg.matmul_ptr(capi.F64, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(),
                      flags=capi.FLAG_ORDER_PRESERVING)
        torch.cuda.synchronize()
        assert bits_equal(y.cpu().numpy(), og.matmul(xh, od))
    cg.close()


def _pg_miss(a):
    return grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                    a["mutation_nodes"], mutation_miss_nodes=a.get("mutation_miss_nodes"),
                                    coal_counts=a.get("coal_counts"))


@pytest.mark.parametrize("ploidy", [3, 4])
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int32, np.int8, np.bool_])
def test_polyploid_by_individual_and_xtx(gpu, ploidy, dtype):
    """Individuals of 3 or 4 haplotypes: by_individual sums `ploidy` consecutive haplotypes
    (grg.h:1448,1463), xtx initialises with 2 x the coalescence counts (grg.h:1369-1385)."""
    from helpers import random_grg

    a = random_grg(40 + ploidy, num_samples=ploidy * 21, num_internal=300, heavy=(90, 150), num_mutations=240,
                   ploidy=ploidy, with_coal=True)
    g = _pg_coal(a)
    og = oracle.OracleGraph(a)
    rng = np.random.default_rng(ploidy)
    exact = np.dtype(dtype).kind in "biu"
    for direction, od in ((UP, oracle.UP), (DOWN, oracle.DOWN)):
        cols = g.num_individuals if direction == UP else g.num_mutations
        for K in (1, 5, 40):
            x = make_input(rng, dtype, (K, cols))
            want = og.matmul(x, od, by_individual=True)
            got = grgl_amd.matmul(g, x, direction, by_individual=True, order_preserving=not exact)
            assert bits_equal(got, want), (direction, K)
        xf = make_input(rng, dtype, (2, g.num_samples if direction == UP else g.num_mutations))
        if dtype == np.bool_:  # as the reference (src/python/_grgl.cpp:236)
            with pytest.raises(RuntimeError):
                grgl_amd.matmul(g, xf, direction, init="xtx")
            continue
        got = grgl_amd.matmul(g, xf, direction, init="xtx", order_preserving=not exact)
        assert bits_equal(got, og.matmul(xf, od, init="xtx")), direction


def _pg_coal(a):
    return grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                    a["mutation_nodes"], coal_counts=a.get("coal_counts"))


def test_large_results_in_pinned_pool(gpu):
    """Results of 1 MB or more come from a pool of page-locked blocks: they are ordinary,
    writable numpy arrays, correct, and their blocks are reused after the arrays are freed
    (values never leak from one result into the next, with or without accumulation)."""
    import gc

    g = grgl_amd.grg_from_coalescent(2000, 120000, 4, seed=3)
    a = g.arrays()
    og = oracle.OracleGraph(a)
    rng = np.random.default_rng(12)
    kept = []
    for i in range(12):
        K = [3, 5, 8][i % 3]
        x = rng.integers(-9, 9, size=(K, g.num_samples)).astype(np.int32)
        out = grgl_amd.matmul(g, x, UP)
        assert out.nbytes >= 1 << 20 and out.flags.writeable and out.flags.c_contiguous
        assert bits_equal(out, og.matmul(x, oracle.UP)), i
        out[0, 0] += 1  # writable
        if i % 2:
            kept.append(out)
        del out
        gc.collect()
    w = rng.integers(-9, 9, size=(4, g.num_mutations)).astype(np.int32)
    miss = np.zeros((4, g.num_mutations), np.int32)
    out = grgl_amd.matmul(g, rng.integers(0, 2, size=(4, g.num_samples)).astype(np.int32), UP, miss=miss)
    assert out.shape == (4, g.num_mutations)
    down = grgl_amd.matmul(g, w, DOWN)
    assert bits_equal(down, og.matmul(w, oracle.DOWN))


def test_large_host_inputs_split_over_threads(gpu):
    """Host inputs of 32 MB or more are uploaded in 2 or 4 slices from as many threads, each
    on its own stream; the traversal waits for all of them (results bit-exact)."""
    g = grgl_amd.grg_from_coalescent(2000, 120000, 4, seed=4)
    og = oracle.OracleGraph(g.arrays())
    rng = np.random.default_rng(13)
    for K in (80, 150):  # 38 MB and 72 MB of int32 input
        w = rng.integers(-50, 50, size=(K, g.num_mutations)).astype(np.int32)
        assert w.nbytes >= 32 << 20
        for _ in range(2):
            assert bits_equal(grgl_amd.matmul(g, w, DOWN), og.matmul(w, oracle.DOWN)), K


def test_stream_ceiling(gpu):
    """grgl_hip_stream_ceiling: measured read and copy rates are positive and under the
    published 8 TB/s (with 5% margin); out-of-range arguments are misuse errors."""
    from grgl_amd import capi

    c = capi.stream_ceiling(0, 256 << 20)
    assert 0 < c["read_GBps"] < 8400 and 0 < c["copy_GBps"] < 8400
    with pytest.raises(RuntimeError):
        capi.stream_ceiling(0, 1024)
    with pytest.raises(RuntimeError):
        capi.stream_ceiling(99, 1 << 20)


def test_graph_without_mutations_or_edges(gpu):
    """Degenerate graphs, as pygrgl.matmul treats them (src/python/_grgl.cpp:196-203): with no
    mutations an UP result has zero columns and a DOWN input cannot be given (zero columns is
    misuse); with no edges, UP reads the mutated samples and DOWN writes them."""
    tree = csr_from_lists(4, [[], [], [], [], [0, 1], [2, 3], [4, 5]], np.zeros(0, np.uint32))
    g = _pg(tree)
    assert g.num_mutations == 0
    for dt in (np.int32, np.float64):
        out = grgl_amd.matmul(g, np.ones((2, 4), dt), UP)
        assert out.shape == (2, 0) and out.dtype == dt
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(g, np.ones((2, 0), np.int32), DOWN)
    # every node's subtree sample count
    nodes = grgl_amd.matmul(g, np.ones((1, 4), np.int32), UP, emit_all_nodes=True)
    assert nodes.tolist() == [[1, 1, 1, 1, 2, 2, 4]]

    flat = csr_from_lists(4, [[], [], [], []], [2, 0, INVALID, 2])
    g2 = _pg(flat)
    og = oracle.OracleGraph(flat)
    rng = np.random.default_rng(8)
    x = rng.integers(-9, 9, size=(3, 4)).astype(np.int32)
    up = grgl_amd.matmul(g2, x, UP)
    assert up.tolist() == np.stack([x[:, 2], x[:, 0], np.zeros(3, np.int32), x[:, 2]], axis=1).tolist()
    assert bits_equal(up, og.matmul(x, oracle.UP))
    w = rng.integers(-9, 9, size=(3, 4)).astype(np.int32)
    down = grgl_amd.matmul(g2, w, DOWN)
    zero = np.zeros(3, np.int32)
    want = np.stack([w[:, 1], zero, w[:, 0] + w[:, 3], zero], axis=1)
    assert down.tolist() == want.tolist()
    assert bits_equal(down, og.matmul(w, oracle.DOWN))
    f = rng.standard_normal((2, 4)).astype(np.float32)
    assert bits_equal(grgl_amd.matmul(g2, f, DOWN), og.matmul(f, oracle.DOWN))


def test_strided_and_fortran_inputs_are_made_contiguous(gpu):
    """matmul() takes its input through py::array::ensure(c_style | forcecast), as the reference
    does (src/python/_grgl.cpp:194): Fortran-ordered and strided views give the results of
    their contiguous copies; the dtype is the input's own, so a Python list is rejected."""
    a = random_grg(61, num_samples=40, num_internal=150, num_mutations=220)
    g = _pg(a)
    rng = np.random.default_rng(21)
    for dt in (np.float64, np.int32):
        x = make_input(rng, dt, (5, 2 * g.num_samples))[:, ::2]
        assert not x.flags.c_contiguous
        want = grgl_amd.matmul(g, np.ascontiguousarray(x), UP)
        assert bits_equal(grgl_amd.matmul(g, x, UP), want)
        w = np.asfortranarray(make_input(rng, dt, (6, g.num_mutations)))
        assert not w.flags.c_contiguous
        assert bits_equal(grgl_amd.matmul(g, w, DOWN), grgl_amd.matmul(g, np.ascontiguousarray(w), DOWN))
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(g, [[1.0] * g.num_samples], UP)


def test_host_alloc_buffers_through_the_c_abi(gpu):
    """grgl_hip_host_alloc / grgl_hip_host_free: page-locked buffers used as host input and
    output of a matmul give the same bits as ordinary numpy buffers; zero bytes is misuse."""
    import ctypes

    from grgl_amd import capi

    a = random_grg(63, num_samples=50, num_internal=200, num_mutations=300)
    cg = capi.Graph(a)
    L = capi.lib()
    rng = np.random.default_rng(4)
    K = 6
    x = make_input(rng, np.float32, (K, cg.num_samples))
    want = cg.matmul_host(x, capi.UP)
    pin_in, pin_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.grgl_hip_host_alloc(x.nbytes, ctypes.byref(pin_in)) == capi.OK and pin_in.value
    assert L.grgl_hip_host_alloc(want.nbytes, ctypes.byref(pin_out)) == capi.OK and pin_out.value
    try:
        xin = np.ctypeslib.as_array((ctypes.c_float * x.size).from_address(pin_in.value)).reshape(x.shape)
        yout = np.ctypeslib.as_array((ctypes.c_float * want.size).from_address(pin_out.value)).reshape(want.shape)
        xin[...] = x
        yout[...] = 0
        cg.matmul_ptr(capi.F32, pin_in.value, cg.num_samples, K, capi.UP, pin_out.value, yout.size,
                      flags=capi.FLAG_HOST_PTRS)
        assert bits_equal(yout.copy(), want)
    finally:
        assert L.grgl_hip_host_free(pin_in) == capi.OK
        assert L.grgl_hip_host_free(pin_out) == capi.OK
    bad = ctypes.c_void_p()
    assert L.grgl_hip_host_alloc(0, ctypes.byref(bad)) == capi.ERR_MISUSE
    assert L.grgl_hip_host_free(None) == capi.OK
    cg.close()
