"""GPU parity: the HIP path against the oracle (and the reference's golden vectors).

Bars (stated here, enforced below):
  * integer and bool dtypes: bit-exact;
  * floats with order_preserving=True: bit-exact (same summation order as the CPU path);
  * floats by default (high fan-in lists split 64 ways):
      - FP64: <= 1e-12 normwise per output row (max |gpu - ref| / max |ref|) against the oracle;
      - FP32, any graph: every element within the a priori rounding bound of a correct FP32
        evaluation (oracle/fp_bounds.py: sum over the nodes reaching the output of
        (fan-in - 1) * 2^-24 * sum |list terms|; bound_ratio <= 1), so no term is dropped,
        repeated or misplaced;
      - FP32 on the coalescent benchmark graphs with N(0,1) inputs (1,000 x 10,000;
        100,000 x 500,000; 200,000 x 1,000,000): <= 1e-6 normwise per row against FP64.
        Arbitrary test DAGs carry no such bar: there the reference's own FP32 result is up to
        2.3e-6 from FP64 (high fan-in lists of nearly cancelling terms).
"""

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import capi
from oracle import fp_bounds
from helpers import (bits_equal, depth3_bin_tree, direct_map_grg, make_input, normwise_rel_err, random_grg,
                     sample8_grg, vcf_fixture)

pytestmark = pytest.mark.gpu

UP, DOWN = grgl_amd.UP, grgl_amd.DOWN
INT_DTYPES = [np.int64, np.int32, np.int16, np.int8, np.uint64, np.uint32, np.uint16, np.uint8, np.bool_]
FLOAT_DTYPES = [np.float64, np.float32]


def _pg(arrays):
    return grgl_amd.grg_from_arrays(arrays["num_samples"], arrays["ploidy"], arrays["down_offsets"],
                                    arrays["down_children"], arrays["mutation_nodes"],
                                    mutation_miss_nodes=arrays.get("mutation_miss_nodes"),
                                    coal_counts=arrays.get("coal_counts"))


def _odir(d):
    return oracle.UP if d == UP else oracle.DOWN


FP32_BAR = 1e-6  # normwise, against FP64, on the benchmark graphs


def _check_float(gpu_out, x, og, direction, bar=None):
    """FP64: <= 1e-12 normwise.  FP32: within the rounding bound, and <= `bar` normwise when
    given.  Returns the FP32 error report (None for FP64)."""
    if x.dtype == np.float64:
        ref64 = og.matmul(x, _odir(direction))
        err = normwise_rel_err(gpu_out, ref64)
        assert err <= 1e-12, err
        return None
    rep = fp_bounds.fp32_report(og, x, _odir(direction), gpu_out, with_reference=False)
    assert rep["bound_ratio"] <= 1.0, rep
    if bar is not None:
        assert rep["normwise"] <= bar, rep
    return rep


# --- golden vectors on the device --------------------------------------------------------

def test_golden_depth3(gpu):
    g = _pg(depth3_bin_tree([1, 5, 6]))
    assert grgl_amd.matmul(g, np.array([[1.0, 2.0, 1.0]]), DOWN).tolist() == [[1.0, 2.0, 3.0, 3.0]]
    assert grgl_amd.matmul(g, np.ones((1, 4)), UP).tolist() == [[1.0, 2.0, 4.0]]
    down = grgl_amd.matmul(g, np.array([[1.0, 2.0, 1.0], [0.0, 1.0, 1.0]]), DOWN)
    assert down.ravel().tolist() == [1.0, 2.0, 3.0, 3.0, 1.0, 1.0, 2.0, 2.0]
    v = np.ones((2, 4))
    v[1] /= 4
    np.testing.assert_allclose(grgl_amd.matmul(g, v, UP).ravel(), [1, 2, 4, 0.25, 0.5, 1.0], atol=1e-3)
    assert grgl_amd.dot_product(g, np.array([1.0, 2.0, 1.0]), DOWN).tolist() == [1.0, 2.0, 3.0, 3.0]


def test_golden_xtx_and_change_samples(gpu):
    g = _pg(depth3_bin_tree([6, 5, 4], coal=[0, 0, 0, 0, 1, 1, 0]))
    assert grgl_amd.matmul(g, np.ones((1, 4), np.int64), UP, init="xtx").tolist() == [[8, 4, 4]]
    assert grgl_amd.matmul(g, np.array([[11, 41, 37, 73]], np.int64), UP).tolist() == [[162, 110, 52]]
    nocoal = _pg(depth3_bin_tree([6, 5, 4]))
    with pytest.raises(RuntimeError, match="Coalescent counts"):
        grgl_amd.matmul(nocoal, np.ones((1, 4), np.int64), UP, init="xtx")


def test_golden_vcf_allele_counts(gpu):
    fx = vcf_fixture()
    g = _pg(direct_map_grg(fx))
    counts = grgl_amd.matmul(g, np.ones((1, g.num_samples), np.int32), UP)[0]
    np.testing.assert_array_equal(counts, fx["expect_alt_count"])
    zyg = grgl_amd.matmul(g, np.ones((1, g.num_individuals), np.int32), UP, by_individual=True)[0]
    np.testing.assert_array_equal(zyg, fx["expect_Aa"] + 2 * fx["expect_aa"])


def test_golden_vcf_built_graph(gpu):
    """The reference's expect files on the graph the constructor builds from the same VCF:
    allele counts, and homozygous counts from the coalescence counts (init="xtx")."""
    fx = vcf_fixture()
    g = _pg(_built_vcf())
    counts = grgl_amd.matmul(g, np.ones((1, g.num_samples), np.int32), UP)[0]
    np.testing.assert_array_equal(counts, fx["expect_alt_count"])
    hom2 = grgl_amd.matmul(g, np.zeros((1, g.num_samples), np.int64), UP, init="xtx")[0]
    np.testing.assert_array_equal(hom2, 2 * fx["expect_aa"])
    zyg = grgl_amd.matmul(g, np.ones((1, g.num_individuals), np.int32), UP, by_individual=True)[0]
    np.testing.assert_array_equal(zyg, fx["expect_Aa"] + 2 * fx["expect_aa"])


def test_missing_data_built_graph(gpu):
    """The reference's missing-data matmul test (test/endtoend/test_missing.py:55-89) on the
    device, on the graph the constructor builds from test-200-samples.miss.igd: UP by individual
    with the `miss` output and DOWN with a `miss` input reproduce the dense genotype matrix with
    missing calls at the allele frequency; int64 results equal the oracle's bit for bit."""
    from test_fast_build import _missing_build, _missing_x

    g, N, pos, is_miss, sets = _missing_build()
    X, freq, _ = _missing_x(N, pos, is_miss, sets)
    a = g.arrays()
    dg = _pg(a)
    og = oracle.OracleGraph(a)
    rng = np.random.default_rng(23)
    rv = rng.standard_normal((7, N // 2))
    M = np.zeros((7, g.num_mutations))
    up = grgl_amd.matmul(dg, rv, UP, by_individual=True, miss=M)
    np.testing.assert_allclose(rv @ X, up + M * freq, rtol=1e-10, atol=1e-10)
    rw = rng.standard_normal((7, g.num_mutations))
    down = grgl_amd.matmul(dg, rw, DOWN, by_individual=True, miss=np.ascontiguousarray(freq[None, :] * rw))
    np.testing.assert_allclose(rw @ X.T, down, rtol=1e-10, atol=1e-10)
    xi = rng.integers(-5, 6, size=(3, N)).astype(np.int64)
    Mg, Mo = np.zeros((3, g.num_mutations), np.int64), np.zeros((3, g.num_mutations), np.int64)
    assert bits_equal(grgl_amd.matmul(dg, xi, UP, miss=Mg), og.matmul(xi, oracle.UP, miss=Mo))
    assert bits_equal(Mg, Mo)


def test_built_graph_equals_tree_graph(gpu):
    """Two graphs of one genotype matrix (tree-direct and constructor-built, 1,000 x 10,000):
    integer products are bit-identical in both directions."""
    src = grgl_amd.grg_from_coalescent(1000, 10000, 1)
    built = _pg(_built_from(src))
    rng = np.random.default_rng(12)
    for direction, cols in ((UP, src.num_samples), (DOWN, src.num_mutations)):
        for K in (1, 16):
            x = rng.integers(-9, 10, size=(K, cols)).astype(np.int32)
            assert bits_equal(grgl_amd.matmul(built, x, direction), grgl_amd.matmul(src, x, direction)), (direction, K)


def test_matmul_bad(gpu):
    """test_grg.cpp:126-161 on the device path."""
    g = _pg(depth3_bin_tree([1, 5]))
    for x, d in ((np.zeros((1, 4)), DOWN), (np.zeros((1, 3)), UP), (np.zeros((9, 1)), UP)):
        with pytest.raises(RuntimeError):
            grgl_amd.matmul(g, x, d)


# --- oracle parity -----------------------------------------------------------------------

def _built_from(src):
    off, car = src.mutation_carriers()
    return grgl_amd.fast_grg_from_samples(src.num_samples, src.ploidy, carrier_offsets=off, carriers=car).arrays()


def _built_vcf():
    fx = vcf_fixture()
    return grgl_amd.fast_grg_from_samples(int(fx["num_samples"]), int(fx["ploidy"]), carrier_offsets=fx["carrier_off"],
                                          carriers=fx["carriers"], positions=fx["positions"]).arrays()


GRAPHS = {
    "sample8": lambda: sample8_grg(list(range(8, 14)) + [0, 3]),
    "random_heavy": lambda: random_grg(7, num_samples=300, num_internal=2000, max_children=8,
                                       heavy=(90, 700, 3000, 1500), num_mutations=3000),
    "tree_cfg1": lambda: grgl_amd.grg_from_coalescent(1000, 10000, 1).arrays(),
    "vcf": lambda: direct_map_grg(vcf_fixture()),
    # the same genotypes built by the haplotype-clustering constructor
    "built_cfg1": lambda: _built_from(grgl_amd.grg_from_coalescent(1000, 10000, 1)),
    "built_vcf": lambda: _built_vcf(),
}
_cache = {}


def _graph(name):
    if name not in _cache:
        arrays = GRAPHS[name]()
        _cache[name] = (arrays, _pg(arrays), oracle.OracleGraph(arrays))
    return _cache[name]


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("dtype", INT_DTYPES + FLOAT_DTYPES)
def test_parity_all_dtypes(gpu, name, dtype):
    arrays, g, og = _graph(name)
    rng = np.random.default_rng(abs(hash((name, np.dtype(dtype).name))) % 2**32)
    Ks = (1, 2, 3, 16, 33) if name != "sample8" else (1, 5)
    for direction in (UP, DOWN):
        cols = g.num_samples if direction == UP else g.num_mutations
        for K in Ks:
            x = make_input(rng, dtype, (K, cols))
            out = grgl_amd.matmul(g, x, direction)
            if np.dtype(dtype).kind in "biu":
                assert bits_equal(out, og.matmul(x, _odir(direction))), (direction, K)
            else:
                _check_float(out, x, og, direction, bar=FP32_BAR if name in ("tree_cfg1", "built_cfg1") else None)
                exact = grgl_amd.matmul(g, x, direction, order_preserving=True)
                assert bits_equal(exact, og.matmul(x, _odir(direction))), (direction, K)


@pytest.mark.parametrize("K", [64, 128, 256, 300, 512])
def test_parity_wide_rows(gpu, K):
    arrays, g, og = _graph("random_heavy")
    rng = np.random.default_rng(K)
    for dtype in (np.float32, np.int32, np.float64):
        for direction in (UP, DOWN):
            x = make_input(rng, dtype, (K, g.num_samples if direction == UP else g.num_mutations))
            out = grgl_amd.matmul(g, x, direction)
            if dtype == np.int32:
                assert bits_equal(out, og.matmul(x, _odir(direction)))
            else:
                _check_float(out, x, og, direction)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int64, np.int32, np.int8, np.uint16])
def test_parity_options(gpu, dtype):
    """by_individual, emit_all_nodes, init (vector/matrix/xtx) and miss against the oracle."""
    arrays = random_grg(13, num_samples=200, num_internal=900, heavy=(150, 400), num_mutations=1200,
                        with_miss=True, with_coal=True)
    g = _pg(arrays)
    og = oracle.OracleGraph(arrays)
    rng = np.random.default_rng(3)
    K = 4
    exactness = {"order_preserving": True} if np.dtype(dtype).kind == "f" else {}

    def same(a, b):
        assert bits_equal(a, b)

    for direction in (UP, DOWN):
        od = _odir(direction)
        cols = g.num_samples if direction == UP else g.num_mutations
        x = make_input(rng, dtype, (K, cols))
        xi = make_input(rng, dtype, (K, g.num_individuals if direction == UP else g.num_mutations))
        same(grgl_amd.matmul(g, xi, direction, by_individual=True, **exactness),
             og.matmul(xi, od, by_individual=True))
        same(grgl_amd.matmul(g, x, direction, emit_all_nodes=True, **exactness),
             og.matmul(x, od, emit_all_nodes=True))
        vec = make_input(rng, dtype, (K,))
        mat = make_input(rng, dtype, (K, g.num_nodes))
        for init in (vec, mat, "xtx"):
            same(grgl_amd.matmul(g, x, direction, init=init, **exactness), og.matmul(x, od, init=init))
        miss = make_input(rng, dtype, (K, g.num_mutations))
        m_gpu, m_cpu = miss.copy(), miss.copy()
        same(grgl_amd.matmul(g, x, direction, miss=m_gpu, **exactness), og.matmul(x, od, miss=m_cpu))
        same(m_gpu, m_cpu)


# --- the C ABI with device pointers: batching, determinism ----------------------------------

def _torch_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


_MISS_GRAPH = {}


def _miss_graph():
    if not _MISS_GRAPH:
        a = random_grg(21, num_samples=120, num_internal=700, heavy=(90, 300), num_mutations=900, with_miss=True)
        _MISS_GRAPH["g"] = (a, oracle.OracleGraph(a))
    return _MISS_GRAPH["g"]


@pytest.mark.parametrize("graph", ["random_heavy", "with_miss"])
@pytest.mark.parametrize("accumulate", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int8, np.uint16, np.bool_])
def test_batch_alone_equals_mixed(gpu, dtype, accumulate, graph):
    """Requests stacked into one traversal (one stacked seed and read-out launch) give the
    same bits as each request alone, with and without accumulation into the output."""
    import torch

    if graph == "random_heavy":
        arrays, _, og = _graph("random_heavy")
    else:
        arrays, og = _miss_graph()
    cg = capi.Graph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    flags = capi.FLAG_ACCUMULATE if accumulate else 0
    rng = np.random.default_rng(8)
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        Ks = [1, 8, 32, 3, 128, 1, 5]
        xs = [_torch_dev(make_input(rng, dtype, (k, cin))) for k in Ks]
        base = [_torch_dev(make_input(rng, dtype, (k, cout))) if accumulate else
                torch.zeros((k, cout), dtype=x.dtype, device="cuda") for k, x in zip(Ks, xs)]
        solo = []
        for k, x, b0 in zip(Ks, xs, base):
            out = b0.clone()
            cg.matmul_ptr(dt, x.data_ptr(), cin, k, direction, out.data_ptr(), out.numel(), flags=flags)
            solo.append(out)
        outs = [b0.clone() for b0 in base]
        cg.matmul_batch_ptr(dt, direction, [(k, x.data_ptr(), o.data_ptr()) for k, x, o in zip(Ks, xs, outs)],
                            flags=flags)
        torch.cuda.synchronize()
        for a, b in zip(solo, outs):
            assert bits_equal(a.cpu().numpy(), b.cpu().numpy())
        if np.dtype(dtype).kind in "iub" and not accumulate:
            for k, x, o in zip(Ks, xs, outs):
                assert bits_equal(o.cpu().numpy(), og.matmul(x.cpu().numpy(), _odir(UP if direction == capi.UP
                                                                                      else DOWN)))
    cg.close()


@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64])
def test_batch_host_buffers_and_order_preserving(gpu, dtype):
    """grgl_hip_matmul_batch with host buffers (FLAG_HOST_PTRS: the batch is staged through
    the workspace), with and without accumulation and order preservation, gives each request
    the bits of the same request alone; integers equal the oracle, and order-preserving floats
    equal the oracle's own float arithmetic."""
    arrays, _, og = _graph("random_heavy")
    cg = capi.Graph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    rng = np.random.default_rng(31)
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        Ks = [4, 1, 17, 64, 2]
        xs = [np.ascontiguousarray(make_input(rng, dtype, (k, cin))) for k in Ks]
        for flags in (0, capi.FLAG_ACCUMULATE, capi.FLAG_ORDER_PRESERVING):
            start = [make_input(rng, dtype, (k, cout)) if flags & capi.FLAG_ACCUMULATE else np.zeros((k, cout), dtype)
                     for k in Ks]
            outs = [np.ascontiguousarray(b.copy()) for b in start]
            cg.matmul_batch_ptr(dt, direction, [(k, x.ctypes.data, o.ctypes.data) for k, x, o in zip(Ks, xs, outs)],
                                flags=flags | capi.FLAG_HOST_PTRS)
            for k, x, b, o in zip(Ks, xs, start, outs):
                alone = np.ascontiguousarray(b.copy())
                cg.matmul_ptr(dt, x.ctypes.data, cin, k, direction, alone.ctypes.data, alone.size,
                              flags=flags | capi.FLAG_HOST_PTRS)
                assert bits_equal(o, alone), (direction, flags, k)
                if np.dtype(dtype).kind == "i" or flags & capi.FLAG_ORDER_PRESERVING:
                    want = og.matmul(x, _odir(UP if direction == capi.UP else DOWN))
                    if flags & capi.FLAG_ACCUMULATE:
                        want = want + b
                    assert bits_equal(o, want.astype(dtype)), (direction, flags, k)
    cg.close()


def test_deterministic_repeat(gpu):
    arrays, g, _ = _graph("random_heavy")
    x = make_input(np.random.default_rng(2), np.float32, (16, g.num_samples))
    a = grgl_amd.matmul(g, x, UP)
    b = grgl_amd.matmul(g, x, UP)
    assert bits_equal(a, b)


def test_device_pointer_accumulate(gpu):
    """FLAG_ACCUMULATE adds into the caller's output, like the reference's `+=`."""
    import torch

    arrays, _, og = _graph("tree_cfg1")
    cg = capi.Graph(arrays)
    x = make_input(np.random.default_rng(4), np.int64, (3, cg.num_mutations))
    prior = make_input(np.random.default_rng(5), np.int64, (3, cg.num_samples))
    out = _torch_dev(prior)
    xd = _torch_dev(x)
    cg.matmul_ptr(capi.I64, xd.data_ptr(), cg.num_mutations, 3, capi.DOWN, out.data_ptr(), out.numel(),
                  flags=capi.FLAG_ACCUMULATE)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), prior + og.matmul(x, oracle.DOWN))
    cg.close()


# --- full-size graphs ----------------------------------------------------------------------

def test_config2_scale_parity(gpu):
    """100,000 haplotypes x 500,000 sites, 20 trees: K=1 and 16, FP64 and FP32, both directions."""
    g = grgl_amd.grg_from_coalescent(100_000, 500_000, 20)
    og = oracle.OracleGraph(g.arrays())
    rng = np.random.default_rng(20261203)
    for K in (1, 16):
        for direction in (UP, DOWN):
            cols = g.num_samples if direction == UP else g.num_mutations
            x64 = rng.standard_normal((K, cols))
            out64 = grgl_amd.matmul(g, x64, direction)
            ref64 = og.matmul(x64, _odir(direction))
            assert normwise_rel_err(out64, ref64) <= 1e-12
            x32 = x64.astype(np.float32)
            rep = _check_float(grgl_amd.matmul(g, x32, direction), x32, og, direction, bar=FP32_BAR)
            print("fp32 100k x 500k K=%d %s: normwise %.3e, bound ratio %.3f" %
                  (K, "UP" if direction == UP else "DOWN", rep["normwise"], rep["bound_ratio"]))
    ones = np.ones((1, g.num_samples), np.int32)
    assert bits_equal(grgl_amd.matmul(g, ones, UP), og.matmul(ones, oracle.UP))


def test_config3_scale_properties(gpu):
    """200,000 x 1,000,000, 50 trees: int32 bit-exact vs oracle, adjoint identity in int64."""
    g = grgl_amd.grg_from_coalescent(200_000, 1_000_000, 50)
    og = oracle.OracleGraph(g.arrays())
    rng = np.random.default_rng(1)
    mask = (rng.random((4, g.num_samples)) < 0.5).astype(np.int32)
    assert bits_equal(grgl_amd.matmul(g, mask, UP), og.matmul(mask, oracle.UP))
    wmask = (rng.random((4, g.num_mutations)) < 0.5).astype(np.int32)
    assert bits_equal(grgl_amd.matmul(g, wmask, DOWN), og.matmul(wmask, oracle.DOWN))
    v = rng.integers(-100, 100, (2, g.num_samples)).astype(np.int64)
    w = rng.integers(-100, 100, (2, g.num_mutations)).astype(np.int64)
    lhs = int((grgl_amd.matmul(g, v, UP) * w).sum())
    rhs = int((v * grgl_amd.matmul(g, w, DOWN)).sum())
    assert lhs == rhs


def test_cpp_shim_reference_unit_tests(gpu, tmp_path):
    """The reference's C++ matmul unit tests through include/grgl_hip.hpp on the device."""
    import subprocess

    from test_host import _build_shim_test

    import os

    exe = _build_shim_test(tmp_path)
    env = dict(os.environ, GRGL_SHIM_LVLZ=str(tmp_path / "shim.lvlz"))
    proc = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120,
                          env=env)
    assert proc.returncode == 0, proc.stdout
