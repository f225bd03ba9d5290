"""The benchmarked paths at their full workload sizes, against the oracle.

* The 100,000 haplotypes x 500,000 sites workload (20 segment trees) built by the
  haplotype-clustering constructor (fast_grg_from_samples, saved and reloaded as its .grg file
  holds it): K = 1 and 16, UP and DOWN; int32 bit-exact against the oracle and equal to the
  tree-direct graph's products (one genotype matrix), FP64 <= 1e-12 and FP32 <= 1e-6 normwise
  against the FP64 oracle, FP32 within the rounding bound.
* Batched right-hand sides, 200,000 haplotypes x 1,000,000 sites (50 segment trees): the K
  ladder {1, 2, ..., 256} stacked as 511 rows in ONE traversal (grgl_hip_matmul_batch), with
  the tuning the autotuner picks for 511 rows, f32 N(0,1) and int32 Bernoulli(0.5), UP and DOWN.
  int32: every request bit-exact.  f32: every request <= 1e-6 normwise per row against the
  FP64 oracle, and the K = 1, 2, 4 requests within the a priori rounding bound.
* Standardized PCA, 50,000 diploids (100,000 haplotypes) x 300,000 sites (15 trees): top-10
  eigenvalues and eigenvectors against the FP64 CPU oracle with the same random block.
* Serving on the three resident graphs (10k x 50k, 100k x 500k, 200k x 1M) with the mixed
  request stream (K in {1, 8, 32, 128} p .4/.3/.2/.1, random direction, 70% f32 / 30% int32),
  each request with its own device output: every result equals the same request run alone,
  bit for bit, and every int32 result equals the oracle.
"""

import concurrent.futures as cf

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import capi
from helpers import bits_equal
from oracle import fp_bounds

pytestmark = pytest.mark.gpu

SEED = 20261203
K_LADDER = [1, 2, 4, 8, 16, 32, 64, 128, 256]
_G = {}


def _odir(d):
    return oracle.UP if d == capi.UP else oracle.DOWN


def oracle_rows(og, x, direction, threads=8):
    """The oracle on row blocks of `x` in parallel threads (rows are independent)."""
    K = x.shape[0]
    step = max(1, -(-K // threads))
    blocks = [(k, min(K, k + step)) for k in range(0, K, step)]
    with cf.ThreadPoolExecutor(len(blocks)) as ex:
        parts = list(ex.map(lambda b: og.matmul(np.ascontiguousarray(x[b[0]:b[1]]), direction), blocks))
    return np.concatenate(parts, axis=0)


def _config3():
    if "c3" not in _G:
        g = grgl_amd.grg_from_coalescent(200_000, 1_000_000, 50, seed=SEED)
        a = g.arrays()
        _G["c3"] = (a, capi.Graph(a), oracle.OracleGraph(a))
    return _G["c3"]


def test_config2_constructed_graph(gpu, tmp_path):
    import os

    src = grgl_amd.grg_from_coalescent(100_000, 500_000, 20, seed=SEED)
    off, car = src.mutation_carriers()
    built = grgl_amd.fast_grg_from_samples(100_000, 2, carrier_offsets=off, carriers=car,
                                           positions=src.arrays()["positions"])
    del off, car
    fn = os.path.join(str(tmp_path), "c2.grg")
    grgl_amd.save_grg(built, fn)
    del built
    g = grgl_amd.load_immutable_grg(fn)
    assert g.num_mutations == src.num_mutations
    og = oracle.OracleGraph(g.arrays())
    rng = np.random.default_rng(SEED + 11)
    for d, cols in ((grgl_amd.UP, g.num_samples), (grgl_amd.DOWN, g.num_mutations)):
        od = oracle.UP if d == grgl_amd.UP else oracle.DOWN
        for K in (1, 16, 40):  # 40: 160-byte rows, grouped as 64-byte column slices UP by default
            xi = rng.integers(-9, 10, size=(K, cols)).astype(np.int32)
            yi = grgl_amd.matmul(g, xi, d)
            assert bits_equal(yi, og.matmul(xi, od)), (d, K)
            assert bits_equal(yi, grgl_amd.matmul(src, xi, d)), (d, K)
            x64 = rng.standard_normal((K, cols))
            assert fp_bounds.normwise_rel_err(grgl_amd.matmul(g, x64, d), og.matmul(x64, od)) <= 1e-12, (d, K)
            x32 = x64.astype(np.float32)
            rep = fp_bounds.fp32_report(og, x32, od, grgl_amd.matmul(g, x32, d), with_reference=False)
            assert rep["bound_ratio"] <= 1.0 and rep["normwise"] <= 1e-6, (d, K, rep)


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_config3_batched_ladder(gpu, dtype):
    import torch

    a, cg, og = _config3()
    cdt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    rng = np.random.default_rng(SEED + (0 if dtype == np.int32 else 1))
    Ktot = sum(K_LADDER)
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        if dtype == np.int32:
            x = (rng.random((Ktot, cin)) < 0.5).astype(np.int32)
        else:
            x = rng.standard_normal((Ktot, cin)).astype(np.float32)
        cg.autotune(direction, cdt, Ktot)
        xd = torch.from_numpy(x).cuda()
        yd = torch.zeros((Ktot, cout), dtype=xd.dtype, device="cuda")
        reqs, k0 = [], 0
        for K in K_LADDER:
            reqs.append((K, xd[k0:k0 + K].data_ptr(), yd[k0:k0 + K].data_ptr()))
            k0 += K
        cg.matmul_batch_ptr(cdt, direction, reqs, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        if dtype == np.int32:
            want = oracle_rows(og, x, _odir(direction))
            assert bits_equal(y, want), direction
        else:
            exact = oracle_rows(og, x.astype(np.float64), _odir(direction))
            k0 = 0
            for K in K_LADDER:
                err = fp_bounds.normwise_rel_err(y[k0:k0 + K], exact[k0:k0 + K])
                assert err <= 1e-6, (direction, K, err)
                print("fp32 200k x 1M batched %s K=%d: normwise %.3e" % ("UP" if direction == capi.UP else "DOWN",
                                                                        K, err))
                k0 += K
            rep = fp_bounds.fp32_report(og, x[:7], _odir(direction), y[:7], with_reference=False)
            assert rep["bound_ratio"] <= 1.0, rep
        cg.set_tuning(direction, cdt, Ktot)


@pytest.mark.parametrize("K", [1, 40])
def test_config3_grouped_levels(gpu, K):
    """Grouped levels at the benchmark graph's size (K = 40: 160-byte rows as 64-byte column
    slices): int32 equal to the oracle, FP32 bit-identical to one launch per level."""
    import torch

    a, cg, og = _config3()
    rng = np.random.default_rng(SEED + 3 + K)
    sp = torch.cuda.current_stream().cuda_stream
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        for dt, x in ((capi.I32, rng.integers(-9, 10, size=(K, cin)).astype(np.int32)),
                      (capi.F32, rng.standard_normal((K, cin)).astype(np.float32))):
            xd = torch.from_numpy(x).cuda()
            outs = []
            for chain in (3, 1):
                print(f"config3 grouped: K={K} dir={direction} dtype={dt} chain={chain}", flush=True)
                cg.set_tuning(direction, dt, K, chain=chain)
                assert cg.get_tuning(direction, dt, K)["chain"] == chain
                yd = torch.zeros((K, cout), dtype=xd.dtype, device="cuda")
                cg.matmul_ptr(dt, xd.data_ptr(), cin, K, direction, yd.data_ptr(), yd.numel(), stream=sp)
                torch.cuda.synchronize()
                outs.append(yd.cpu().numpy())
            cg.set_tuning(direction, dt, K)
            assert bits_equal(outs[0], outs[1]), (K, direction, dt)
            if dt == capi.I32:
                assert bits_equal(outs[0], oracle_rows(og, x, _odir(direction))), (K, direction)


def test_config4_pca_full_size(gpu):
    from grgl_amd.pca import DeviceStdOperator, eigen_errors, randomized_pca
    from oracle.pca_oracle import OracleStdOperator
    from oracle.pca_oracle import randomized_pca as oracle_pca

    a = grgl_amd.grg_from_coalescent(100_000, 300_000, 15, seed=SEED + 4).arrays()
    assert a["num_samples"] == 100_000 and len(a["mutation_nodes"]) == 300_000
    cg = capi.Graph(a)
    op = DeviceStdOperator(cg)
    assert op.shape == (50_000, 300_000)
    ref = OracleStdOperator(a)
    np.testing.assert_allclose(op.freq.cpu().numpy(), ref.freq, rtol=1e-15, atol=0)
    e, U = randomized_pca(op, k=10)
    omega = op.randn(op.shape[1], 20, 20261203).cpu().numpy()
    re, rU = oracle_pca(ref, k=10, omega=omega)
    rel, ang = eigen_errors(e, U, re, rU)
    print("pca 50k x 300k: max eigenvalue rel err %.3e, max eigenvector angle %.3e rad" % (rel.max(), ang.max()))
    assert rel.max() < 1e-6, rel
    assert ang.max() < 1e-4, ang
    cg.close()


SERVE_GRAPHS = [(10_000, 50_000, 10), (100_000, 500_000, 20), (200_000, 1_000_000, 50)]


def test_config5_serving_full_size(gpu):
    import torch

    graphs = []
    for n, m, t in SERVE_GRAPHS:
        a = grgl_amd.grg_from_coalescent(n, m, t, seed=SEED + n).arrays()
        graphs.append((a, capi.Graph(a), oracle.OracleGraph(a)))
    srv = capi.Server(num_streams=2, max_batch_rows=512, hbm_budget_bytes=48 << 30)
    gids = [srv.add_graph(cg) for _, cg, _ in graphs]
    rng = np.random.default_rng(SEED + 5)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 5)
    reqs = []
    for _ in range(160):
        gi = int(rng.integers(0, 3))
        K = int(rng.choice([1, 8, 32, 128], p=[0.4, 0.3, 0.2, 0.1]))
        d = capi.UP if rng.random() < 0.5 else capi.DOWN
        i32 = rng.random() >= 0.7
        cg = graphs[gi][1]
        cin, cout = (cg.num_samples, cg.num_mutations) if d == capi.UP else (cg.num_mutations, cg.num_samples)
        if i32:
            x = (torch.rand((K, cin), device="cuda", generator=gen) < 0.5).to(torch.int32)
        else:
            x = torch.randn((K, cin), device="cuda", generator=gen)
        y = torch.zeros((K, cout), dtype=x.dtype, device="cuda")
        reqs.append((gi, K, d, capi.I32 if i32 else capi.F32, x, y, cin, cout))
    torch.cuda.synchronize()
    ids = [srv.submit_device(gids[gi], cdt, d, K, x.data_ptr(), y.data_ptr())
           for gi, K, d, cdt, x, y, _, _ in reqs]
    for rid in ids:
        assert srv.wait(rid) >= 0.0
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] < len(reqs)
    sp = torch.cuda.current_stream().cuda_stream
    by_key = {}
    for i, (gi, K, d, cdt, x, y, cin, cout) in enumerate(reqs):
        alone = torch.zeros_like(y)
        graphs[gi][1].matmul_ptr(cdt, x.data_ptr(), cin, K, d, alone.data_ptr(), alone.numel(), stream=sp)
        torch.cuda.synchronize()
        assert torch.equal(alone.view(torch.int32), y.view(torch.int32)), ("served != alone", i, gi, K, d, cdt)
        if cdt == capi.I32:
            by_key.setdefault((gi, d), []).append(i)
    for (gi, d), idx in by_key.items():
        x = np.concatenate([reqs[i][4].cpu().numpy() for i in idx], axis=0)
        y = np.concatenate([reqs[i][5].cpu().numpy() for i in idx], axis=0)
        assert bits_equal(y, oracle_rows(graphs[gi][2], x, _odir(d))), (gi, d)
    srv.close()
    for _, cg, _ in graphs:
        cg.close()
