"""Chained traversals (grgl_tuning::chain = 2: every level of a direction in one launch,
blocks taking the levels' work in order and waiting on per-level completion counters):
bit-identical to the level-by-level launches for every eligible dtype and row width, with and
without node initialisation, in batches and with host buffers; the host-set abort word makes
the kernel give up and return, the graph is marked, calls fail until it is reset.

The abort is forced from the host before the launch (blocks read the word when they start),
so every kernel still exits normally."""

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import capi
from helpers import bits_equal, csr_from_lists, make_input, random_grg

pytestmark = pytest.mark.gpu

_G = {}


def _graphs():
    if not _G:
        _G["tree"] = grgl_amd.grg_from_coalescent(3000, 30000, 3, seed=5).arrays()
        _G["heavy"] = random_grg(31, num_samples=300, num_internal=2500, max_children=9, heavy=(100, 900, 2500),
                                 num_mutations=2000)
        # Levels of more than 512 heavy rows (one heavy row per wavefront): 900 nodes of 70-100
        # sample children (UP level 1), so every sample has about 125 parents (DOWN).
        rng = np.random.default_rng(41)
        N = 600
        kids = [[] for _ in range(N)]
        for i in range(900):
            kids.append(sorted(rng.choice(N, size=70 + i % 31, replace=False).tolist()))
        kids.append(list(range(N, N + 900, 7)))
        muts = rng.integers(N, len(kids), size=1500)
        _G["wide"] = csr_from_lists(N, kids, muts)
    return _G


def _run(cg, dt, x, K, direction, cin, cout, stream, **kw):
    import torch

    y = torch.zeros((K, cout), dtype=x.dtype, device="cuda")
    cg.matmul_ptr(dt, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), stream=stream, **kw)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("name", ["tree", "heavy", "wide"])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int64, np.uint32])
def test_chained_equals_level_launches(gpu, name, dtype):
    import torch

    arrays = _graphs()[name]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    es = np.dtype(dtype).itemsize
    rng = np.random.default_rng(3)
    sp = torch.cuda.current_stream().cuda_stream
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        for K in [k for k in (1, 2, 3, 4, 5, 8, 12, 16) if k * es <= 64]:
            x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()
            cg.set_tuning(direction, dt, K, chain=1)
            ref = _run(cg, dt, x, K, direction, cin, cout, sp)
            if np.dtype(dtype).kind in "iu":
                assert bits_equal(ref, og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN))
            tws = [t for t in (1, 2, 4, 8)]
            for tw in tws:
                cg.set_tuning(direction, dt, K, team_lanes=tw, chain=2)
                for rep in range(2):  # the state words are reused
                    got = _run(cg, dt, x, K, direction, cin, cout, sp)
                    assert bits_equal(got, ref), (name, direction, K, tw, rep)
            # node initialisation (rows start from their init value) and order preservation
            vec = torch.from_numpy(make_input(rng, dtype, (K,))).cuda()
            for kw in (dict(init_ptr=vec.data_ptr(), init_mode=capi.INIT_VECTOR),
                       dict(flags=capi.FLAG_ORDER_PRESERVING)):
                cg.set_tuning(direction, dt, K, chain=1)
                want = _run(cg, dt, x, K, direction, cin, cout, sp, **kw)
                cg.set_tuning(direction, dt, K, chain=2)
                assert bits_equal(_run(cg, dt, x, K, direction, cin, cout, sp, **kw), want), (name, direction, K, kw)
            cg.set_tuning(direction, dt, K)
    assert cg.chain_control() == 0
    cg.close()


def test_chained_batch_and_host_buffers(gpu):
    """Stacked requests (one chained traversal for all) and host buffers (staged, synchronous)."""
    import torch

    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    rng = np.random.default_rng(8)
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        Ks = [1, 2, 5, 3]  # 11 stacked rows of 4 bytes (44-byte rows)
        xs = [torch.from_numpy(make_input(rng, np.int32, (k, cin))).cuda() for k in Ks]
        for chain in (1, 2):
            for k in range(1, 65):
                cg.set_tuning(direction, capi.I32, k, chain=chain)
            outs = [torch.zeros((k, cout), dtype=torch.int32, device="cuda") for k in Ks]
            cg.matmul_batch_ptr(capi.I32, direction, [(k, x.data_ptr(), o.data_ptr()) for k, x, o in zip(Ks, xs, outs)],
                                stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            for x, o in zip(xs, outs):
                want = og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN)
                assert bits_equal(o.cpu().numpy(), want), (direction, chain)
            xh = make_input(rng, np.float32, (4, cin))
            cg.set_tuning(direction, capi.F32, 4, chain=1)
            a = cg.matmul_host(xh, direction)
            cg.set_tuning(direction, capi.F32, 4, chain=2)
            assert bits_equal(cg.matmul_host(xh, direction), a)
    cg.close()


def test_chained_abort_word(gpu):
    """The host sets the abort word: the chained kernel gives up at once and returns, the graph is
    marked, the next call fails; after the reset calls work and are exact again."""
    import torch

    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    sp = torch.cuda.current_stream().cuda_stream
    cin, cout = cg.num_mutations, cg.num_samples
    x = torch.from_numpy(make_input(np.random.default_rng(1), np.int32, (4, cin))).cuda()
    cg.set_tuning(capi.DOWN, capi.I32, 4, chain=2)
    assert cg.chain_control(abort=1) == 0
    y = torch.zeros((4, cout), dtype=torch.int32, device="cuda")
    cg.matmul_ptr(capi.I32, x.data_ptr(), cin, 4, capi.DOWN, y.data_ptr(), y.numel(), stream=sp)
    torch.cuda.synchronize()  # the kernel returned
    assert cg.chain_control() == 1
    with pytest.raises(capi.GrglError, match="gave up"):
        cg.matmul_ptr(capi.I32, x.data_ptr(), cin, 4, capi.DOWN, y.data_ptr(), y.numel(), stream=sp)
    assert cg.chain_control(abort=0, reset=True) == 1
    assert cg.chain_control() == 0
    y.zero_()
    cg.matmul_ptr(capi.I32, x.data_ptr(), cin, 4, capi.DOWN, y.data_ptr(), y.numel(), stream=sp)
    torch.cuda.synchronize()
    assert bits_equal(y.cpu().numpy(), og.matmul(x.cpu().numpy(), oracle.DOWN))
    assert cg.chain_control() == 0
    cg.close()
