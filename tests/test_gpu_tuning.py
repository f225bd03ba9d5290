"""Tuning never changes results: every team width, hipGraph replay on/off and both level-kernel
block orders give bit-identical outputs (the per-column arithmetic depends only on the graph
and the heavy threshold).  Captured level sequences are replayed several times, with eager
work in between, and every replay is checked."""

import numpy as np
import pytest

import oracle
from grgl_amd import capi
from helpers import bits_equal, make_input, random_grg

pytestmark = pytest.mark.gpu

_G = {}


def _graphs():
    if not _G:
        import grgl_amd

        _G["heavy"] = random_grg(31, num_samples=300, num_internal=2500, max_children=9, heavy=(100, 900, 2500),
                                 num_mutations=2000)
        _G["tree"] = grgl_amd.grg_from_coalescent(3000, 30000, 3, seed=5).arrays()
    return _G


def _team_widths(K, es):
    """Every team width the library accepts for K columns of es bytes: powers of two up to
    the row's 16-byte slices rounded up, and at least up to 8 (narrower accesses, then idle
    lanes, on short rows)."""
    raw = K * es
    lanes = -(-raw // 16) if raw > 8 else 1
    top = min(64, max(8, 1 << (lanes - 1).bit_length()))
    return [t for t in (1, 2, 4, 8, 16, 32, 64) if t <= top]


def _run(cg, dt, x, K, direction, cin, cout, stream):
    import torch

    y = torch.zeros((K, cout), dtype=x.dtype, device="cuda")
    cg.matmul_ptr(dt, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), stream=stream)
    return y


@pytest.mark.parametrize("name", ["heavy", "tree"])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int8])
def test_every_tuning_gives_the_same_bits(gpu, name, dtype):
    import torch

    arrays = _graphs()[name]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    es = np.dtype(dtype).itemsize
    rng = np.random.default_rng(11)
    stream = torch.cuda.current_stream().cuda_stream
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        for K in (1, 2, 3, 4, 5, 64, 200, 511):
            x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()
            cg.set_tuning(direction, dt, K)  # library default
            ref = _run(cg, dt, x, K, direction, cin, cout, stream).cpu().numpy()
            if np.dtype(dtype).kind == "i":
                assert bits_equal(ref, og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN))
            tws = _team_widths(K, es)
            for tw in tws:
                for use_graph in (0, 1):
                    for xcd in (1, 2):
                        cg.set_tuning(direction, dt, K, team_lanes=tw, use_graph=use_graph, xcd_map=xcd)
                        for rep in range(3 if use_graph else 1):
                            got = _run(cg, dt, x, K, direction, cin, cout, stream)
                            torch.cuda.synchronize()
                            # eager work between replays
                            _ = (x.float() * 2).sum().item()
                            assert bits_equal(got.cpu().numpy(), ref), (name, direction, K, tw, use_graph, xcd, rep)
            cg.set_tuning(direction, dt, K)
    cg.close()


def test_autotune_choice_is_used_and_exact(gpu):
    import torch

    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(2)
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        chosen, ms = cg.autotune(direction, capi.I32, 300)
        assert chosen["team_lanes"] in (1, 2, 4, 8, 16, 32, 64) and ms > 0
        assert chosen["xcd_map"] in (1, 2) and chosen["use_graph"] in (0, 1)
        x = torch.from_numpy(make_input(rng, np.int32, (300, cin))).cuda()
        got = _run(cg, capi.I32, x, 300, direction, cin, cout, stream).cpu().numpy()
        assert bits_equal(got, og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN))
    with pytest.raises(RuntimeError):
        cg.set_tuning(capi.UP, capi.I32, 8, xcd_map=7)
    cg.close()


@pytest.mark.parametrize("name", ["heavy", "tree"])
def test_split_levels_give_the_same_bits(gpu, name):
    """Levels cut into several launches (max_launch_entries; the library does this for levels
    of more than 2^30 gather entries, whose list offsets are then read relative to each
    launch's first entry) give the same bits, with and without hipGraph replay."""
    import torch

    arrays = _graphs()[name]
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    rng = np.random.default_rng(12)
    stream = torch.cuda.current_stream().cuda_stream
    for dtype in (np.float32, np.int32, np.float64):
        dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
        for direction in (capi.UP, capi.DOWN):
            cin = cg.num_samples if direction == capi.UP else cg.num_mutations
            cout = cg.num_mutations if direction == capi.UP else cg.num_samples
            for K in (1, 3, 40):
                x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()
                cg.set_tuning(direction, dt, K, max_launch_entries=1 << 30)
                ref = _run(cg, dt, x, K, direction, cin, cout, stream).cpu().numpy()
                if np.dtype(dtype).kind == "i":
                    assert bits_equal(ref, og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN))
                for chunk in (1, 7, 100, 5000):
                    for use_graph in (0, 1):
                        cg.set_tuning(direction, dt, K, use_graph=use_graph, max_launch_entries=chunk)
                        for rep in range(2 if use_graph else 1):
                            got = _run(cg, dt, x, K, direction, cin, cout, stream)
                            torch.cuda.synchronize()
                            assert bits_equal(got.cpu().numpy(), ref), (direction, K, chunk, use_graph, rep)
                cg.set_tuning(direction, dt, K, max_launch_entries=1 << 30)
    cg.close()


def test_tuning_cache_file(gpu, tmp_path):
    """Autotuned choices saved to a file and loaded into a new handle of the same graph give
    the same tuning (and results); a graph of another structure refuses the file."""
    import torch

    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    chosen_up, _ = cg.autotune(capi.UP, capi.F32, 300)
    chosen_down, _ = cg.autotune(capi.DOWN, capi.I32, 7)
    cg.set_tuning(capi.DOWN, capi.F64, 40, team_lanes=2, use_graph=1, heavy_fanin=100, max_launch_entries=5000)
    path = str(tmp_path / "tree.tuning")
    cg.save_tuning(path)
    want = {key: cg.get_tuning(*key) for key in ((capi.UP, capi.F32, 300), (capi.DOWN, capi.I32, 7),
                                                 (capi.DOWN, capi.F64, 40))}
    assert want[(capi.UP, capi.F32, 300)]["team_lanes"] == chosen_up["team_lanes"]
    assert want[(capi.DOWN, capi.F64, 40)]["heavy_fanin"] == 100
    assert want[(capi.DOWN, capi.F64, 40)]["max_launch_entries"] == 5000

    fresh = capi.Graph(arrays)
    assert fresh.get_tuning(capi.DOWN, capi.F64, 40)["heavy_fanin"] == 64
    fresh.load_tuning(path)
    for key, t in want.items():
        assert fresh.get_tuning(*key) == t, key
    # same results on both handles after loading
    rng = np.random.default_rng(4)
    stream = torch.cuda.current_stream().cuda_stream
    x = torch.from_numpy(make_input(rng, np.float64, (40, cg.num_mutations))).cuda()
    a = _run(cg, capi.F64, x, 40, capi.DOWN, cg.num_mutations, cg.num_samples, stream)
    b = _run(fresh, capi.F64, x, 40, capi.DOWN, cg.num_mutations, cg.num_samples, stream)
    torch.cuda.synchronize()
    assert bits_equal(a.cpu().numpy(), b.cpu().numpy())

    other = capi.Graph(_graphs()["heavy"])
    with pytest.raises(RuntimeError):
        other.load_tuning(path)
    bad = tmp_path / "bad.tuning"
    bad.write_text("grgl_hip_tuning 1\ngraph 0000\n")
    with pytest.raises(RuntimeError):
        fresh.load_tuning(str(bad))
    for h in (cg, fresh, other):
        h.close()


def test_graph_cache_threads_past_its_bound(gpu):
    """Captured level sequences from several threads at once, over more keys than the cache
    keeps (64), while another thread swaps the DOWN launch table (max_launch_entries, which
    never changes results): every replay equals the direct launches, bit for bit.  Entries are
    evicted only when no traversal holds them and their last launch has completed."""
    import concurrent.futures as cf

    import torch

    arrays = _graphs()["tree"]
    cg = capi.Graph(arrays)
    rng = np.random.default_rng(21)
    Ks = list(range(1, 91))
    cases = []
    for i, K in enumerate(Ks):
        dtype = np.float32 if i % 2 == 0 else np.int32
        direction = capi.UP if i % 3 else capi.DOWN
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        cases.append((K, dtype, direction, cin, cout, torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()))
    stream = torch.cuda.current_stream().cuda_stream
    refs = []
    for K, dtype, direction, cin, cout, x in cases:
        refs.append(_run(cg, capi.NUMPY_TO_DTYPE[np.dtype(dtype)], x, K, direction, cin, cout, stream).cpu().numpy())
    for K, dtype, direction, _, _, _ in cases:
        cg.set_tuning(direction, capi.NUMPY_TO_DTYPE[np.dtype(dtype)], K, use_graph=1)
    stop = [False]

    def retune():
        n = 0
        while not stop[0]:
            cg.set_tuning(capi.DOWN, capi.I8, 1, max_launch_entries=(5000 if n % 2 else 1 << 30))
            n += 1
        return n

    def worker(t):
        s = torch.cuda.Stream()
        bad = []
        with torch.cuda.stream(s):
            for rep in range(3):
                for i in range(t, len(cases), 4):
                    K, dtype, direction, cin, cout, x = cases[i]
                    y = torch.zeros((K, cout), dtype=x.dtype, device="cuda")
                    cg.matmul_ptr(capi.NUMPY_TO_DTYPE[np.dtype(dtype)], x.data_ptr(), cin, K, direction,
                                  y.data_ptr(), y.numel(), stream=s.cuda_stream)
                    s.synchronize()
                    if not bits_equal(y.cpu().numpy(), refs[i]):
                        bad.append((i, rep))
        return bad

    with cf.ThreadPoolExecutor(5) as ex:
        tuner = ex.submit(retune)
        try:
            futs = [ex.submit(worker, t) for t in range(4)]
            bad = [b for f in futs for b in f.result()]
        finally:
            stop[0] = True  # the retuning thread ends whatever the workers did
        assert tuner.result() > 0
    assert bad == []
    cg.close()


@pytest.mark.parametrize("name", ["heavy", "tree"])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int64])
def test_tile_shapes_give_the_same_bits(gpu, name, dtype):
    """Wide-team level kernels with 256 or 512 threads per block and 4 or 8 row loads in flight
    (direct and graph launches) give the bits of the default shape; get_tuning reports the
    shape that runs."""
    import torch

    arrays = _graphs()[name]
    cg = capi.Graph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    rng = np.random.default_rng(17)
    stream = torch.cuda.current_stream().cuda_stream
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        for K in (40, 200, 511):
            x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()
            cg.set_tuning(direction, dt, K)
            ref = _run(cg, dt, x, K, direction, cin, cout, stream).cpu().numpy()
            for tw in (8, 16, 32):
                for blk in (256, 512):
                    for unr in (4, 8):
                        for use_graph in ((0, 1) if (blk, unr) == (512, 8) else (0,)):
                            # One launch per level (block_threads reports the level kernels').
                            cg.set_tuning(direction, dt, K, team_lanes=tw, block_threads=blk, loads_in_flight=unr,
                                          use_graph=use_graph, chain=1)
                            t = cg.get_tuning(direction, dt, K)
                            if t["team_lanes"] == tw and K * np.dtype(dtype).itemsize >= 16 * tw:
                                assert (t["block_threads"], t["loads_in_flight"]) == (blk, unr), t
                            got = _run(cg, dt, x, K, direction, cin, cout, stream)
                            torch.cuda.synchronize()
                            assert bits_equal(got.cpu().numpy(), ref), (name, direction, K, tw, blk, unr, use_graph)
            cg.set_tuning(direction, dt, K)
    with pytest.raises(RuntimeError):
        cg.set_tuning(capi.UP, dt, 8, block_threads=384)
    cg.close()
