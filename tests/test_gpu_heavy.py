"""Heavy rows (more than 64 gather entries) have one arithmetic however they are launched: 64
interleaved partials (entry i goes to partial i mod 64), each summed in list order, then a fixed
binary tree.  A launch with at least 512 heavy rows and rows of at most 16 bytes (32 bytes in
heavy-dominated levels) gives each heavy row one wavefront whose lanes are the partials; other
launches give each heavy row a whole block.  The graph here has levels of 600-800 heavy rows in
both directions, so the default launch takes the wavefront path; cutting each level into
launches of at most 20,000 gather entries leaves fewer than 512 heavy rows per launch and takes
the block path.  Both must give the same bits, for every team width; integers also equal the
oracle."""

import numpy as np
import pytest

import oracle
from grgl_amd import capi
from helpers import INVALID, bits_equal, csr_from_lists, make_input

_G = {}


def _wide_graph():
    """3,000 samples; 800 nodes of 65-400 sample children (height 1); 300 nodes of 2-6
    children among the samples and the first group, at least one from the first group (height
    2); 600 nodes of 65-200 children among the nodes before them, at least one from the second
    group (height 3); 4,000 mutations (5% unmapped)."""
    if "arrays" not in _G:
        rng = np.random.default_rng(2026)
        N = 3000
        kids = [[] for _ in range(N)]
        g1 = []
        for _ in range(800):
            g1.append(len(kids))
            kids.append(rng.choice(N, size=int(rng.integers(65, 401)), replace=False).tolist())
        g2 = []
        first = len(kids)
        for _ in range(300):
            v = len(kids)
            f = int(rng.integers(2, 7))
            c = set(rng.choice(first, size=f - 1, replace=False).tolist())
            c.add(int(rng.choice(g1)))
            g2.append(v)
            kids.append(sorted(c))
        below = len(kids)
        for _ in range(600):
            f = int(rng.integers(65, 201))
            c = set(rng.choice(below, size=f - 1, replace=False).tolist())
            c.add(int(rng.choice(g2)))
            kids.append(sorted(c))
        V = len(kids)
        muts = rng.integers(0, V, size=4000).astype(np.int64)
        muts[rng.random(4000) < 0.05] = INVALID
        _G["arrays"] = csr_from_lists(N, kids, muts)
    return _G["arrays"]


def _heavy_per_level(arrays):
    """Rows of more than 64 gather entries in each topological level, per direction: UP rows
    are nodes by height above the samples with their children as entries; DOWN rows are nodes
    by depth below the roots with their parents and mutations as entries."""
    N = arrays["num_samples"]
    off = arrays["down_offsets"].astype(np.int64)
    ch = arrays["down_children"].astype(np.int64)
    V = len(off) - 1
    height = np.zeros(V, np.int64)
    for n in range(N, V):
        c = ch[off[n]:off[n + 1]]
        height[n] = 1 + (height[c].max() if len(c) else -1)
    parents = np.bincount(ch, minlength=V)
    muts = arrays["mutation_nodes"].astype(np.int64)
    seeds = np.bincount(muts[muts != INVALID], minlength=V)
    depth = np.zeros(V, np.int64)
    for n in range(V - 1, -1, -1):
        c = ch[off[n]:off[n + 1]]
        depth[c] = np.maximum(depth[c], depth[n] + 1)
    fan_up = off[1:] - off[:-1]
    up = np.bincount(height[fan_up > 64], minlength=height.max() + 1)
    down = np.bincount(depth[(parents + seeds) > 64], minlength=depth.max() + 1)
    return up, down


def test_wide_graph_has_levels_of_many_heavy_rows():
    up, down = _heavy_per_level(_wide_graph())
    assert (up >= 512).sum() >= 2, up
    assert (down >= 512).any(), down


def _team_widths(K, es):
    raw = K * es
    lanes = -(-raw // 16) if raw > 8 else 1
    top = min(64, max(8, 1 << (lanes - 1).bit_length()))
    return [t for t in (1, 2, 4, 8, 16, 32, 64) if t <= top]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
def test_heavy_rows_same_bits_per_wavefront_and_per_block(gpu, dtype):
    import torch

    arrays = _wide_graph()
    cg = capi.Graph(arrays)
    og = oracle.OracleGraph(arrays)
    dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
    es = np.dtype(dtype).itemsize
    rng = np.random.default_rng(17)
    stream = torch.cuda.current_stream().cuda_stream
    for direction in (capi.UP, capi.DOWN):
        cin = cg.num_samples if direction == capi.UP else cg.num_mutations
        cout = cg.num_mutations if direction == capi.UP else cg.num_samples
        # row bytes 4..32 (wavefront path by default), 64 (block path always)
        for K in sorted({1, 2, 4, 8, 16} if es == 4 else {1, 2, 4, 8}):
            x = torch.from_numpy(make_input(rng, dtype, (K, cin))).cuda()

            def run():
                y = torch.zeros((K, cout), dtype=x.dtype, device="cuda")
                cg.matmul_ptr(dt, x.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(), stream=stream)
                torch.cuda.synchronize()
                return y.cpu().numpy()

            cg.set_tuning(direction, dt, K)
            ref = run()
            if np.dtype(dtype).kind == "i":
                assert bits_equal(ref, og.matmul(x.cpu().numpy(), oracle.UP if direction == capi.UP else oracle.DOWN))
            else:
                want = og.matmul(x.cpu().numpy().astype(np.float64),
                                 oracle.UP if direction == capi.UP else oracle.DOWN)
                tol = 1e-5 if es == 4 else 1e-12
                assert np.abs(ref - want).max() <= tol * max(1.0, np.abs(want).max())
            for tw in _team_widths(K, es):
                for chunk in (1 << 30, 20000):
                    cg.set_tuning(direction, dt, K, team_lanes=tw, max_launch_entries=chunk)
                    assert bits_equal(run(), ref), (direction, K, tw, chunk)
            cg.set_tuning(direction, dt, K)
    cg.close()
