"""GRG construction: MutableGRG + map_mutations (the reference's construction API,
src/python/_grgl.cpp:759-925 and :1029-1062; src/map_mutations.cpp:609-876).

The mapping heuristic is this library's own, so graph shapes are not compared with the
reference's.  What any correct construction must satisfy is pinned instead:
  * the reference's own expect files: UP(ones) on a GRG of test-200-samples.vcf.gz gives the
    10,893 ALT counts of freq.txt, and by-individual UP(ones) gives Aa + 2*aa of
    zygosity.txt (test/endtoend/test_basic.py:71-96,134-149);
  * the genotype matrix read back from the graph (path counts) equals the input carrier
    sets exactly, which also proves every sample reaches a mutation node by one path;
  * duplicated sets reuse a node, missing-data rows give missingness nodes, batch sizes
    change only the shape, and coalescence counts match a brute-force count.
"""

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import dense_x, make_input, vcf_fixture

INVALID = grgl_amd.INVALID_NODE


def _carrier_lists(fx):
    off, car = fx["carrier_off"], fx["carriers"]
    return [car[off[m]:off[m + 1]].astype(np.uint32).tolist() for m in range(len(off) - 1)]


def _map(num_samples, sets, ploidy=2, positions=None, alleles=None, batch=1, size_order=True):
    g = grgl_amd.MutableGRG(num_samples, ploidy)
    muts = [grgl_amd.Mutation(float(positions[i] if positions is not None else i),
                              alleles[i] if alleles is not None else "A", "G") for i in range(len(sets))]
    stats = grgl_amd.map_mutations(g, muts, sets, mutation_batch_size=batch, size_order=size_order)
    return g, stats


def _x_from_sets(num_samples, sets):
    X = np.zeros((num_samples, len(sets)))
    for m, s in enumerate(sets):
        X[list(s), m] = 1
    return X


def test_vcf_fixture_matches_reference_expect_files():
    fx = vcf_fixture()
    N = int(fx["num_samples"])
    sets = _carrier_lists(fx)
    g, stats = _map(N, sets, ploidy=int(fx["ploidy"]), positions=fx["positions"])
    a = g.arrays()
    assert g.num_mutations == len(sets) == 10893 and stats.total_mutations == len(sets)
    og = oracle.OracleGraph(a)
    counts = og.matmul(np.ones((1, N), np.int64), oracle.UP)[0]
    assert np.array_equal(counts, fx["expect_alt_count"])
    zyg = og.matmul(np.ones((1, N // 2), np.int64), oracle.UP, by_individual=True)[0]
    assert np.array_equal(zyg, fx["expect_Aa"] + 2 * fx["expect_aa"])
    assert np.array_equal(dense_x(a), _x_from_sets(N, sets))
    # Sharing: far fewer edges than one edge per (mutation, carrier).
    assert g.num_edges < 0.5 * sum(len(s) for s in sets)
    assert stats.reused_nodes > 0 and stats.reused_exactly > 0


def test_random_overlapping_sets_reconstruct_exactly():
    rng = np.random.default_rng(3)
    N = 60
    sets = []
    for _ in range(400):
        k = int(rng.integers(0, N + 1))
        sets.append(sorted(rng.choice(N, size=k, replace=False).tolist()))
    sets += [sets[5], sets[17], [], [3], list(range(N))]
    g, stats = _map(N, sets)
    a = g.arrays()
    assert np.array_equal(dense_x(a), _x_from_sets(N, sets))
    assert stats.empty_mutations == sum(1 for s in sets if not s)
    assert stats.mutations_with_one_sample == sum(1 for s in sets if len(s) == 1)
    assert a["mutation_nodes"][len(sets) - 3] == INVALID
    assert a["mutation_nodes"][len(sets) - 2] == 3
    # Duplicated sets map onto the node made for the first copy.
    assert a["mutation_nodes"][400] == a["mutation_nodes"][5]


def test_laminar_sets_from_trees_give_a_compact_graph():
    tree = grgl_amd.grg_from_coalescent(300, 3000, 2, seed=11)
    t = tree.arrays()
    X = dense_x(t)
    sets = [np.nonzero(X[:, m])[0].tolist() for m in range(X.shape[1])]
    g, _ = _map(300, sets)
    a = g.arrays()
    assert np.array_equal(dense_x(a), X)
    assert g.num_edges <= 1.2 * tree.num_edges


def test_batch_size_changes_shape_not_genotypes():
    fx = vcf_fixture()
    N = int(fx["num_samples"])
    sets = _carrier_lists(fx)[:2000]
    X = _x_from_sets(N, sets)
    edges = {}
    for batch in (0, 1, 7, 500):
        for size_order in (False, True):
            g, _ = _map(N, sets, batch=batch, size_order=size_order)
            assert np.array_equal(dense_x(g.arrays()), X), (batch, size_order)
            edges[(batch, size_order)] = g.num_edges
    # One batch over an empty graph has nothing to reuse (as in the reference): one edge per
    # carrier of every mutation with two or more carriers (singletons map to the sample).
    # Smallest-first, one mutation per batch, shares the most.
    flat = sum(len(s) for s in sets if len(s) >= 2)
    assert edges[(0, False)] == edges[(0, True)] == flat
    assert edges[(1, True)] == min(edges.values()) < 0.5 * X.sum()


def test_missing_rows_and_existing_graph():
    N = 8
    g = grgl_amd.MutableGRG(N, 2)
    # A hand-built node over samples {0, 1} that mapping can reuse.
    n = g.make_node()
    g.connect(n, 0)
    g.connect(n, 1)
    muts = [grgl_amd.Mutation(100, ".", "A"), grgl_amd.Mutation(100, "T", "A"), grgl_amd.Mutation(100, "C", "A"),
            grgl_amd.Mutation(200, "G", "A")]
    sets = [[4, 5, 6], [0, 1], [0, 1, 2, 7], [2, 3]]
    stats = grgl_amd.map_mutations(g, muts, sets)
    a = g.arrays()
    assert g.num_mutations == 3  # the "." row is not a mutation
    assert a["mutation_nodes"][0] == n and stats.reused_exactly == 1
    miss = a["mutation_miss_nodes"]
    assert miss[0] == miss[1] != INVALID and miss[2] == INVALID
    below = dense_x({**a, "mutation_nodes": np.array([miss[0]], np.uint32)})[:, 0]
    assert below.tolist() == [0, 0, 0, 0, 1, 1, 1, 0]
    assert np.array_equal(dense_x(a), _x_from_sets(N, sets[1:]))
    assert g.get_mutation_by_id(1).allele == "C" and g.get_mutation_by_id(1).position == 100


def test_coalescence_counts_brute_force():
    rng = np.random.default_rng(9)
    N = 40
    sets = [sorted(rng.choice(N, size=int(rng.integers(2, N)), replace=False).tolist()) for _ in range(150)]
    g, _ = _map(N, sets)
    a = g.arrays()
    off, ch, coal = a["down_offsets"], a["down_children"], a["coal_counts"]
    V = len(off) - 1
    below = [{v} if v < N else set() for v in range(V)]
    for v in range(N, V):
        for c in ch[off[v]:off[v + 1]]:
            below[v] |= below[c]
    for v in range(N, V):
        owner = {}
        for c in ch[off[v]:off[v + 1]]:
            for s in below[c]:
                owner[s] = c
        want = sum(1 for i in range(N // 2) if 2 * i in owner and 2 * i + 1 in owner
                   and owner[2 * i] != owner[2 * i + 1])
        assert coal[v] == want, v


def test_round_trip_and_oracle_on_constructed_graph(tmp_path):
    fx = vcf_fixture()
    N = int(fx["num_samples"])
    g, _ = _map(N, _carrier_lists(fx)[:3000], positions=fx["positions"][:3000])
    fn = str(tmp_path / "built.grg")
    grgl_amd.save_grg(g, fn)
    h = grgl_amd.load_immutable_grg(fn)
    a, b = oracle.OracleGraph(g.arrays()), oracle.OracleGraph(h.arrays())
    x = make_input(np.random.default_rng(2), np.float64, (2, N))
    assert np.array_equal(a.matmul(x, oracle.UP), b.matmul(x, oracle.UP))
    assert h.get_mutation_by_id(7).position == int(fx["positions"][7])


def test_argument_errors():
    g = grgl_amd.MutableGRG(4, 2)
    with pytest.raises(RuntimeError, match="same length"):
        grgl_amd.map_mutations(g, [grgl_amd.Mutation(1, "A")], [])
    with pytest.raises(RuntimeError, match="out of range"):
        grgl_amd.map_mutations(g, [grgl_amd.Mutation(1, "A")], [[0, 9]])
    n = g.make_node()
    with pytest.raises(RuntimeError, match="smaller node ID"):
        g.connect(0, n)
    with pytest.raises(RuntimeError):
        grgl_amd.MutableGRG(5, 2)


@pytest.mark.gpu
def test_gpu_matmul_on_mapped_graph(gpu):
    fx = vcf_fixture()
    N = int(fx["num_samples"])
    g, _ = _map(N, _carrier_lists(fx))
    og = oracle.OracleGraph(g.arrays())
    ones = np.ones((1, N), np.int32)
    assert np.array_equal(grgl_amd.matmul(g, ones, grgl_amd.UP)[0], fx["expect_alt_count"])
    w = make_input(np.random.default_rng(5), np.int64, (3, g.num_mutations))
    assert np.array_equal(grgl_amd.matmul(g, w, grgl_amd.DOWN), og.matmul(w, oracle.DOWN))
    # Growing the graph afterwards invalidates the device copy.
    n = g.make_node()
    g.connect(n, 0)
    g.add_mutation(grgl_amd.Mutation(1e9, "T"), n)
    assert grgl_amd.matmul(g, ones, grgl_amd.UP)[0, -1] == 1
