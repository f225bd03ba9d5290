"""Graph construction from haplotypes (fast_grg_from_samples; the reference's fastGRGFromSamples,
src/fast_build.cpp:794-986) on the CPU.

Pinned by what a correct build must satisfy, independent of how it clusters:
  * the reference's own expect files for its 200-individual test VCF (allele counts and
    AA/Aa/aa per variant, test/endtoend/test_basic.py:71-149; fixture tests/golden/test200.npz):
    allele counts are the UP product of ones, homozygous counts the UP product of zeros seeded
    with twice the coalescence counts (init="xtx"), through the oracle and the reference's own
    compiled matmul (oracle/_ref) where built;
  * every mutation's carrier set reproduced exactly, every sample reached by at most one path;
  * coalescence counts equal to a brute-force count on the finished graph;
  * the tree-splitting options (count, faster modes, low-frequency direct mapping) and the
    edge cases (empty rows, duplicates, ploidy 1, errors);
  * missing data: the reference's test-200-samples.miss.igd (fixture tests/golden/
    test200_miss.npz) gives the missing-sample and missing-individual totals its test states,
    and the matmul identities with `miss` of test/endtoend/test_missing.py:55-89."""

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import INVALID, dense_x, vcf_fixture

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference library oracle/_ref not built here")


def _csr(sets):
    off = np.zeros(len(sets) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in sets])
    flat = np.concatenate([np.asarray(s, np.uint32) for s in sets]) if sets else np.zeros(0, np.uint32)
    return off, flat


def _carrier_lists(off, car):
    return [car[off[i]:off[i + 1]] for i in range(len(off) - 1)]


def _coalescent_input(n, m, trees, seed):
    src = grgl_amd.grg_from_coalescent(n, m, trees, seed=seed)
    off, car = src.mutation_carriers()
    return src, off.astype(np.int64), car


def _build(n, off, car, ploidy=2, **kw):
    return grgl_amd.fast_grg_from_samples(n, ploidy, carrier_offsets=off, carriers=car, **kw)


def _assert_same_carriers(g, off, car):
    o2, c2 = g.mutation_carriers()
    assert np.array_equal(o2.astype(np.int64), off.astype(np.int64))
    assert np.array_equal(c2, car)


def _brute_coals(arrays):
    """Individuals whose two haplotypes reach the node through different children, per node."""
    N = int(arrays["num_samples"])
    off, ch = arrays["down_offsets"], arrays["down_children"]
    V = len(off) - 1
    below = np.zeros((V, N), bool)
    below[np.arange(N), np.arange(N)] = True
    out = np.zeros(V, np.int64)
    for v in range(N, V):
        kids = ch[off[v]:off[v + 1]]
        owner = np.full(N, -1)
        for i, c in enumerate(kids):
            below[v] |= below[c]
            owner[below[c]] = i
        both = below[v][0::2] & below[v][1::2]
        out[v] = int(np.sum(both & (owner[0::2] != owner[1::2])))
    return out


@pytest.fixture(scope="module")
def test200():
    fx = vcf_fixture()
    off = fx["carrier_off"].astype(np.int64)
    return fx, off, fx["carriers"]


@pytest.mark.parametrize("kw", [{}, {"trees": 3}, {"trees": "faster2"}, {"lf_no_tree": 0.01}, {"lf_no_tree": 5}])
def test_reference_vcf_counts(test200, kw):
    """Allele counts and zygosity of every variant equal the reference's expect files."""
    fx, off, car = test200
    N = int(fx["num_samples"])
    g, st = _build(N, off, car, positions=fx["positions"], return_stats=True, **kw)
    assert g.num_mutations == len(off) - 1
    _assert_same_carriers(g, off, car)
    for arrays in (g.arrays(), None):
        if arrays is None:  # the graph as its file holds it (simplified on write)
            import tempfile, os

            with tempfile.TemporaryDirectory() as d:
                path = os.path.join(d, "t.grg")
                grgl_amd.save_grg(g, path)
                h = grgl_amd.load_immutable_grg(path)
            arrays = h.arrays()
            # the file orders mutations by (position, REF, ALT): the VCF order here
            assert np.array_equal(arrays["positions"], fx["positions"].astype(np.float64))
        og = oracle.OracleGraph(arrays)
        ones = np.ones((1, N), np.int64)
        assert np.array_equal(og.matmul(ones, oracle.UP)[0], fx["expect_alt_count"])
        hom2 = og.matmul(np.zeros((1, N), np.int64), oracle.UP, init="xtx")[0]
        assert np.array_equal(hom2, 2 * fx["expect_aa"])
        het = fx["expect_alt_count"] - 2 * fx["expect_aa"]
        assert np.array_equal(het, fx["expect_Aa"])
        assert np.array_equal(N // 2 - fx["expect_aa"] - het, fx["expect_AA"])
    if "trees" in kw and isinstance(kw["trees"], int):
        assert st.trees == kw["trees"]


@needs_ref
def test_reference_matmul_on_built_graph(test200):
    """The reference's own compiled matmul on the built graph gives the expect-file counts."""
    fx, off, car = test200
    N = int(fx["num_samples"])
    og = oracle.OracleGraph(_build(N, off, car).arrays())
    assert np.array_equal(og.ref_matmul(np.ones((1, N), np.int64), oracle.UP)[0], fx["expect_alt_count"])
    hom2 = og.ref_matmul(np.zeros((1, N), np.int64), oracle.UP, init="xtx")[0]
    assert np.array_equal(hom2, 2 * fx["expect_aa"])
    x = np.random.default_rng(2).standard_normal((2, len(off) - 1))
    assert np.array_equal(og.ref_matmul(x, oracle.DOWN), og.matmul(x, oracle.DOWN))


@pytest.mark.parametrize("kw", [{}, {"trees": 1}, {"trees": 5}, {"trees": "faster1"}, {"lf_no_tree": 0.02}])
def test_coalescent_unique_paths_and_coals(kw):
    """Genotypes exact, one path per (mutation, sample), coalescence counts by brute force."""
    N = 240
    src, off, car = _coalescent_input(N, 1500, 3, seed=11)
    g = _build(N, off, car, **kw)
    _assert_same_carriers(g, off, car)
    a = g.arrays()
    X = dense_x(a)  # path counts
    want = np.zeros_like(X)
    for m, c in enumerate(_carrier_lists(off, car)):
        want[c, m] = 1.0
    assert np.array_equal(X, want)
    coal = a["coal_counts"]
    assert coal is not None
    brute = _brute_coals(a)
    assert np.array_equal(coal[N:].astype(np.int64), brute[N:])
    # products through the oracle equal the dense genotype matrix's
    og = oracle.OracleGraph(a)
    v = np.random.default_rng(3).integers(-5, 6, size=(2, N)).astype(np.int64)
    assert np.array_equal(og.matmul(v, oracle.UP), (v @ want).astype(np.int64))


def test_threads_do_not_change_the_graph():
    N = 300
    _, off, car = _coalescent_input(N, 3000, 4, seed=5)
    one = _build(N, off, car, threads=1).arrays()
    many = _build(N, off, car, threads=8).arrays()
    for k in ("down_offsets", "down_children", "mutation_nodes", "coal_counts"):
        assert np.array_equal(one[k], many[k]), k


def test_tree_splitting_options():
    N = 200
    _, off, car = _coalescent_input(N, 4000, 4, seed=9)
    M = len(off) - 1
    _, s_opt = _build(N, off, car, return_stats=True)
    _, s_f1 = _build(N, off, car, trees="faster1", return_stats=True)
    _, s_f2 = _build(N, off, car, trees="faster2", return_stats=True)
    assert s_opt.trees <= s_f1.trees <= s_f2.trees
    assert s_opt.trees >= 2
    # With a count, trees take ceil(M / count) rows each and the last one the rest.
    for t in (1, 2, 7):
        g, s = _build(N, off, car, trees=t, return_stats=True)
        assert s.trees == t
        _assert_same_carriers(g, off, car)
    # More trees than rows: one row each.
    _, s = _build(N, off[:4], car[:off[3]], trees=10, return_stats=True)
    assert s.trees == 3
    # Low-frequency rows are mapped directly: by count, and by frequency (floor(N * f)).
    sizes = np.diff(off)
    _, s = _build(N, off, car, lf_no_tree=6, return_stats=True)
    assert s.direct_mapped == int(np.sum(sizes < 6))
    assert s.tree_mapped + s.unmapped == M - s.direct_mapped
    _, s = _build(N, off, car, lf_no_tree=0.05, return_stats=True)
    assert s.direct_mapped == int(np.sum(sizes < int(N * 0.05)))
    # Without tree mapping the graph comes back without mutations.
    g = _build(N, off, car, no_tree_map=True)
    assert g.num_mutations == 0 and g.num_nodes > N


def test_edge_cases():
    N = 20
    # no rows
    g = _build(N, np.zeros(1, np.int64), np.zeros(0, np.uint32))
    assert g.num_mutations == 0 and g.num_nodes == N
    # empty rows stay unmapped; duplicate carriers are listed once; full rows; lists API
    sets = [[], [3, 3, 1], list(range(N)), [0], [], [5, 6, 7, 8]]
    muts = [grgl_amd.Mutation(100 + i, "T", "C") for i in range(len(sets))]
    g = grgl_amd.fast_grg_from_samples(N, 2, muts, sets)
    a = g.arrays()
    assert a["mutation_nodes"][0] == INVALID and a["mutation_nodes"][4] == INVALID
    o, c = g.mutation_carriers()
    assert [list(c[o[i]:o[i + 1]]) for i in range(len(sets))] == [sorted(set(s)) for s in sets]
    assert g.get_mutation_by_id(1).position == 101 and g.get_mutation_by_id(1).ref_allele == "C"
    # window edges: 253, 254 and 255 rows (one full window plus a partial one)
    rng = np.random.default_rng(1)
    for m in (253, 254, 255, 509):
        sets = [sorted(rng.choice(N, size=rng.integers(1, N), replace=False).tolist()) for _ in range(m)]
        off, car = _csr(sets)
        g, s = _build(N, off, car, return_stats=True)
        assert s.windows == (m + 253) // 254
        _assert_same_carriers(g, off, car)
    # ploidy 1: no coalescence counts
    off, car = _csr([[0, 1], [1, 2, 3]])
    g = _build(5, off, car, ploidy=1)
    assert g.arrays()["coal_counts"] is None
    _assert_same_carriers(g, off, car)
    with pytest.raises(RuntimeError, match="out of range"):
        _build(4, *_csr([[0, 4]]))
    # a lone missing-data row: no mutation, nothing fails
    g = grgl_amd.fast_grg_from_samples(4, 2, [grgl_amd.Mutation(1, ".", "A")], [[0]])
    assert g.num_mutations == 0
    with pytest.raises(RuntimeError, match="non-zero"):
        _build(4, *_csr([[0]]), trees=0)
    with pytest.raises(ValueError):
        _build(4, *_csr([[0]]), trees="fastest")
    with pytest.raises(RuntimeError, match="same length"):
        grgl_amd.fast_grg_from_samples(4, 2, [grgl_amd.Mutation(1, "G", "A")], [[0], [1]])


def _missing_build(**kw):
    from helpers import missing_fixture_rows

    N, P, pos, is_miss, sets = missing_fixture_rows()
    muts = [grgl_amd.Mutation(float(p), "." if m else "G", "A") for p, m in zip(pos, is_miss)]
    g = grgl_amd.fast_grg_from_samples(N, P, muts, [s.tolist() for s in sets], **kw)
    return g, N, pos, is_miss, sets


def _missing_x(N, pos, is_miss, sets):
    """Individual dosage matrix of the fixture with each missing call set to the allele frequency
    among the called haplotypes (grg2X with diploid=True, test/endtoend/testing_utils.py:66-101,
    built here from the input rows), and per mutation its missing haplotypes."""
    muts = [i for i in range(len(sets)) if not is_miss[i]]
    X = np.zeros((N // 2, len(muts)))
    freq = np.zeros(len(muts))
    missing = []
    cur = None
    k = 0
    for i in range(len(sets)):
        if is_miss[i]:
            cur = (pos[i], sets[i])
            continue
        ms = cur[1] if cur is not None and cur[0] == pos[i] else np.zeros(0, np.uint32)
        np.add.at(X[:, k], sets[i] // 2, 1)
        freq[k] = len(sets[i]) / (N - len(ms)) if N > len(ms) else 0.0
        missing.append(ms)
        k += 1
    for k, ms in enumerate(missing):
        np.add.at(X[:, k], ms // 2, freq[k])
    return X, freq, missing


@pytest.mark.parametrize("kw", [{}, {"trees": 4}, {"lf_no_tree": 3}])
def test_missing_data_rows(kw):
    """The reference's missing-data test (test/endtoend/test_missing.py:19-90) on the graph the
    constructor builds from test-200-samples.miss.igd: missingness nodes with exactly the missing
    haplotypes, 25 missing samples and 21 missing individuals over the mutations that have a node
    and a missingness node, and the matmul identities with `miss`."""
    g, N, pos, is_miss, sets = _missing_build(**kw)
    assert g.num_mutations == int(np.sum(~is_miss))
    a = g.arrays()
    V = len(a["down_offsets"]) - 1
    X, freq, missing = _missing_x(N, pos, is_miss, sets)
    # every mutation's carriers and missingness node
    o, c = g.mutation_carriers()
    plain = [s for s, m in zip(sets, is_miss) if not m]
    for k, s in enumerate(plain):
        assert np.array_equal(c[o[k]:o[k + 1]], s), k
    miss_nodes = a["mutation_miss_nodes"]
    assert miss_nodes is not None
    below = {}
    total = total_ind = 0
    for k in range(g.num_mutations):
        node, mn = a["mutation_nodes"][k], miss_nodes[k]
        if len(missing[k]) == 0:
            assert mn == INVALID, k
            continue
        assert mn < V, k
        if mn not in below:
            below[mn] = np.sort(_samples_below(a, mn))
        assert np.array_equal(below[mn], missing[k]), k
        if node != INVALID:
            total += len(missing[k])
            total_ind += len(set((missing[k] // 2).tolist()))
    assert (total, total_ind) == (25, 21)  # MISSING_SAMPLES, MISSING_INDIVS (test_missing.py:21-22)
    og = oracle.OracleGraph(a)
    rng = np.random.default_rng(6)
    K = 7
    rv = rng.standard_normal((K, N // 2))
    Mm = np.zeros((K, g.num_mutations))
    up = og.matmul(rv, oracle.UP, by_individual=True, miss=Mm)
    np.testing.assert_allclose(rv @ X, up + Mm * freq, rtol=1e-10, atol=1e-10)
    rw = rng.standard_normal((K, g.num_mutations))
    down = og.matmul(rw, oracle.DOWN, by_individual=True, miss=np.ascontiguousarray(freq[None, :] * rw))
    np.testing.assert_allclose(rw @ X.T, down, rtol=1e-10, atol=1e-10)


def _samples_below(a, node):
    N = int(a["num_samples"])
    off, ch = a["down_offsets"], a["down_children"]
    out, stack, seen = [], [int(node)], set()
    while stack:
        n = stack.pop()
        if n < N:
            out.append(n)
            continue
        for c in ch[off[n]:off[n + 1]]:
            if int(c) not in seen:
                seen.add(int(c))
                stack.append(int(c))
    return np.array(out, dtype=np.uint32)


def test_missing_row_order_checked():
    muts = [grgl_amd.Mutation(5, "G", "A"), grgl_amd.Mutation(5, ".", "A")]
    with pytest.raises(RuntimeError, match="first row of its site"):
        grgl_amd.fast_grg_from_samples(4, 2, muts, [[0], [1]])


def test_unphased_flag():
    """phased=False marks the graph unphased and changes nothing else."""
    src, off, car = _coalescent_input(100, 500, 1, seed=3)
    a = _build(100, off, car).arrays()
    g = _build(100, off, car, phased=False)
    assert not g.is_phased
    b = g.arrays()
    for k in ("down_offsets", "down_children", "mutation_nodes", "coal_counts"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("seed", range(12))
def test_random_inputs_keep_every_invariant(seed):
    """Random haplotype tables (copied blocks with noise, duplicate, empty and full rows) under
    small windows and every index-rebuild regime (often, sometimes, never): carriers exact, one
    path per (mutation, sample), coalescence counts by brute force, thread count irrelevant."""
    rng = np.random.default_rng(seed)
    N = int(rng.choice([16, 40, 64]))
    M = int(rng.integers(50, 400))
    founders = rng.random((6, M)) < rng.uniform(0.05, 0.5)
    H = founders[rng.integers(0, 6, size=N)].copy()
    H ^= rng.random((N, M)) < 0.03
    H[:, rng.integers(0, M, size=3)] = False  # empty rows
    H[:, rng.integers(0, M, size=2)] = True   # full rows
    H[:, 1] = H[:, 0]                          # duplicate rows
    sets = [np.nonzero(H[:, m])[0].astype(np.uint32) for m in range(M)]
    off, car = _csr(sets)
    kw = dict(hap_length=int(rng.choice([3, 8, 31, 254])), rebuild_proportion=float(rng.choice([0.01, 0.1, 2.0])),
              trees=rng.choice(["optimal", 1, 3]).item() if seed % 2 else "optimal",
              lf_no_tree=float(rng.choice([0.0, 3.0])))
    if isinstance(kw["trees"], str) and kw["trees"].isdigit():
        kw["trees"] = int(kw["trees"])
    g = _build(N, off, car, **kw)
    _assert_same_carriers(g, off, car)
    a = g.arrays()
    X = dense_x(a)
    assert np.array_equal(X, H.astype(np.float64)), kw
    assert np.array_equal(a["coal_counts"][N:].astype(np.int64), _brute_coals(a)[N:]), kw
    b = _build(N, off, car, threads=1, **kw).arrays()
    for k in ("down_offsets", "down_children", "mutation_nodes"):
        assert np.array_equal(a[k], b[k]), (k, kw)
