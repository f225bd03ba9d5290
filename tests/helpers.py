"""Graph builders and comparison helpers shared by the tests.

Graphs are plain-array GRGs (the content of an ordered CSRGRG): samples 0..N-1, every
child ID below its parent's, children sorted ascending.
"""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
INVALID = 0x7FFFFFFF


def csr_from_lists(num_samples, children_lists, mutation_nodes, ploidy=2, miss=None, coal=None):
    off = np.zeros(len(children_lists) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(c) for c in children_lists])
    child = np.array([c for cl in children_lists for c in sorted(cl)], dtype=np.uint32)
    d = {
        "num_samples": num_samples,
        "ploidy": ploidy,
        "down_offsets": off,
        "down_children": child,
        "mutation_nodes": np.asarray(mutation_nodes, dtype=np.uint32),
        "mutation_miss_nodes": None if miss is None else np.asarray(miss, dtype=np.uint32),
        "coal_counts": None if coal is None else np.asarray(coal, dtype=np.uint32),
    }
    return d


def depth3_bin_tree(mutation_nodes, coal=None):
    """test/unit/common_grgs.h:13-26: samples 0-3, 4->(0,1), 5->(2,3), 6->(4,5)."""
    kids = [[], [], [], [], [0, 1], [2, 3], [4, 5]]
    return csr_from_lists(4, kids, mutation_nodes, ploidy=2, coal=coal)


def sample8_grg(mutation_nodes):
    """test/unit/common_grgs.h:35-58."""
    kids = [[]] * 8 + [[0, 1], [8, 2], [3, 4, 7], [8, 10, 6], [10, 5], [12, 9]]
    return csr_from_lists(8, kids, mutation_nodes, ploidy=2)


def random_grg(seed, num_samples=64, num_internal=200, max_children=6, heavy=(), num_mutations=300,
               with_miss=False, with_coal=False, ploidy=2, unmapped_frac=0.05):
    """A random ordered DAG GRG.  `heavy` lists fan-ins for extra high-fan-in nodes."""
    rng = np.random.default_rng(seed)
    N = num_samples
    kids = [[] for _ in range(N)]
    fanins = [int(rng.integers(2, max_children + 1)) for _ in range(num_internal)] + list(heavy)
    rng.shuffle(fanins)
    for f in fanins:
        v = len(kids)
        f = min(f, v)
        kids.append(sorted(rng.choice(v, size=f, replace=False).tolist()))
    V = len(kids)
    muts = rng.integers(0, V, size=num_mutations).astype(np.int64)
    muts[rng.random(num_mutations) < unmapped_frac] = INVALID
    miss = None
    if with_miss:
        miss = rng.integers(N, V, size=num_mutations).astype(np.int64)
        miss[rng.random(num_mutations) < 0.5] = INVALID
    coal = None
    if with_coal:
        coal = rng.integers(0, max(1, N // ploidy // 4), size=V).astype(np.int64)
        coal[:N] = 0
    return csr_from_lists(N, kids, muts, ploidy=ploidy, miss=miss, coal=coal)


def vcf_fixture():
    return np.load(os.path.join(GOLDEN, "test200.npz"))


def direct_map_grg(fx):
    """GRG of the reference's test VCF: one node per distinct carrier set, mutations on it."""
    N = int(fx["num_samples"])
    off, car = fx["carrier_off"], fx["carriers"]
    kids = [[] for _ in range(N)]
    node_of_set = {}
    muts = []
    for m in range(len(off) - 1):
        c = tuple(int(x) for x in car[off[m]:off[m + 1]])
        if not c:
            muts.append(INVALID)
            continue
        if c not in node_of_set:
            node_of_set[c] = len(kids)
            kids.append(list(c))
        muts.append(node_of_set[c])
    return csr_from_lists(N, kids, muts, ploidy=int(fx["ploidy"]))


def dense_x(arrays, by_individual=False):
    """Genotype matrix X (N x M, or individuals x M): X[s, m] = number of down-paths from
    m's node to sample s.  On a valid GRG every path is unique, so X is the 0/1 genotype
    matrix (the reference's grg2X, test/endtoend/testing_utils.py:66-101); on an arbitrary
    DAG the graph product counts paths, which is what this returns."""
    N, P = arrays["num_samples"], arrays["ploidy"]
    off, child = arrays["down_offsets"], arrays["down_children"]
    V = len(off) - 1
    below = np.zeros((V, N), dtype=np.float64)
    for s in range(N):
        below[s, s] = 1.0
    for v in range(N, V):
        for c in child[off[v]:off[v + 1]]:
            below[v] += below[c]
    mut = arrays["mutation_nodes"]
    M = len(mut)
    X = np.zeros((N, M))
    for m in range(M):
        if mut[m] != INVALID:
            X[:, m] = below[mut[m]]
    if by_individual:
        X = X.reshape(N // P, P, M).sum(axis=1)
    return X


def make_input(rng, dtype, shape):
    dtype = np.dtype(dtype)
    if dtype == np.bool_:
        return rng.random(shape) < 0.5
    if np.issubdtype(dtype, np.floating):
        return rng.standard_normal(shape).astype(dtype)
    info = np.iinfo(dtype)
    lo, hi = max(info.min, -1000), min(info.max, 1000)
    return rng.integers(lo, hi, size=shape, endpoint=True).astype(dtype)


def normwise_rel_err(a, b):
    """max over rows of max|a-b| / max|b| (0 where a row of b is all zero and a == b)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    num = np.abs(a - b).max(axis=1)
    den = np.abs(b).max(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, num / den, num)
    return float(r.max()) if r.size else 0.0


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def missing_fixture_rows():
    """The rows of the reference's test-200-samples.miss.igd (tests/golden/test200_miss.npz over
    tests/golden/test200.npz): per site its missing-data row first (haplotypes with no call),
    then the site's variants without those haplotypes.  Returns (num_samples, ploidy,
    positions, is_missing, carrier lists)."""
    fx = vcf_fixture()
    m = np.load(os.path.join(GOLDEN, "test200_miss.npz"))
    miss = {int(p): m["miss_carriers"][m["miss_off"][i]:m["miss_off"][i + 1]].astype(np.uint32)
            for i, p in enumerate(m["miss_positions"])}
    off, car, P = fx["carrier_off"], fx["carriers"], fx["positions"]
    pos, is_miss, sets = [], [], []
    for i, p in enumerate(P.tolist()):
        if p in miss and (i == 0 or P[i - 1] != p):
            pos.append(p)
            is_miss.append(True)
            sets.append(miss[p])
        row = car[off[i]:off[i + 1]]
        if p in miss:
            row = np.setdiff1d(row, miss[p]).astype(np.uint32)
        pos.append(p)
        is_miss.append(False)
        sets.append(row)
    return int(fx["num_samples"]), int(fx["ploidy"]), np.array(pos), np.array(is_miss), sets
