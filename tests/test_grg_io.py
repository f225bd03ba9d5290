"""The .grg file format (reference format 5.x, src/serialize.cpp:518-669 and :697-890)."""

import os
import struct

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import INVALID, csr_from_lists, dense_x, depth3_bin_tree, bits_equal, make_input, random_grg

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference library oracle/_ref not built here")


def _arrays_equal(a, b):
    for k in ("down_offsets", "down_children", "mutation_nodes"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("mutation_miss_nodes", "coal_counts"):
        if a[k] is None:
            assert b[k] is None, k
        else:
            assert np.array_equal(a[k], b[k]), k


def _rich_graph():
    a = random_grg(3, num_samples=40, num_internal=120, heavy=(50,), num_mutations=200, with_miss=True,
                   with_coal=True)
    rng = np.random.default_rng(4)
    refs = ["A", "CT", "G" * 9, ""] * 50
    alts = ["G", "C", "A", "TTTTTTTTTTT"] * 50
    g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                 a["mutation_nodes"], mutation_miss_nodes=a["mutation_miss_nodes"],
                                 coal_counts=a["coal_counts"],
                                 positions=np.sort(rng.integers(0, 10**7, 200)).astype(np.float64),
                                 ref_alleles=refs, alt_alleles=alts, populations=["EUR", "AFR"],
                                 population_ids=rng.integers(0, 2, 40).astype(np.uint16),
                                 individual_ids=[f"ind{i}" for i in range(20)])
    return a, g, refs, alts


def test_round_trip_everything(tmp_path):
    a, g, refs, alts = _rich_graph()
    fn = str(tmp_path / "rich.grg")
    grgl_amd.save_grg(g, fn, allow_simplify=False)
    h = grgl_amd.load_immutable_grg(fn)
    _arrays_equal(g.arrays(), h.arrays())
    for m in (0, 1, 2, 3, 199):
        mu, mv = g.get_mutation_by_id(m), h.get_mutation_by_id(m)
        assert (mu.position, mu.ref_allele, mu.allele) == (mv.position, mv.ref_allele, mv.allele)
        assert (mv.ref_allele, mv.allele) == (refs[m], alts[m])
    assert h.get_populations() == ["EUR", "AFR"]
    assert h.has_individual_ids and h.get_individual_id(7) == "ind7"
    assert h.has_missing_data and h.has_individual_coals


def test_loaded_graph_multiplies_like_the_original(tmp_path):
    g = grgl_amd.grg_from_coalescent(500, 4000, 3, seed=8)
    fn = str(tmp_path / "tree.grg")
    grgl_amd.save_grg(g, fn)
    h = grgl_amd.load_immutable_grg(fn)
    a, b = oracle.OracleGraph(g.arrays()), oracle.OracleGraph(h.arrays())
    x = make_input(np.random.default_rng(1), np.float64, (3, 500))
    assert bits_equal(a.matmul(x, oracle.UP), b.matmul(x, oracle.UP))
    assert g.get_mutation_by_id(10).position == h.get_mutation_by_id(10).position


def _sections(fn):
    data = open(fn, "rb").read()
    hdr = struct.unpack("<QHHHHQQQQQQQ", data[:72])
    nodes = hdr[7]
    ebytes = struct.unpack("<Q", data[128:136])[0]
    edges = np.frombuffer(data[136:136 + ebytes], np.uint8)
    idx = np.frombuffer(data[136 + ebytes:136 + ebytes + 8 * (nodes + 1)], np.uint64)
    return data, hdr, edges, idx


@needs_ref
def test_edge_section_bytes_match_reference_encoder(tmp_path):
    """The varint CSR edge section equals what the reference's EagerCSREdges32 writes."""
    for a in (grgl_amd.grg_from_coalescent(800, 6000, 4, seed=2).arrays(),
              random_grg(6, num_samples=100, num_internal=600, heavy=(300, 5000), num_mutations=100)):
        g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                     a["mutation_nodes"])
        fn = str(tmp_path / "e.grg")
        grgl_amd.save_grg(g, fn, allow_simplify=False)
        _, hdr, edges, idx = _sections(fn)
        ref_edges, ref_idx = oracle.OracleGraph(a).ref_edge_section()
        assert np.array_equal(edges, ref_edges)
        assert np.array_equal(idx, ref_idx)
        assert hdr[0] == 0xE9366C64DDC8C5B0 and (hdr[1], hdr[2]) == (5, 3)
        # Simplified on write: the section holds the simplified graph, encoded the same way.
        grgl_amd.save_grg(g, fn)
        _, hdr, edges, idx = _sections(fn)
        ref_edges, ref_idx = oracle.OracleGraph(grgl_amd.load_immutable_grg(fn).arrays()).ref_edge_section()
        assert np.array_equal(edges, ref_edges) and np.array_equal(idx, ref_idx)
        assert hdr[7] <= len(a["down_offsets"]) - 1 and hdr[8] <= len(a["down_children"])


def test_reads_version_5_1_small_index(tmp_path):
    """Files up to version 5.1 store u32 node offsets (serialize.cpp:85, :783-790)."""
    g = grgl_amd.grg_from_coalescent(300, 2000, 2, seed=4)
    fn = str(tmp_path / "v53.grg")
    grgl_amd.save_grg(g, fn)
    data, hdr, edges, idx = _sections(fn)
    nodes = hdr[7]
    start = 136 + len(edges)
    old = bytearray(data[:start]) + idx.astype(np.uint32).tobytes() + data[start + 8 * (nodes + 1):]
    struct.pack_into("<H", old, 10, 1)  # versionMinor = 1
    fn51 = str(tmp_path / "v51.grg")
    open(fn51, "wb").write(bytes(old))
    _arrays_equal(g.arrays(), grgl_amd.load_immutable_grg(fn51).arrays())


def test_corrupt_files_raise(tmp_path):
    g = grgl_amd.grg_from_coalescent(100, 500, 1, seed=1)
    fn = str(tmp_path / "ok.grg")
    grgl_amd.save_grg(g, fn)
    data = open(fn, "rb").read()
    bad = str(tmp_path / "bad.grg")
    open(bad, "wb").write(b"\0" * 8 + data[8:])
    with pytest.raises(RuntimeError, match="Invalid file header"):
        grgl_amd.load_immutable_grg(bad)
    open(bad, "wb").write(data[: len(data) // 2])
    with pytest.raises(RuntimeError, match="Malformed"):
        grgl_amd.load_immutable_grg(bad)
    with pytest.raises(RuntimeError):
        grgl_amd.load_immutable_grg(str(tmp_path / "missing.grg"))


def test_corrupt_counts_fail_before_allocating(tmp_path):
    """Counts that the file cannot hold (the edge section's byte count, the node count, a list
    longer than its bytes) are malformed-file errors, raised before any buffer of that size is
    allocated."""
    import struct

    g = grgl_amd.grg_from_coalescent(100, 500, 1, seed=1)
    fn = str(tmp_path / "ok.grg")
    grgl_amd.save_grg(g, fn)
    data = bytearray(open(fn, "rb").read())
    bad = str(tmp_path / "bad.grg")
    # edge section byte count (first field after the 128-byte header): 2^60
    d = bytearray(data)
    d[128:136] = struct.pack("<Q", 1 << 60)
    open(bad, "wb").write(bytes(d))
    with pytest.raises(RuntimeError, match="Malformed"):
        grgl_amd.load_immutable_grg(bad)
    # node count (header offset 32): 2^31 - 2 nodes, far more index entries than the file holds
    d = bytearray(data)
    d[32:40] = struct.pack("<Q", (1 << 31) - 2)
    open(bad, "wb").write(bytes(d))
    with pytest.raises(RuntimeError, match="Malformed"):
        grgl_amd.load_immutable_grg(bad)
    # the first non-empty list claims 2^28 children in a few bytes (varint 0x80 0x80 0x80 0x80 0x01)
    edge_bytes = struct.unpack("<Q", data[128:136])[0]
    edges = data[136:136 + edge_bytes]
    first = next(i for i in range(len(edges)) if edges[i] != 0)
    d = bytearray(data)
    d[136 + first:136 + first + 5] = bytes([0x80, 0x80, 0x80, 0x80, 0x01])
    open(bad, "wb").write(bytes(d))
    with pytest.raises(RuntimeError, match="Malformed"):
        grgl_amd.load_immutable_grg(bad)


def test_truncated_and_flipped_files_fail_cleanly(tmp_path):
    """Every prefix of a small file, and random byte flips of it, either load as a valid graph or
    raise RuntimeError: no crash and no other exception type.  (The same corpus was run through
    the reader, validation and levelization built with -fsanitize=address,undefined.)"""
    g = grgl_amd.grg_from_coalescent(60, 300, 2, seed=4)
    fn = str(tmp_path / "ok.grg")
    grgl_amd.save_grg(g, fn)
    data = open(fn, "rb").read()
    rng = np.random.default_rng(5)
    bad = str(tmp_path / "bad.grg")
    variants = [data[:c] for c in range(0, len(data), max(1, len(data) // 300))]
    for _ in range(300):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 4))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        variants.append(bytes(b))
    loaded = 0
    for v in variants:
        open(bad, "wb").write(v)
        try:
            h = grgl_amd.load_immutable_grg(bad)
        except RuntimeError:
            continue
        loaded += 1
        assert h.num_nodes >= h.num_samples
    assert 0 < loaded < len(variants)


@pytest.mark.gpu
def test_gpu_matmul_on_loaded_file(gpu, tmp_path):
    g = grgl_amd.grg_from_coalescent(2000, 20000, 4, seed=12)
    fn = str(tmp_path / "g.grg")
    grgl_amd.save_grg(g, fn)
    h = grgl_amd.load_immutable_grg(fn)
    o = oracle.OracleGraph(g.arrays())
    x = make_input(np.random.default_rng(3), np.float32, (4, 2000))
    y = make_input(np.random.default_rng(4), np.int32, (2, h.num_mutations))
    assert bits_equal(grgl_amd.matmul(h, x, grgl_amd.UP, order_preserving=True), o.matmul(x, oracle.UP))
    assert bits_equal(grgl_amd.matmul(h, y, grgl_amd.DOWN), o.matmul(y, oracle.DOWN))


@needs_ref
def test_mutation_records_match_reference_layout(tmp_path):
    """24-byte mutation records and deferred allele strings equal the raw bytes of the
    reference's own Mutation objects (src/serialize.cpp:608-624); for long alleles the
    first 8 bytes hold the writer's heap pointer, so only its stolen bit 0 is compared."""
    a = grgl_amd.grg_from_coalescent(100, 12, 1, seed=3).arrays()
    M = len(a["mutation_nodes"])
    refs = ["A", "CT", "GGGGGGGGG", "", "ACG", "T", "AAAA", "", "C", "G" * 300, "A", "TT"][:M]
    alts = ["G", "C", "A", "TTTTTTTTTTT", "TTTT", "TTTTTTT", "CCC", "", "T", "A", "ACGTACG", "A"][:M]
    pos = np.arange(M, dtype=np.float64) * 1000 + 123456789
    g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                 a["mutation_nodes"], positions=pos, ref_alleles=refs, alt_alleles=alts)
    fn = str(tmp_path / "m.grg")
    grgl_amd.save_grg(g, fn)
    data, hdr, edges, idx = _sections(fn)
    start = 136 + len(edges) + 8 * len(idx) + 4 * M
    recs = np.frombuffer(data[start:start + 24 * M], np.uint8).reshape(M, 24)
    tail = data[start + 24 * M:]
    deferred = []
    for m in range(M):
        want, short, storage = oracle.ref_mutation_record(int(pos[m]), refs[m], alts[m])
        if short:
            assert np.array_equal(recs[m], want), m
        else:
            assert recs[m][0] & 1 and np.array_equal(recs[m][8:], want[8:]), m
            deferred.append(storage)
    # Deferred strings: LEB128 length + bytes, in mutation order (serialize.cpp:620-622).
    pos_t = 0
    for st in deferred:
        n, shift = 0, 0
        while True:
            b = tail[pos_t]
            pos_t += 1
            n |= (b & 0x7F) << shift
            shift += 7
            if not b & 0x80:
                break
        assert tail[pos_t:pos_t + n].decode() == st
        pos_t += n


# --- simplify-on-write (src/serialize.cpp:387-467) ------------------------------------------

def _save_load(g, tmp_path, name, **kw):
    fn = str(tmp_path / name)
    grgl_amd.save_grg(g, fn, **kw)
    return grgl_amd.load_immutable_grg(fn)


def test_simplify_rule_hand_cases(tmp_path):
    """The reference's keep/drop rule on small graphs, node by node.
    depth3 tree (test/unit/common_grgs.h:13-26) with mutations on 4 and 6: node 4 (2 children,
    1 parent) carries a mutation and stays; node 5 (2 children, 1 parent, no mutation) is
    dropped and its children move to 6; node 6 (no parents) carries a mutation and stays."""
    a = depth3_bin_tree([4, 6], coal=[0, 0, 0, 0, 1, 2, 0])
    g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                 a["mutation_nodes"], coal_counts=a["coal_counts"])
    h = _save_load(g, tmp_path, "d3.grg").arrays()
    assert h["down_offsets"].tolist() == [0, 0, 0, 0, 0, 2, 5]
    assert h["down_children"].tolist() == [0, 1, 2, 3, 4]
    assert h["mutation_nodes"].tolist() == [4, 5]
    assert h["coal_counts"][4:].tolist() == [1, 0 + 2]  # 6 takes the dropped node 5's count
    # Two children and two parents: dropped; a node with one child is dropped too; a childless
    # node's mutation becomes unmapped.
    kids = [[], [], [], [], [0, 1], [4, 2], [4, 3], [2], [], [7, 6]]
    a = csr_from_lists(4, kids, [5, 6, 8, 9], ploidy=2)
    g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                 a["mutation_nodes"])
    h = _save_load(g, tmp_path, "x.grg").arrays()
    # 4: two children, two parents -> dropped.  5 -> (0,1,2), a root with a mutation: kept as 4.
    # 6 -> (0,1,3), one parent, a mutation: kept as 5.  7: one child -> dropped.  8: no
    # children -> dropped, its mutation unmapped.  9 -> 2 (through 7) and 6: kept as 6.
    assert h["down_offsets"].tolist() == [0, 0, 0, 0, 0, 3, 6, 8]
    assert h["down_children"].tolist() == [0, 1, 2, 0, 1, 3, 2, 5]
    assert h["mutation_nodes"].tolist() == [4, 5, INVALID, 6]


def test_simplify_keeps_genotypes_and_unique_paths(tmp_path):
    """On random DAGs and mapped graphs the simplified file multiplies exactly like the input
    (int64, both directions), with no more nodes or edges."""
    graphs = [random_grg(s, num_samples=80, num_internal=400, max_children=6, num_mutations=500, unmapped_frac=0.1)
              for s in (1, 2, 3)]
    graphs.append(grgl_amd.grg_from_coalescent(300, 3000, 2, seed=7).arrays())
    for a in graphs:
        g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                     a["mutation_nodes"])
        h = _save_load(g, tmp_path, "r.grg")
        assert h.num_nodes <= g.num_nodes and h.num_edges <= g.num_edges
        oa, ob = oracle.OracleGraph(a), oracle.OracleGraph(h.arrays())
        rng = np.random.default_rng(0)
        x = rng.integers(-50, 50, (3, g.num_samples)).astype(np.int64)
        w = rng.integers(-50, 50, (3, g.num_mutations)).astype(np.int64)
        assert np.array_equal(oa.matmul(x, oracle.UP), ob.matmul(x, oracle.UP))
        assert np.array_equal(oa.matmul(w, oracle.DOWN), ob.matmul(w, oracle.DOWN))
    # the coalescent graphs are already in simplified form: the file is the graph
    a = graphs[-1]
    g = grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                 a["mutation_nodes"])
    _arrays_equal(g.arrays(), _save_load(g, tmp_path, "t.grg").arrays())


def _brute_coals(off, ch, N):
    """Per node: individuals (haplotypes 2i, 2i+1) whose two haplotypes arrive through
    different children."""
    V = len(off) - 1
    below = [{v} if v < N else set() for v in range(V)]
    for v in range(N, V):
        for c in ch[off[v]:off[v + 1]]:
            below[v] |= below[c]
    out = [0] * V
    for v in range(N, V):
        owner = {}
        for c in ch[off[v]:off[v + 1]]:
            for smp in below[c]:
                owner[smp] = c
        out[v] = sum(1 for i in range(N // 2) if 2 * i in owner and 2 * i + 1 in owner
                     and owner[2 * i] != owner[2 * i + 1])
    return out


def test_simplified_coalescence_counts_brute_force(tmp_path):
    """Coalescence counts of dropped nodes move to the nodes that take their children; on the
    simplified graph they equal a brute-force count (individuals whose two haplotypes arrive
    through different children).  Random binary trees, mutations on 40% of the nodes."""
    for seed in range(4):
        rng = np.random.default_rng(seed)
        N = 40
        kids = [[] for _ in range(N)]
        active = list(range(N))
        while len(active) > 1:
            i, j = sorted(rng.choice(len(active), size=2, replace=False).tolist(), reverse=True)
            a_, b_ = active.pop(i), active.pop(j)
            kids.append(sorted([a_, b_]))
            active.append(len(kids) - 1)
        V = len(kids)
        muts = [v for v in range(N, V) if rng.random() < 0.4] + [V - 1]
        a = csr_from_lists(N, kids, muts, ploidy=2)
        coal = np.array(_brute_coals(a["down_offsets"], a["down_children"], N), np.uint32)
        g = grgl_amd.grg_from_arrays(N, 2, a["down_offsets"], a["down_children"], a["mutation_nodes"],
                                     coal_counts=coal)
        h = _save_load(g, tmp_path, "c.grg").arrays()
        assert len(h["down_offsets"]) - 1 < V
        want = _brute_coals(h["down_offsets"], h["down_children"], N)
        assert h["coal_counts"][N:].tolist() == want[N:], seed
        assert np.array_equal(dense_x(h), dense_x(a))


def test_mapped_graph_mutations_written_in_position_order(tmp_path):
    """A graph grown by map_mutations has unordered mutation IDs; the file lists them by
    position (then REF, ALT), as the reference's writer does (grg.h:528-545), with their nodes."""
    N = 20
    rng = np.random.default_rng(4)
    pos = [500.0, 10.0, 300.0, 10.0, 42.0]
    refs = ["A", "C", "G", "A", "T"]
    sets = [sorted(rng.choice(N, size=5, replace=False).tolist()) for _ in pos]
    g = grgl_amd.MutableGRG(N, 2)
    grgl_amd.map_mutations(g, [grgl_amd.Mutation(p, "G", r) for p, r in zip(pos, refs)], sets)
    h = _save_load(g, tmp_path, "m.grg")
    order = sorted(range(len(pos)), key=lambda m: (pos[m], refs[m], "G"))
    assert [h.get_mutation_by_id(i).position for i in range(len(pos))] == [int(pos[m]) for m in order]
    X = dense_x(h.arrays())
    for i, m in enumerate(order):
        assert sorted(np.nonzero(X[:, i])[0].tolist()) == sets[m]
