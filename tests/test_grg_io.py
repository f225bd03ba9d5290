"""The .grg file format (reference format 5.x, src/serialize.cpp:518-669 and :697-890)."""

import os
import struct

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import bits_equal, make_input, random_grg

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
    grgl_amd.save_grg(g, fn)
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
        grgl_amd.save_grg(g, fn)
        _, hdr, edges, idx = _sections(fn)
        ref_edges, ref_idx = oracle.OracleGraph(a).ref_edge_section()
        assert np.array_equal(edges, ref_edges)
        assert np.array_equal(idx, ref_idx)
        assert hdr[0] == 0xE9366C64DDC8C5B0 and (hdr[1], hdr[2]) == (5, 3)


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
