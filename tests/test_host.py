"""Host-side logic on the CPU: the C ABI library, the generator, the level plans and the
pygrgl-compatible argument checking.  No GPU compute here."""

import ctypes
import os
import re

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import capi
from helpers import INVALID, bits_equal, depth3_bin_tree, make_input, random_grg, vcf_fixture, direct_map_grg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_abi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "grgl_hip.h")).read()
    declared = sorted(set(re.findall(r"\b(grgl_hip_[a-z_]+)\s*\(", header)))
    assert declared, "no declarations parsed"
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"
    assert sorted(capi.EXPORTED_SYMBOLS) == declared


def test_generator_deterministic_and_valid():
    a = grgl_amd.grg_from_coalescent(500, 3000, 3, seed=5).arrays()
    b = grgl_amd.grg_from_coalescent(500, 3000, 3, seed=5).arrays()
    c = grgl_amd.grg_from_coalescent(500, 3000, 3, seed=6).arrays()
    for k in ("down_offsets", "down_children", "mutation_nodes", "coal_counts"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["down_children"], c["down_children"])
    off, child = a["down_offsets"], a["down_children"]
    for v in range(len(off) - 1):
        kids = child[off[v]:off[v + 1]]
        assert np.all(kids < v) and np.all(np.diff(kids.astype(np.int64)) > 0)
    assert len(a["mutation_nodes"]) == 3000
    assert np.all(np.diff(a["positions"]) >= 0)


def test_generator_config1_shape():
    g = grgl_amd.grg_from_coalescent(1000, 10000, 1)
    st = g.graph_stats()
    assert g.num_samples == 1000 and g.num_mutations == 10000 and g.num_individuals == 500
    assert 1000 < g.num_nodes < 3000 and st["up_levels"] > 5
    # Single tree: every sample has at most one parent, every node at most one parent.
    counts = np.bincount(g.arrays()["down_children"], minlength=g.num_nodes)
    assert counts.max() <= 1


def test_generator_coal_counts_match_xtx_definition():
    """Per node, 2*coals + sum(children) reproduces X^T X diagonal terms (test_matmul.py:110-143)."""
    g = grgl_amd.grg_from_coalescent(64, 200, 1, seed=3)
    a = g.arrays()
    og = oracle.OracleGraph(a)
    out = og.matmul(np.ones((1, 64), np.int64), oracle.UP, init="xtx")[0]
    off, child = a["down_offsets"], a["down_children"]
    xx = np.zeros(g.num_nodes, np.int64)
    xx[:64] = 1
    for n in range(64, g.num_nodes):
        xx[n] = sum(xx[c] for c in child[off[n]:off[n + 1]]) + 2 * int(a["coal_counts"][n])
    for m, node in enumerate(a["mutation_nodes"]):
        assert out[m] == xx[node]
    # The count of X^T X: carriers + 2 * (individuals with both haplotypes carrying).
    X = np.zeros((64, g.num_mutations), np.int64)
    from helpers import dense_x
    X = dense_x(a).astype(np.int64)
    hom = (X.reshape(32, 2, -1).sum(axis=1) == 2).sum(axis=0)
    np.testing.assert_array_equal(out, X.sum(axis=0) + 2 * hom)


def test_grg_from_arrays_rejects_unordered():
    bad = depth3_bin_tree([1])
    bad["down_children"] = np.array([0, 1, 2, 3, 4, 6], np.uint32)  # node 6 lists itself
    with pytest.raises(RuntimeError):
        grgl_amd.grg_from_arrays(bad["num_samples"], 2, bad["down_offsets"], bad["down_children"],
                                 bad["mutation_nodes"])


def _execute_plan(arrays, plan, direction, x, dtype):
    """Run a level plan on the CPU: every row summed front to back over its gather list."""
    N = arrays["num_samples"]
    V = len(arrays["down_offsets"]) - 1
    M = len(arrays["mutation_nodes"])
    K = x.shape[0]
    rows = np.zeros((V + 2 * M, K), dtype=dtype)
    off, src = plan["off"], plan["src"]
    if direction == grgl_amd.UP:
        rows[:N] = x.T
        first = plan["first_pulled_level"]
    else:
        rows[V:V + M] = x.T
        first = 0
    ls = plan["level_start"]
    for lvl in range(first, plan["num_levels"]):
        for j in range(ls[lvl], ls[lvl + 1]):
            acc = np.zeros(K, dtype=dtype)
            for s in src[off[j]:off[j + 1]]:
                acc = (acc + rows[s]).astype(dtype)
            rows[j] = acc
    out_row = plan["out_row"]
    out = np.zeros((K, len(out_row)), dtype=dtype)
    for c, r in enumerate(out_row):
        if r != INVALID:
            out[:, c] = rows[r]
    return out


@pytest.mark.parametrize("which", ["random", "tree", "vcf"])
def test_level_plan_reproduces_oracle_bitwise(which):
    if which == "random":
        arrays = random_grg(31, num_samples=40, num_internal=150, heavy=(80,), num_mutations=200)
    elif which == "tree":
        arrays = grgl_amd.grg_from_coalescent(300, 2000, 2, seed=2).arrays()
    else:
        arrays = direct_map_grg(vcf_fixture())
    g = grgl_amd.grg_from_arrays(arrays["num_samples"], arrays["ploidy"], arrays["down_offsets"],
                                 arrays["down_children"], arrays["mutation_nodes"])
    og = oracle.OracleGraph(arrays)
    rng = np.random.default_rng(1)
    for direction in (grgl_amd.UP, grgl_amd.DOWN):
        plan = grgl_amd._grgl_amd._level_plan(g, direction)
        ls = plan["level_start"]
        level_of = np.zeros(g.num_nodes, np.int64)
        for lvl in range(plan["num_levels"]):
            level_of[ls[lvl]:ls[lvl + 1]] = lvl
            fan = np.diff(plan["off"][ls[lvl]:ls[lvl + 1] + 1].astype(np.int64))
            if lvl >= plan["first_pulled_level"]:
                assert np.all(np.diff(fan) <= 0), "rows not sorted by list length"
        # Every gathered node row belongs to an earlier level.
        for lvl in range(plan["first_pulled_level"], plan["num_levels"]):
            for j in range(ls[lvl], ls[lvl + 1]):
                for s in plan["src"][plan["off"][j]:plan["off"][j + 1]]:
                    if s < g.num_nodes:
                        assert level_of[s] < lvl
        d = oracle.UP if direction == grgl_amd.UP else oracle.DOWN
        cols = g.num_samples if direction == grgl_amd.UP else g.num_mutations
        for dt in (np.float32, np.float64):
            x = make_input(rng, dt, (2, cols))
            assert bits_equal(_execute_plan(arrays, plan, direction, x, dt), og.matmul(x, d))


def _expect_runtime_error(fn, fragment=None):
    with pytest.raises(RuntimeError) as ei:
        fn()
    if fragment is not None:
        assert fragment in str(ei.value), str(ei.value)


def test_pygrgl_argument_checks():
    """Argument errors of _grgl.cpp:100-241 surface as RuntimeError before any device work."""
    g = grgl_amd.grg_from_arrays(4, 2, *[depth3_bin_tree([1, 5, 6])[k] for k in
                                         ("down_offsets", "down_children", "mutation_nodes")])
    up = grgl_amd.UP
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones(4), up), "two dimensional")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((0, 4)), up), "non-zero")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4), np.complex64), up), "Unsupported numpy dtype")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4)), up, init="test"), "Unexpected init value")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4), np.int32), up, init=np.ones((1, 4))),
                          "must match the dtype")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((9, 4)), up, init=np.ones(10)), "single dimension")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4)), up, init=np.ones((2, 4))), "ROW x NODES")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4), bool), up, init=np.ones(1, bool)),
                          "boolean")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4)), up, emit_all_nodes=True,
                                                  miss=np.zeros((1, 7))), "emit_all_nodes")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4)), up, miss=np.zeros((2, 3))), "rows")
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 4)), up, miss=np.zeros((1, 4))),
                          "columns in the output")
    _expect_runtime_error(lambda: grgl_amd.dot_product(g, np.ones((1, 4)), up), "one-dimensional")
    _expect_runtime_error(lambda: grgl_amd.dot_product(g, np.ones(4, np.int8), up), "Unsupported numpy dtype")


def test_pygrgl_properties():
    arrays = depth3_bin_tree([1, 5, 6], coal=[0, 0, 0, 0, 1, 1, 0])
    g = grgl_amd.grg_from_arrays(4, 2, arrays["down_offsets"], arrays["down_children"], arrays["mutation_nodes"],
                                 coal_counts=arrays["coal_counts"])
    assert (g.num_samples, g.num_nodes, g.num_edges, g.num_mutations, g.num_individuals) == (4, 7, 6, 3, 2)
    assert g.shape == (2, 3) and g.ploidy == 2
    assert g.get_down_edges(6) == [4, 5] and g.get_up_edges(4) == [6] and g.num_up_edges(6) == 0
    assert g.get_mutations_for_node(5) == [1] and g.get_mutation_node_pairs() == [(0, 1), (1, 5), (2, 6)]
    assert g.is_sample(3) and not g.is_sample(4)
    assert g.get_num_individual_coals(4) == 1 and g.get_num_individual_coals(2) == 0
    assert g.has_individual_coals and not g.has_missing_data


@pytest.mark.skipif(grgl_amd.device_count() > 0, reason="checks the no-GPU failure mode")
def test_no_gpu_fails_loudly():
    g = grgl_amd.grg_from_coalescent(100, 200)
    _expect_runtime_error(lambda: grgl_amd.matmul(g, np.ones((1, 100)), grgl_amd.UP))


@pytest.mark.parametrize("es", [1, 2, 4, 8])
def test_stacked_requests_never_share_a_vector(es):
    """Batched requests: a request >= 16 bytes wide starts 16-byte aligned and owns whole
    16-byte slots, so its vector accesses past its last row touch no other request."""
    rng = np.random.default_rng(es)
    stack = grgl_amd._grgl_amd._stack_offsets
    off, total = stack([1, 2, 4, 8, 16, 32, 64, 128, 256], 4)
    assert total == 511
    for _ in range(200):
        rows = [int(x) for x in rng.choice([1, 2, 3, 5, 8, 17, 31, 64, 100], size=int(rng.integers(1, 12)))]
        off, total = stack(rows, es)
        v = max(1, 16 // es)
        used = np.zeros(total + 64, dtype=np.int64) - 1
        for i, (o, r) in enumerate(zip(off, rows)):
            span = r
            if r * es >= 16:
                assert o % v == 0
                span = -(-r // v) * v
            assert np.all(used[o:o + span] == -1), "overlapping slots"
            used[o:o + span] = i
        assert total == max(o + r for o, r in zip(off, rows))


def _build_shim_test(tmp_path):
    exe = tmp_path / "shim_matmul_test"
    lib_dir = os.path.join(ROOT, "grgl_amd")
    import subprocess

    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "cpp", "shim_matmul_test.cpp"), f"-L{lib_dir}", "-lgrgl_hip",
                    f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    return exe


def test_cpp_shim_compiles_and_links(tmp_path):
    """include/grgl_hip.hpp (GRG::matMul / matrixMultiplication surface) builds against the library."""
    assert _build_shim_test(tmp_path).exists()


def test_graph_create_rejects_missing_arrays():
    """grgl_hip_graph_create checks its descriptor before any device work: missing offsets,
    children (when there are edges) or mutation nodes (when there are mutations) are misuse."""
    off = np.array([0, 0, 0, 2], np.uint64)
    child = np.array([0, 1], np.uint32)
    muts = np.array([2], np.uint32)
    L = capi.lib()
    cases = [(None, child, muts, "down_offsets"), (off, None, muts, "down_children"), (off, child, None, "mutation_nodes")]
    for o, c, m, what in cases:
        d = capi.GraphDesc(2, 3, 2, 1, None if o is None else o.ctypes.data, None if c is None else c.ctypes.data,
                           None if m is None else m.ctypes.data, None, None)
        out = ctypes.c_void_p()
        assert L.grgl_hip_graph_create(ctypes.byref(d), 0, ctypes.byref(out)) == capi.ERR_MISUSE
        assert what in L.grgl_hip_last_error().decode()
        assert not out.value


def test_import_does_not_load_torch():
    """`import grgl_amd` loads neither torch nor anything that needs it; with GRGL_AMD_NO_TORCH=1
    the numpy surface (generator, arrays, .grg files) works the same."""
    import subprocess
    import sys

    code = ("import sys, os, tempfile\n"
            "import grgl_amd\n"
            "assert 'torch' not in sys.modules, 'torch imported'\n"
            "g = grgl_amd.grg_from_coalescent(200, 1000, 1, seed=3)\n"
            "a = g.arrays()\n"
            "d = tempfile.mkdtemp()\n"
            "grgl_amd.save_grg(g, os.path.join(d, 'x.grg'))\n"
            "h = grgl_amd.load_immutable_grg(os.path.join(d, 'x.grg'))\n"
            "assert (h.num_nodes, h.num_edges) == (g.num_nodes, g.num_edges)\n"
            "assert 'torch' not in sys.modules\n"
            "print('ok')\n")
    for env_extra in ({}, {"GRGL_AMD_NO_TORCH": "1"}):
        env = dict(os.environ, **env_extra)
        proc = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, stdout=subprocess.PIPE,
                              stderr=subprocess.STDOUT, text=True, timeout=120)
        assert proc.returncode == 0 and proc.stdout.strip().endswith("ok"), proc.stdout
