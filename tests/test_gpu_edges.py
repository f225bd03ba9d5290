"""Edge cases the reference's matmul accepts or rejects (src/python/_grgl.cpp:100-241,
include/grgl/internal_mult.h:130-158): graphs without mutations or internal nodes, unmapped
mutations, a single haploid sample, the largest and one-too-large row counts, empty inputs."""

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import bits_equal, csr_from_lists, make_input

pytestmark = pytest.mark.gpu
INVALID = grgl_amd.INVALID_NODE
UP, DOWN = grgl_amd.UP, grgl_amd.DOWN


def _pg(a):
    return grgl_amd.grg_from_arrays(a["num_samples"], a["ploidy"], a["down_offsets"], a["down_children"],
                                    a["mutation_nodes"])


def _both(a, x, direction, **kw):
    got = grgl_amd.matmul(_pg(a), x, direction, **kw)
    want = oracle.OracleGraph(a).matmul(x, oracle.UP if direction == UP else oracle.DOWN, **kw)
    return got, want


def test_no_mutations(gpu):
    a = csr_from_lists(6, [[] for _ in range(6)] + [[0, 1], [2, 3, 6]], np.zeros(0, np.int64))
    g = _pg(a)
    out = grgl_amd.matmul(g, np.ones((3, 6), np.float32), UP)
    assert out.shape == (3, 0)
    out = grgl_amd.matmul(g, np.ones((2, 6), np.int32), UP, emit_all_nodes=True)
    assert out[:, 7].tolist() == [4, 4]
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(g, np.ones((1, 0), np.float64), DOWN)


def test_samples_only_graph(gpu):
    kids = [[] for _ in range(10)]
    muts = np.array([0, 3, 3, 9, INVALID, 5], np.int64)
    a = csr_from_lists(10, kids, muts)
    rng = np.random.default_rng(1)
    for dt in (np.float64, np.int32, np.bool_):
        x = make_input(rng, dt, (4, 10))
        got, want = _both(a, x, UP)
        assert bits_equal(got, want)
        w = make_input(rng, dt, (4, 6))
        got, want = _both(a, w, DOWN)
        assert bits_equal(got, want)


def test_all_mutations_unmapped(gpu):
    a = csr_from_lists(4, [[], [], [], [], [0, 1], [2, 3, 4]], np.full(5, INVALID, np.int64))
    up = grgl_amd.matmul(_pg(a), np.ones((2, 4), np.float32), UP)
    assert up.shape == (2, 5) and not up.any()
    down = grgl_amd.matmul(_pg(a), np.ones((2, 5), np.float32), DOWN)
    assert down.shape == (2, 4) and not down.any()


def test_single_haploid_sample(gpu):
    a = csr_from_lists(1, [[], [0], [1]], np.array([0, 1, 2, INVALID], np.int64), ploidy=1)
    got, want = _both(a, np.array([[2.5]]), UP)
    assert bits_equal(got, want) and got.tolist() == [[2.5, 2.5, 2.5, 0.0]]
    got, want = _both(a, np.array([[1.0, 2.0, 4.0, 8.0]]), DOWN)
    assert bits_equal(got, want) and got.tolist() == [[7.0]]


def test_row_limits(gpu):
    a = csr_from_lists(4, [[], [], [], [], [0, 1], [2, 3], [4, 5]], np.array([4, 6, 1], np.int64))
    x = np.ones((65536, 4), np.int8)
    got, want = _both(a, x, UP)
    assert bits_equal(got, want)
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(_pg(a), np.ones((65537, 4), np.int8), UP)
    with pytest.raises(RuntimeError):
        grgl_amd.matmul(_pg(a), np.ones((0, 4), np.int8), UP)
