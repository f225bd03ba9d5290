"""The FP32 rounding bound used by the GPU parity tests (oracle/fp_bounds.py), pinned on the CPU:
correct FP32 evaluations meet it (the reference's own sequential FP32 result, the oracle's FP32
restatement, the exact result rounded to FP32), and results with a planted arithmetic mistake
(a dropped term, a doubled term, a term of the wrong sign) do not."""

import numpy as np
import pytest

import grgl_amd
import oracle
from helpers import random_grg
from oracle import fp_bounds

GRAPHS = {
    "tree": lambda: grgl_amd.grg_from_coalescent(1000, 10000, 1, seed=20261203).arrays(),
    "random_heavy": lambda: random_grg(7, num_samples=300, num_internal=2000, max_children=8,
                                       heavy=(90, 700, 3000, 1500), num_mutations=3000),
}


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("direction", [oracle.UP, oracle.DOWN])
def test_bound_separates_rounding_from_mistakes(name, direction):
    og = oracle.OracleGraph(GRAPHS[name]())
    rng = np.random.default_rng(3)
    cols = og.num_samples if direction == oracle.UP else og.num_mutations
    x = rng.standard_normal((3, cols)).astype(np.float32)
    cpu32 = og.matmul(x, direction)
    rep = fp_bounds.fp32_report(og, x, direction, cpu32)
    assert rep["bound_ratio"] <= 1.0, rep
    if oracle.ref_available():
        assert rep["ref_fp32_kind"] == "reference"
        assert fp_bounds.fp32_report(og, x, direction, og.ref_matmul(x, direction),
                                     with_reference=False)["bound_ratio"] <= 1.0
    rounded = og.matmul(x.astype(np.float64), direction).astype(np.float32)
    assert fp_bounds.fp32_report(og, x, direction, rounded, with_reference=False)["bound_ratio"] <= 1.0
    # Planted mistakes, applied to one input column that reaches some output.
    col = int(np.argmax(np.abs(og.matmul(np.eye(1, cols, 0, dtype=np.float64), direction)).sum() > 0))
    for scale in (0.0, 2.0, -1.0):
        bad_in = x.astype(np.float64)
        bad_in[:, col] *= scale
        bad = og.matmul(bad_in, direction).astype(np.float32)
        r = fp_bounds.fp32_report(og, x, direction, bad, with_reference=False)
        assert r["bound_ratio"] > 10.0, (scale, r)
