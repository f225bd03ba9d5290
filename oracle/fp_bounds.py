"""ORACLE — TEST INFRASTRUCTURE ONLY.

FP32 error of a traversal result against the FP64 oracle, and an a priori rounding bound that
any correct FP32 evaluation of the graph product must meet.

Bound (first order in the unit roundoff u = 2^-24).  Each node's value is a sum of its gather
list: UP, the node's children (include/grgl/grg.h:1324-1334); DOWN, its mutations' input values
followed by its parents (grg.h:1406-1431).  Summing f terms in FP32, in any order and any
bracketing, errs by at most (f - 1) u sum|term| (Higham, "Accuracy and Stability of Numerical
Algorithms", 2nd ed., §4.2; recursive, pairwise and the 64-partial split are all such
bracketings).  That error then flows to every output the node reaches, once per path.  So an
output o of a correct FP32 evaluation satisfies

    |fp32(o) - o| <= sum over nodes n reaching o of (f_n - 1) u A_n    (+ O(u^2)),

with A_n = sum |exact value of each list term|.  The right-hand side is itself a graph product:
the oracle's traversal of a zero input with node initialisation E_n = (f_n - 1) u A_n
(init = matrix, grg.h:1369-1402) sums E over each output's cone.  Reads of the input and of
the final node values are exact copies and add nothing.

A dropped, repeated or misplaced term moves an output by about the size of that term, orders
of magnitude above this bound, so `bound_ratio <= 1` separates rounding from arithmetic mistakes.
"""

import numpy as np

import oracle

U32 = 2.0 ** -24


def _node_values(og, x64, direction):
    """Exact (FP64) value of every node, K x numNodes, in node order."""
    return og.matmul(x64, direction, emit_all_nodes=True)


def _parents_csr(og):
    off = og.off.astype(np.int64)
    child = og.child.astype(np.int64)
    V = og.num_nodes
    parent_of_entry = np.repeat(np.arange(V, dtype=np.int64), np.diff(off))
    order = np.argsort(child, kind="stable")
    poff = np.zeros(V + 1, np.int64)
    np.add.at(poff, child + 1, 1)
    poff = np.cumsum(poff)
    return poff, parent_of_entry[order]


def _segment_sums(vals, off):
    """vals: E values, off: E-segment offsets (V + 1) -> V sums (empty segments 0)."""
    cs = np.zeros(len(vals) + 1)
    np.cumsum(vals, out=cs[1:])
    return cs[off[1:]] - cs[off[:-1]]


def rounding_bound(og, x, direction, u=U32):
    """K x outCols first-order bound on |FP32 result - exact result| (plain traversal)."""
    x64 = np.ascontiguousarray(x, dtype=np.float64)
    K = x64.shape[0]
    V = og.num_nodes
    vals = _node_values(og, x64, direction)
    A = np.zeros((K, V))
    if direction == oracle.UP:
        off = og.off.astype(np.int64)
        child = og.child.astype(np.int64)
        f = np.diff(off).astype(np.float64)
        for k in range(K):
            A[k] = _segment_sums(np.abs(vals[k])[child], off)
    else:
        poff, parents = _parents_csr(og)
        mut = og.mut.astype(np.int64)
        mapped = mut != oracle.INVALID_NODE
        f = np.diff(poff).astype(np.float64)
        np.add.at(f, mut[mapped], 1.0)
        for k in range(K):
            A[k] = _segment_sums(np.abs(vals[k])[parents], poff)
            np.add.at(A[k], mut[mapped], np.abs(x64[k, mapped]))
    E = np.maximum(f - 1.0, 0.0)[None, :] * u * A
    zero = np.zeros_like(x64)
    return og.matmul(zero, direction, init=np.ascontiguousarray(E))


def normwise_rel_err(a, b):
    """max over rows of max|a - b| / max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    num = np.abs(a - b).max(axis=1)
    den = np.abs(b).max(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, num / den, num)
    return float(r.max()) if r.size else 0.0


def fp32_report(og, x32, direction, result, with_reference=True):
    """Error of an FP32 `result` of the traversal of `x32`:
    normwise      max over rows of max|result - exact| / max|exact| (exact = FP64 oracle);
    bound_ratio   max over elements of |result - exact| / rounding_bound (<= 1 for a correct
                  FP32 evaluation);
    ref_fp32_normwise  the same normwise error of the reference's own FP32 result
                  (oracle/_ref, when built), or of the oracle's FP32 restatement."""
    x32 = np.ascontiguousarray(x32, dtype=np.float32)
    exact = og.matmul(x32.astype(np.float64), direction)
    err = np.abs(np.asarray(result, np.float64) - exact)
    bound = rounding_bound(og, x32, direction) * 1.01
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(bound > 0, err / bound, np.where(err > 0, np.inf, 0.0))
    rep = {"normwise": normwise_rel_err(result, exact), "bound_ratio": float(ratio.max()) if ratio.size else 0.0}
    if with_reference:
        if oracle.ref_available():
            cpu = og.ref_matmul(x32, direction)
            rep["ref_fp32_kind"] = "reference"
        else:
            cpu = og.matmul(x32, direction)
            rep["ref_fp32_kind"] = "oracle"
        rep["ref_fp32_normwise"] = normwise_rel_err(cpu, exact)
    return rep
