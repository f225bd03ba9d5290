"""Grouped-level plans (grgl_amd/csrc/grouping.cpp) on the CPU.

Each plan is executed the way pull_group_kernel runs it (leading HBM entries first, then the
rest segment by segment from LDS slots, heavy rows as 64 partials and a fixed tree) and
compared bit for bit in float32 with the level plan executed the way the level kernels run
it.  The binding validates every plan it returns (validateGroupPlan) as well."""

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import _grgl_amd as M
from helpers import random_grg

TAG = 0x80000000
PARTIALS = 64


def _graph(arrays):
    return grgl_amd.grg_from_arrays(arrays["num_samples"], arrays["ploidy"], arrays["down_offsets"],
                                    arrays["down_children"], arrays["mutation_nodes"])


def _rows0(g, plan, direction, x):
    V, N, M_ = g.num_nodes, g.num_samples, g.num_mutations
    rows = np.zeros((V + M_, x.shape[0]), np.float32)
    if direction == grgl_amd.UP:
        rows[:N] = x.T
    else:
        rows[V:V + M_] = x.T
    return rows


def _sum_row(rows, lst, heavy):
    """One row as the level kernels compute it: in order, or 64 partials and a fixed tree."""
    K = rows.shape[1]
    if len(lst) <= heavy:
        acc = np.zeros(K, np.float32)
        for s in lst:
            acc = acc + rows[s]
        return acc
    part = [np.zeros(K, np.float32) for _ in range(PARTIALS)]
    for i, s in enumerate(lst):
        part[i % PARTIALS] = part[i % PARTIALS] + rows[s]
    st = PARTIALS // 2
    while st >= 1:
        for t in range(st):
            part[t] = part[t] + part[t + st]
        st //= 2
    return part[0]


def _run_levels(rows, plan, lo, hi, heavy):
    off, src, ls = plan["off"].astype(np.int64), plan["src"], plan["level_start"]
    for lvl in range(lo, hi):
        for j in range(ls[lvl], ls[lvl + 1]):
            rows[j] = _sum_row(rows, src[off[j]:off[j + 1]], heavy)


def _run_groups(rows, gp):
    """pull_group_kernel, block after block (blocks never read each other's rows)."""
    off, pre, src, row = gp["off"].astype(np.int64), gp["pre"], gp["src"], gp["row"]
    rem_off, rem = gp["rem_off"].astype(np.int64), gp["rem"]
    K = rows.shape[1]
    heavy = gp["heavy_fanin"]
    for b in range(gp["blocks"]):
        r0, r1 = gp["block_row"][b], gp["block_row"][b + 1]
        h0, h1 = gp["block_heavy"][b], gp["block_heavy"][b + 1]
        i0, i1 = gp["block_import"][b], gp["block_import"][b + 1]
        n, nh = r1 - r0, h1 - h0
        import_base = n + PARTIALS * nh
        slots = {}
        # Phase 0: imports, light prefixes, heavy partial prefixes (HBM rows only).
        for t in range(i1 - i0):
            slots[import_base + t] = rows[gp["import_row"][i0 + t]].copy()
        for i in range(r0, r1):
            lb, le = off[i], off[i + 1]
            if le - lb > heavy:
                continue
            acc = np.zeros(K, np.float32)
            for k in range(lb, lb + pre[i]):
                assert not int(src[k]) & TAG
                acc = acc + rows[src[k]]
            slots[i - r0] = acc
        for k in range(nh):
            i = gp["heavy_row"][h0 + k]
            lb = off[i]
            for q in range(PARTIALS):
                acc = np.zeros(K, np.float32)
                for t in range(lb + q, lb + pre[i], PARTIALS):
                    acc = acc + rows[src[t]]
                slots[n + PARTIALS * k + q] = acc
        # Phase 1: slot lists, segment by segment.
        for s in range(gp["block_seg"][b], gp["block_seg"][b + 1]):
            g0, g1 = gp["seg_row"][s], gp["seg_row"][s + 1]
            hc, hord = gp["seg_heavy"][s], gp["seg_heavy_ord"][s]
            for h in range(hc):
                i = g0 + h
                k = hord + h
                p = int(pre[i])
                part = [slots[n + PARTIALS * k + q] for q in range(PARTIALS)]
                lst = rem[rem_off[i]:rem_off[i + 1]]
                for j, slot in enumerate(lst):
                    q = (p + j) % PARTIALS
                    part[q] = part[q] + slots[int(slot)]
                st = PARTIALS // 2
                while st >= 1:
                    for t in range(st):
                        part[t] = part[t] + part[t + st]
                    st //= 2
                slots[i - r0] = part[0]
            for i in range(g0 + hc, g1):
                acc = slots[i - r0]
                for slot in rem[rem_off[i]:rem_off[i + 1]]:
                    acc = acc + slots[int(slot)]
                slots[i - r0] = acc
        for i in range(r0, r1):
            rows[row[i]] = slots[i - r0]


def _graphs():
    yield "random", random_grg(41, num_samples=60, num_internal=300, heavy=(90, 150), num_mutations=400)
    yield "tree", grgl_amd.grg_from_coalescent(400, 3000, 3, seed=7).arrays()


@pytest.mark.parametrize("name,arrays", list(_graphs()))
@pytest.mark.parametrize("heavy,lds,max_entries", [(64, 65536, 16384), (4, 65536, 16384), (8, 2048, 300)])
def test_group_plans_compute_the_level_plan_bitwise(name, arrays, heavy, lds, max_entries):
    g = _graph(arrays)
    rng = np.random.default_rng(3)
    for direction in (grgl_amd.UP, grgl_amd.DOWN):
        plan = M._level_plan(g, direction)
        plans = M._group_plans(g, direction, heavy_fanin=heavy, lds_bytes=lds, max_entries=max_entries,
                               max_blocks=16)
        cols = g.num_samples if direction == grgl_amd.UP else g.num_mutations
        x = rng.standard_normal((3, cols)).astype(np.float32)
        first, L = plan["first_pulled_level"], plan["num_levels"]
        ref = _rows0(g, plan, direction, x)
        _run_levels(ref, plan, first, L, heavy)
        grouped_any = False
        for gp in plans:
            assert gp["heavy_fanin"] == heavy
            if gp["level_begin"] >= gp["level_end"]:
                continue
            grouped_any = True
            assert gp["lds_bytes"] <= lds
            rows = _rows0(g, plan, direction, x)
            _run_levels(rows, plan, first, gp["level_begin"], heavy)
            _run_groups(rows, gp)
            _run_levels(rows, plan, gp["level_end"], L, heavy)
            assert np.array_equal(rows.view(np.uint32), ref.view(np.uint32)), (name, direction, gp["slot_bytes"])
        assert grouped_any
        # The level plan itself against the oracle (integers, exact).
        og = oracle.OracleGraph(arrays)
        xi = rng.integers(-5, 6, (2, cols)).astype(np.int64)
        d = oracle.UP if direction == grgl_amd.UP else oracle.DOWN
        rows = np.zeros((g.num_nodes + g.num_mutations, 2), np.int64)
        if direction == grgl_amd.UP:
            rows[:g.num_samples] = xi.T
        else:
            rows[g.num_nodes:] = xi.T
        off, src, ls = plan["off"].astype(np.int64), plan["src"], plan["level_start"]
        for lvl in range(first, L):
            for j in range(ls[lvl], ls[lvl + 1]):
                rows[j] = rows[src[off[j]:off[j + 1]]].sum(axis=0)
        out = np.zeros((2, len(plan["out_row"])), np.int64)
        for c, r in enumerate(plan["out_row"]):
            if r != 0x7fffffff:
                out[:, c] = rows[r]
        assert np.array_equal(out, og.matmul(xi, d))


@pytest.mark.parametrize("name,arrays", list(_graphs()))
def test_level_row_limit(name, arrays):
    """A row limit per level (the DOWN plan for wide rows takes only levels of at most 8,192
    rows) starts the grouped levels after the last larger level, ends them where the unlimited
    plan ends, and still computes the level plan bit for bit."""
    g = _graph(arrays)
    rng = np.random.default_rng(9)
    for direction in (grgl_amd.UP, grgl_amd.DOWN):
        plan = M._level_plan(g, direction)
        ls, first, L = plan["level_start"], plan["first_pulled_level"], plan["num_levels"]
        sizes = np.diff(ls)
        full = M._group_plans(g, direction, max_blocks=16)
        cols = g.num_samples if direction == grgl_amd.UP else g.num_mutations
        x = rng.standard_normal((2, cols)).astype(np.float32)
        ref = _rows0(g, plan, direction, x)
        _run_levels(ref, plan, first, L, 64)
        for limit in sorted(set(int(v) for v in sizes[first:]))[:4]:
            limited = M._group_plans(g, direction, max_blocks=16, max_level_rows=limit)
            for a, b in zip(full, limited):
                if b["level_begin"] >= b["level_end"]:
                    continue
                assert a["level_begin"] < a["level_end"] and b["level_end"] == a["level_end"]
                assert b["level_begin"] >= a["level_begin"]
                assert sizes[b["level_begin"]:b["level_end"]].max() <= limit
                rows = _rows0(g, plan, direction, x)
                _run_levels(rows, plan, first, b["level_begin"], 64)
                _run_groups(rows, b)
                _run_levels(rows, plan, b["level_end"], L, 64)
                assert np.array_equal(rows.view(np.uint32), ref.view(np.uint32)), (name, direction, limit)


def test_down_plan_puts_childless_nodes_last():
    for arrays in (random_grg(5, num_samples=50, num_internal=200), grgl_amd.grg_from_coalescent(300, 2000, 2, seed=1).arrays()):
        g = _graph(arrays)
        plan = M._level_plan(g, grgl_amd.DOWN)
        ls = plan["level_start"]
        childless = np.diff(arrays["down_offsets"]) == 0
        last = np.zeros(g.num_nodes, bool)
        last[plan["order"][ls[-2]:ls[-1]]] = True
        assert np.array_equal(last, childless)


def test_group_limits_shape_the_plans():
    """Smaller LDS budgets and entry caps never group more levels; wider slots never more than narrower."""
    g = grgl_amd.grg_from_coalescent(2000, 12000, 4, seed=11)
    for direction in (grgl_amd.UP, grgl_amd.DOWN):
        wide = M._group_plans(g, direction)
        tight = M._group_plans(g, direction, lds_bytes=4096, max_entries=512)
        spans = [p["level_end"] - p["level_begin"] if p["level_begin"] < p["level_end"] else 0 for p in wide]
        assert spans[0] >= 2
        assert all(spans[c] >= spans[c + 1] for c in range(len(spans) - 1))
        for a, b in zip(wide, tight):
            sa = a["level_end"] - a["level_begin"] if a["level_begin"] < a["level_end"] else 0
            sb = b["level_end"] - b["level_begin"] if b["level_begin"] < b["level_end"] else 0
            assert sb <= sa
            if sb:
                assert b["lds_bytes"] <= 4096
        # DOWN groups end before the level of childless nodes; UP groups run to the last level.
        plan = M._level_plan(g, direction)
        for p in wide:
            if p["level_begin"] < p["level_end"]:
                assert p["level_end"] == plan["num_levels"] - (1 if direction == grgl_amd.DOWN else 0)
