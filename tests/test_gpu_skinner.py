"""Five-table chain join with selective filters (incl. TEXT prefix / LIKE),
executed on the HIP path as Skinner's per-episode work does it — filtered
children materialized as selections, column indexes on the join columns,
index-nested-loop expansion one table at a time, result tuples deduplicated
in a tuple-id set — then MIN over TEXT and INT columns. Checked against the
oracle (hash joins over the same filtered children).

Bars: result tuple set identical (tuple ids = positions in the filtered
children, as in the reference); MIN values identical."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import (Column, ColumnIndex, Context, TextColumn, TupleSet, gather, index_join_expand,
                                text_minmax, text_select, tuple_rows)
from kushdb_amd.chain import chain_filters, chain_join, chain_query
from kushdb_amd.workloads import skinner_tables
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def gpu_filters(ctx, tabs, spec):
    return chain_filters(ctx, tabs, spec)


def gpu_chain(ctx, tabs, sels, order):
    """Join-order tuples put in table order (test-side, to compare with the
    oracle's table-order tuples)."""
    res, cols, _, _ = chain_join(ctx, tabs, sels, order)
    assert cols == list(order)
    return res[:, [cols.index(t) for t in range(len(tabs))]].contiguous()


def host_filters(ht, spec):
    """Selection of one table on the CPU (oracle scan-select / TEXT select)."""
    rows = None
    for kind, col, op, lit in spec:
        if kind == "text":
            rows = po.text_select(po.TextCol(ht[col + "_off"], ht[col + "_heap"]), op, lit, rows)
        elif rows is None:
            rows = po.scan_select([po.Col(po.INT, ht[col])], [(0, op, lit)])
        else:
            keep = po.scan_select([po.Col(po.INT, ht[col][rows])], [(0, op, lit)])
            rows = rows[keep]
    if rows is None:
        rows = np.arange(len(ht["id"]), dtype=np.int32)
    return rows


def oracle_chain(htabs, hsels):
    """Positions of all 5-tuples satisfying the chain joins over the selections."""
    # start from table 0 and extend with hash joins (build on the new table)
    cur = np.arange(len(hsels[0]), dtype=np.int32).reshape(-1, 1)
    for t in range(1, len(htabs)):
        prev_rows = hsels[t - 1][cur[:, t - 1]]
        probe = po.Col(po.INT, htabs[t - 1]["id"][prev_rows])
        build = po.Col(po.INT, htabs[t]["fk"][hsels[t]])
        pr, br = po.hash_join([build], [probe])
        cur = np.concatenate([cur[pr], br.reshape(-1, 1)], axis=1)
    return cur


def run_case(scale, theta, order):
    ctx = Context(0)
    tabs, spec, htabs = skinner_tables(scale=scale, theta=theta, device=DEV, seed=11 + int(theta * 10))
    sels = gpu_filters(ctx, tabs, spec)
    res = gpu_chain(ctx, tabs, sels, order)
    ts = TupleSet(ctx, 5, max(res.shape[0], 1))
    is_new = ts.insert(res)
    assert int(is_new.sum().item()) == res.shape[0]  # a chain join yields no duplicates
    got = ts.extract().cpu().numpy()
    # oracle
    hsels = [host_filters(ht, sp) for ht, sp in zip(htabs, spec)]
    for a, b in zip(sels, hsels):
        assert np.array_equal(np.arange(len(b), dtype=np.int32) if a is None else a.cpu().numpy(), b)
    want = oracle_chain(htabs, hsels)
    g = got[np.lexsort(got.T[::-1])]
    w = want[np.lexsort(want.T[::-1])]
    assert np.array_equal(g, w)
    # MIN(t0.name), MIN(t1.s), MIN(t4.v)
    if len(w):
        r0 = tuple_rows(ctx, res, 0, sels[0])
        pos, val = text_minmax(ctx, tabs[0]["name"], r0)
        names = htabs[0]["name_list"]
        assert val.decode() == min(names[r] for r in hsels[0][w[:, 0]])
        r4 = tuple_rows(ctx, res, 4, sels[4])
        v4 = gather(ctx, tabs[4]["v"], r4)
        _, vals, _ = ctx.scan_group_agg([v4], [], [], [(abi.KG_AGG_MIN, 0)])
        assert vals[0, 0] == htabs[4]["v"][hsels[4][w[:, 4]]].min()
    return len(w)


@pytest.mark.parametrize("theta", [0.0, 1.0])
def test_chain_join_small(theta):
    assert run_case(0.01, theta, [0, 1, 2, 3, 4]) > 0


def test_chain_join_orders_agree():
    a = run_case(0.01, 0.0, [4, 3, 2, 1, 0])
    b = run_case(0.01, 0.0, [2, 1, 3, 0, 4])
    assert a == b


def test_chain_query_matches_oracle():
    """The packaged pipeline (filters, chain join, dedup, MIN) on the same data."""
    ctx = Context(0)
    tabs, spec, htabs = skinner_tables(scale=0.01, theta=1.0, device=DEV, seed=21)
    tuples, mn_text, mn_int, stats = chain_query(ctx, tabs, spec, [0, 1, 2, 3, 4])
    hsels = [host_filters(ht, sp) for ht, sp in zip(htabs, spec)]
    want = oracle_chain(htabs, hsels)
    got = tuples.cpu().numpy()
    assert np.array_equal(got[np.lexsort(got.T[::-1])], want[np.lexsort(want.T[::-1])])
    assert stats["tuples"] == len(want) and stats["selected"] == [len(h) for h in hsels]
    if len(want):
        assert mn_text.decode() == min(htabs[0]["name_list"][r] for r in hsels[0][want[:, 0]])
        assert mn_int == htabs[4]["v"][hsels[4][want[:, 4]]].min()
