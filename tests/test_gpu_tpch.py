"""TPC-H Q3 end to end on the device operators, composed as the reference's
q03_test.cc plan: TEXT select (kg_text_select), scan-selects (kg_scan_select),
two hash joins (kg_ht_build / kg_ht_probe over gathered keys), the derived
revenue (kg_expr_eval), GROUP BY over three hashed keys (kg_hash_group_agg)
and ORDER BY revenue DESC, o_orderdate (kg_order_by). Checked against the
reference's q03_expected.tbl (CHECK_EQ_TBL: REAL within 1e-5) and against the
oracle's plan (same rows in the same order; revenue within 1e-12 relative)."""
import numpy as np
import pytest
import torch

import kdb
import tpch_q03
from kushdb_amd import abi
from kushdb_amd.expr import ColRef, Literal, Mul, Sub, expr_eval
from kushdb_amd.runtime import Column, Context, HashTable, TextColumn, gather, order_by, text_select
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def dev(v, typ):
    return Column(typ, torch.from_numpy(np.ascontiguousarray(v)).to(DEV))


def device_q03(ctx, c):
    d = po.build_date(*tpch_q03.DATE)
    seg = TextColumn.from_strings(c["c_mktsegment"], device=DEV)
    rows_c = text_select(ctx, seg, abi.KG_TEXT_EQ, "AUTOMOBILE")
    odate = dev(c["o_orderdate"], abi.KG_DATE)
    rows_o = ctx.scan_select([odate], [(0, abi.KG_LT, d)])
    # customer JOIN orders ON c_custkey = o_custkey (build: selected customers)
    ck = gather(ctx, dev(c["c_custkey"], abi.KG_INT), rows_c)
    ok = gather(ctx, dev(c["o_custkey"], abi.KG_INT), rows_o)
    ht = HashTable(ctx, max(len(ck), 1))
    ht.insert(ck)
    pi, _ = ht.probe(ok)
    ht.free()
    co = rows_o[pi.long()].contiguous()                    # orders rows of the join
    # ... JOIN lineitem ON o_orderkey = l_orderkey (build: customer-orders)
    rows_l = ctx.scan_select([dev(c["l_shipdate"], abi.KG_DATE)], [(0, abi.KG_GT, d)])
    bk = gather(ctx, dev(c["o_orderkey"], abi.KG_INT), co)
    pk = gather(ctx, dev(c["l_orderkey"], abi.KG_INT), rows_l)
    ht = HashTable(ctx, max(len(bk), 1))
    ht.insert(bk)
    pj, bj = ht.probe(pk)
    ht.free()
    li = rows_l[pj.long()].contiguous()
    orow = co[bj.long()].contiguous()
    k1 = gather(ctx, dev(c["l_orderkey"], abi.KG_INT), li)
    k2 = gather(ctx, odate, orow)
    k3 = gather(ctx, dev(c["o_shippriority"], abi.KG_INT), orow)
    ext = gather(ctx, dev(c["l_extendedprice"], abi.KG_REAL), li)
    disc = gather(ctx, dev(c["l_discount"], abi.KG_REAL), li)
    rev = expr_eval(ctx, len(li), [ext, disc],
                    Mul(ColRef(0, abi.KG_REAL), Sub(Literal(1.0), ColRef(1, abi.KG_REAL))))
    first, vals, _ = ctx.hash_group_agg([k1, k2, k3, Column(abi.KG_REAL, rev.data)], [], [0, 1, 2],
                                        [(abi.KG_AGG_SUM, 3)], max_groups=len(li) + 1)
    frows = torch.from_numpy(first).to(DEV)
    g1, g2, g3 = gather(ctx, k1, frows), gather(ctx, k2, frows), gather(ctx, k3, frows)
    grev = Column(abi.KG_REAL, torch.from_numpy(vals[:, 0].view(np.float64).copy()).to(DEV))
    perm = order_by(ctx, [grev, g2], [False, True]).cpu().numpy()
    a1, a2, a3 = g1.data.cpu().numpy(), g2.data.cpu().numpy(), g3.data.cpu().numpy()
    r = grev.data.cpu().numpy()
    return [(int(a1[i]), float(r[i]), int(a2[i]), int(a3[i])) for i in perm]


def test_tpch_q03():
    ctx = Context(0)
    c = tpch_q03.columns()
    got = device_q03(ctx, c)
    tpch_q03.check(got, kdb.expected("tpch/q03_expected.tbl"))
    want = tpch_q03.oracle_q03(c)
    assert [(k, d, p) for k, _, d, p in got] == [(k, d, p) for k, _, d, p in want]
    assert np.allclose([r for _, r, _, _ in got], [r for _, r, _, _ in want], rtol=1e-12, atol=0)
    ctx.close()
