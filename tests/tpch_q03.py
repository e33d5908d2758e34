"""TPC-H Q3 as the reference's end_to_end_test/tpch/q03_test.cc plans it:
Select(customer: c_mktsegment = 'AUTOMOBILE') JOIN ScanSelect(orders:
o_orderdate < 1995-03-30) ON c_custkey = o_custkey, JOIN ScanSelect(lineitem:
l_shipdate > 1995-03-30) ON o_orderkey = l_orderkey, GROUP BY (l_orderkey,
o_orderdate, o_shippriority) SUM(l_extendedprice * (1.0 - l_discount)),
ORDER BY revenue DESC, o_orderdate. Rows: (l_orderkey, revenue, o_orderdate,
o_shippriority)."""
import numpy as np

import kdb
from oracle import pyoracle as po

DATE = (1995, 3, 30)


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("c_custkey", "int"), ("o_orderkey", "int"), ("o_custkey", "int"), ("o_orderdate", "date"),
        ("o_shippriority", "int"), ("l_orderkey", "int"), ("l_extendedprice", "real"), ("l_discount", "real"),
        ("l_shipdate", "date")]}
    c["c_mktsegment"] = kdb.text("c_mktsegment")[0]
    return c


def oracle_q03(c):
    """The plan on the oracle: scan-selects, two hash joins (reference output
    order: probe order, then build insertion order), the aggregate hash table,
    then ORDER BY."""
    d = po.build_date(*DATE)
    rows_c = np.array([i for i, s in enumerate(c["c_mktsegment"]) if s == "AUTOMOBILE"], dtype=np.int32)
    rows_o = po.scan_select([po.Col(po.DATE, c["o_orderdate"])], [(0, po.LT, d)])
    pr, br = po.hash_join([po.Col(po.INT, c["c_custkey"][rows_c])], [po.Col(po.INT, c["o_custkey"][rows_o])])
    co = rows_o[pr]                                   # orders rows of the customer-orders join
    rows_l = po.scan_select([po.Col(po.DATE, c["l_shipdate"])], [(0, po.GT, d)])
    pr2, br2 = po.hash_join([po.Col(po.INT, c["o_orderkey"][co])], [po.Col(po.INT, c["l_orderkey"][rows_l])])
    li, orow = rows_l[pr2], co[br2]
    rev = c["l_extendedprice"][li] * (1.0 - c["l_discount"][li])
    keys = [po.Col(po.INT, c["l_orderkey"][li]), po.Col(po.DATE, c["o_orderdate"][orow]),
            po.Col(po.INT, c["o_shippriority"][orow])]
    kv, _, vals, _ = po.group_by(keys, [(po.SUM, po.Col(po.REAL, rev))], max_groups=len(li) + 1)
    revenue = vals[:, 0].view(np.float64)
    order = sorted(range(len(kv)), key=lambda g: (-revenue[g], kv[g, 1]))
    return [(int(kv[g, 0]), float(revenue[g]), int(kv[g, 1]), int(kv[g, 2])) for g in order]


def check(rows, expected_lines, tol=1e-5):
    """CHECK_EQ_TBL (util/test_util.h:82-175): same rows in order, integers and
    dates exact, REAL within the tolerance."""
    assert len(rows) == len(expected_lines)
    for (k, r, dt, p), line in zip(rows, expected_lines):
        e = line.split("|")
        assert int(e[0]) == k and e[2] == kdb.fmt_date(dt) and int(e[3]) == p, (line, k, dt, p)
        assert abs(float(e[1]) - r) < tol, (line, r)
