"""TPC-H Q5 as end_to_end_test/tpch/q05_test.cc plans it: region (r_name =
'MIDDLE EAST') JOIN nation ON r_regionkey = n_regionkey, JOIN customer ON
n_nationkey = c_nationkey, JOIN orders (o_orderdate in [1993-01-01,
1994-01-01)) ON c_custkey = o_custkey, JOIN lineitem ON o_orderkey =
l_orderkey, JOIN supplier ON (s_nationkey, s_suppkey) = (n_nationkey,
l_suppkey), GROUP BY n_name SUM(l_extendedprice * (1.0 - l_discount)),
ORDER BY revenue DESC. Rows: (n_name, revenue)."""
import numpy as np

import kdb
from oracle import pyoracle as po

LO, HI = (1993, 1, 1), (1994, 1, 1)


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("r_regionkey", "int"), ("n_nationkey", "int"), ("n_regionkey", "int"), ("c_custkey", "int"),
        ("c_nationkey", "int"), ("o_orderkey", "int"), ("o_custkey", "int"), ("o_orderdate", "date"),
        ("l_orderkey", "int"), ("l_suppkey", "int"), ("l_extendedprice", "real"), ("l_discount", "real"),
        ("s_suppkey", "int"), ("s_nationkey", "int")]}
    c["r_name"] = kdb.text("r_name")[0]
    c["n_name"] = kdb.text("n_name")[0]
    return c


def _join(bkeys, pkeys):
    pr, br = po.hash_join([po.Col(po.INT, np.ascontiguousarray(bkeys))], [po.Col(po.INT, np.ascontiguousarray(pkeys))])
    return pr, br


def oracle_q05(c):
    lo, hi = po.build_date(*LO), po.build_date(*HI)
    reg = np.array([i for i, s in enumerate(c["r_name"]) if s == "MIDDLE EAST"], dtype=np.int32)
    pr, br = _join(c["r_regionkey"][reg], c["n_regionkey"])
    nat = pr                                                     # nation rows
    pr, br = _join(c["n_nationkey"][nat], c["c_nationkey"])
    cust, cnat = pr, nat[br]                                     # customer rows, their nation rows
    ords = po.scan_select([po.Col(po.DATE, c["o_orderdate"])], [(0, po.GEQ, lo), (0, po.LT, hi)])
    pr, br = _join(c["c_custkey"][cust], c["o_custkey"][ords])
    orow, onat = ords[pr], cnat[br]
    pr, br = _join(c["o_orderkey"][orow], c["l_orderkey"])
    lrow, lnat = pr, onat[br]
    pr, br = _join(c["l_suppkey"][lrow], c["s_suppkey"])         # supplier probes the join on s_suppkey ...
    srow, jl, jn = pr, lrow[br], lnat[br]
    keep = c["s_nationkey"][srow] == c["n_nationkey"][jn]        # ... and s_nationkey = n_nationkey
    jl, jn = jl[keep], jn[keep]
    rev = c["l_extendedprice"][jl] * (1.0 - c["l_discount"][jl])
    groups = {}
    for name, r in zip((c["n_name"][i] for i in jn), rev):
        groups[name] = groups.get(name, 0.0) + r
    return sorted(groups.items(), key=lambda kv: -kv[1])


def check(rows, expected_lines, tol=1e-5):
    assert len(rows) == len(expected_lines)
    for (name, r), line in zip(rows, expected_lines):
        e = line.split("|")
        assert e[0] == name and abs(float(e[1]) - r) < tol, (line, name, r)
