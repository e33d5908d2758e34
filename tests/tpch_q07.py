"""TPC-H Q7 as end_to_end_test/tpch/q07_test.cc plans it: nation (n_name =
'EGYPT' or 'KENYA') JOIN supplier (n1), nation JOIN customer (n2) JOIN orders
JOIN Select(lineitem: l_shipdate in [1995-01-01, 1996-12-31]) JOIN the
nation-supplier join ON s_suppkey = l_suppkey, Select((n1 = EGYPT AND n2 =
KENYA) OR (n1 = KENYA AND n2 = EGYPT)), GROUP BY (supp_nation, cust_nation,
EXTRACT(YEAR l_shipdate)) SUM(l_extendedprice * (1.0 - l_discount)), ORDER BY
the three keys. Rows: (supp_nation, cust_nation, l_year, revenue)."""
import numpy as np

import kdb
from oracle import pyoracle as po

LO, HI = (1995, 1, 1), (1996, 12, 31)
NATIONS = ("EGYPT", "KENYA")


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("n_nationkey", "int"), ("s_suppkey", "int"), ("s_nationkey", "int"), ("c_custkey", "int"),
        ("c_nationkey", "int"), ("o_orderkey", "int"), ("o_custkey", "int"), ("l_orderkey", "int"),
        ("l_suppkey", "int"), ("l_shipdate", "date"), ("l_extendedprice", "real"), ("l_discount", "real")]}
    c["n_name"] = kdb.text("n_name")[0]
    return c


def _join(b, p):
    return po.hash_join([po.Col(po.INT, np.ascontiguousarray(b))], [po.Col(po.INT, np.ascontiguousarray(p))])


def oracle_q07(c):
    nat = np.array([i for i, s in enumerate(c["n_name"]) if s in NATIONS], dtype=np.int64)
    pr, br = _join(c["n_nationkey"][nat], c["s_nationkey"])
    srow, snat = pr, nat[br]
    pr, br = _join(c["n_nationkey"][nat], c["c_nationkey"])
    crow, cnat = pr, nat[br]
    pr, br = _join(c["c_custkey"][crow], c["o_custkey"])
    orow, onat = pr, cnat[br]
    li = po.scan_select([po.Col(po.DATE, c["l_shipdate"])],
                        [(0, po.GEQ, po.build_date(*LO)), (0, po.LEQ, po.build_date(*HI))])
    pr, br = _join(c["o_orderkey"][orow], c["l_orderkey"][li])
    lrow, lnat = li[pr], onat[br]
    pr, br = _join(c["s_suppkey"][srow], c["l_suppkey"][lrow])
    lrow, n2, n1 = lrow[pr], lnat[pr], snat[br]
    groups = {}
    for k in range(len(lrow)):
        a, b = c["n_name"][n1[k]], c["n_name"][n2[k]]
        if not ((a == "EGYPT" and b == "KENYA") or (a == "KENYA" and b == "EGYPT")):
            continue
        y = po.split_date(int(c["l_shipdate"][lrow[k]]))[0]
        r = c["l_extendedprice"][lrow[k]] * (1.0 - c["l_discount"][lrow[k]])
        groups[(a, b, y)] = groups.get((a, b, y), 0.0) + r
    return [(a, b, y, v) for (a, b, y), v in sorted(groups.items(), key=lambda kv: (kv[0][0].encode(),
                                                                                  kv[0][1].encode(), kv[0][2]))]


def check(rows, expected_lines, tol=1e-5):
    assert len(rows) == len(expected_lines)
    for (a, b, y, v), line in zip(rows, expected_lines):
        e = line.split("|")
        assert e[0] == a and e[1] == b and int(e[2]) == y and abs(float(e[3]) - v) < tol, (line, a, b, y, v)
