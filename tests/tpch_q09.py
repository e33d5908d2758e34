"""TPC-H Q9 as end_to_end_test/tpch/q09_test.cc plans it: nation JOIN
supplier, Select(part: CONTAINS(p_name, 'moccasin')) JOIN partsupp, the two
joined ON s_suppkey = ps_suppkey, JOIN lineitem ON (p_partkey, s_suppkey) =
(l_partkey, l_suppkey), JOIN orders ON l_orderkey = o_orderkey, GROUP BY
(n_name, EXTRACT(YEAR o_orderdate)) SUM(l_extendedprice * (1.0 - l_discount)
- ps_supplycost * l_quantity), ORDER BY nation, o_year DESC. Rows: (nation,
o_year, sum_profit)."""
import numpy as np

import kdb
from oracle import pyoracle as po


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("n_nationkey", "int"), ("s_suppkey", "int"), ("s_nationkey", "int"), ("p_partkey", "int"),
        ("ps_partkey", "int"), ("ps_suppkey", "int"), ("ps_supplycost", "real"), ("l_partkey", "int"),
        ("l_suppkey", "int"), ("l_orderkey", "int"), ("l_quantity", "real"), ("l_extendedprice", "real"),
        ("l_discount", "real"), ("o_orderkey", "int"), ("o_orderdate", "date")]}
    c["n_name"] = kdb.text("n_name")[0]
    c["p_name"] = kdb.text("p_name")[0]
    return c


def _join(b, p):
    return po.hash_join([po.Col(po.INT, np.ascontiguousarray(b))], [po.Col(po.INT, np.ascontiguousarray(p))])


def oracle_q09(c):
    pr, br = _join(c["n_nationkey"], c["s_nationkey"])
    srow, snat = pr, br
    part = np.array([i for i, s in enumerate(c["p_name"]) if "moccasin" in s], dtype=np.int64)
    pr, br = _join(c["p_partkey"][part], c["ps_partkey"])
    psrow, prow = pr, part[br]
    pr, br = _join(c["s_suppkey"][srow], c["ps_suppkey"][psrow])
    psrow, prow, srow2, nrow = psrow[pr], prow[pr], srow[br], snat[br]
    pr, br = _join(c["p_partkey"][prow], c["l_partkey"])            # then l_suppkey = s_suppkey
    keep = c["l_suppkey"][pr] == c["s_suppkey"][srow2[br]]
    lrow, psrow, nrow = pr[keep], psrow[br][keep], nrow[br][keep]
    pr, br = _join(c["l_orderkey"][lrow], c["o_orderkey"])
    orow, lrow, psrow, nrow = pr, lrow[br], psrow[br], nrow[br]
    amount = (c["l_extendedprice"][lrow] * (1.0 - c["l_discount"][lrow])
              - c["ps_supplycost"][psrow] * c["l_quantity"][lrow])
    groups = {}
    for k in range(len(lrow)):
        key = (c["n_name"][nrow[k]], po.split_date(int(c["o_orderdate"][orow[k]]))[0])
        groups[key] = groups.get(key, 0.0) + amount[k]
    return [(n, y, v) for (n, y), v in sorted(groups.items(), key=lambda kv: (kv[0][0].encode(), -kv[0][1]))]


def check(rows, expected_lines, tol=1e-5):
    assert len(rows) == len(expected_lines)
    for (n, y, v), line in zip(rows, expected_lines):
        e = line.split("|")
        assert e[0] == n and int(e[1]) == y and abs(float(e[2]) - v) < tol, (line, n, y, v)
