"""TPC-H Q10 as end_to_end_test/tpch/q10_test.cc plans it: Select(orders:
o_orderdate in [1995-01-01, 1995-04-01)) JOIN customer ON o_custkey =
c_custkey, JOIN nation ON n_nationkey = c_nationkey, JOIN Select(lineitem:
l_returnflag = 'R') ON o_orderkey = l_orderkey, GROUP BY (c_custkey,
c_name, c_acctbal, c_phone, n_name, c_address, c_comment)
SUM(l_extendedprice * (1.0 - l_discount)), ORDER BY revenue DESC. Rows:
(c_custkey, c_name, revenue, c_acctbal, n_name, c_address, c_phone,
c_comment)."""
import numpy as np

import kdb
from oracle import pyoracle as po

LO, HI = (1995, 1, 1), (1995, 4, 1)
TEXT = ("c_name", "c_phone", "c_address", "c_comment", "n_name", "l_returnflag")


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("o_orderdate", "date"), ("o_custkey", "int"), ("o_orderkey", "int"), ("c_custkey", "int"),
        ("c_nationkey", "int"), ("c_acctbal", "real"), ("n_nationkey", "int"), ("l_orderkey", "int"),
        ("l_extendedprice", "real"), ("l_discount", "real")]}
    for t in TEXT:
        c[t] = kdb.text(t)[0]
    return c


def _join(b, p):
    return po.hash_join([po.Col(po.INT, np.ascontiguousarray(b))], [po.Col(po.INT, np.ascontiguousarray(p))])


def oracle_q10(c):
    ords = po.scan_select([po.Col(po.DATE, c["o_orderdate"])],
                          [(0, po.GEQ, po.build_date(*LO)), (0, po.LT, po.build_date(*HI))])
    pr, br = _join(c["o_custkey"][ords], c["c_custkey"])          # build: orders, probe: customer
    crow, orow = pr, ords[br]
    pr, br = _join(c["n_nationkey"], c["c_nationkey"][crow])        # build: nation
    crow, orow, nrow = crow[pr], orow[pr], br
    li = np.array([i for i, f in enumerate(c["l_returnflag"]) if f == "R"], dtype=np.int64)
    pr, br = _join(c["o_orderkey"][orow], c["l_orderkey"][li])
    lrow, crow, nrow = li[pr], crow[br], nrow[br]
    rev = c["l_extendedprice"][lrow] * (1.0 - c["l_discount"][lrow])
    groups = {}
    for k, r in zip(range(len(lrow)), rev):
        cr, nr = crow[k], nrow[k]
        key = (int(c["c_custkey"][cr]), c["c_name"][cr], float(c["c_acctbal"][cr]), c["c_phone"][cr],
               c["n_name"][nr], c["c_address"][cr], c["c_comment"][cr])
        groups[key] = groups.get(key, 0.0) + r
    out = [(k[0], k[1], v, k[2], k[4], k[5], k[3], k[6]) for k, v in groups.items()]
    return sorted(out, key=lambda t: -t[2])


def check(rows, expected_lines, tol=1e-5):
    """CHECK_EQ_TBL: same rows in order; REAL columns within the tolerance."""
    assert len(rows) == len(expected_lines)
    for row, line in zip(rows, expected_lines):
        e = line.split("|")
        assert int(e[0]) == row[0] and e[1] == row[1] and e[4:] == list(row[4:]), (line, row)
        assert abs(float(e[2]) - row[2]) < tol and abs(float(e[3]) - row[3]) < tol, (line, row)
