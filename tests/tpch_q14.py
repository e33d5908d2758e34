"""TPC-H Q14 as end_to_end_test/tpch/q14_test.cc plans it: Select(lineitem:
l_shipdate in [1995-12-01, 1996-01-01)) JOIN part ON l_partkey = p_partkey,
ungrouped Aggregate of SUM(CASE STARTS_WITH(p_type, 'PROMO') THEN
l_extendedprice * (1.0 - l_discount) ELSE 0.0) and SUM(l_extendedprice *
(1.0 - l_discount)), output 100.0 * agg1 / agg2."""
import numpy as np

import kdb
from oracle import pyoracle as po

LO, HI = (1995, 12, 1), (1996, 1, 1)


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("l_shipdate", "date"), ("l_extendedprice", "real"), ("l_discount", "real"), ("l_partkey", "int"),
        ("p_partkey", "int")]}
    c["p_type"] = kdb.text("p_type")[0]
    return c


def oracle_q14(c):
    rows = po.scan_select([po.Col(po.DATE, c["l_shipdate"])],
                          [(0, po.GEQ, po.build_date(*LO)), (0, po.LT, po.build_date(*HI))])
    pr, br = po.hash_join([po.Col(po.INT, np.ascontiguousarray(c["l_partkey"][rows]))], [po.Col(po.INT, c["p_partkey"])])
    li = rows[br]
    rev = c["l_extendedprice"][li] * (1.0 - c["l_discount"][li])
    promo = np.array([c["p_type"][p].startswith("PROMO") for p in pr], dtype=bool)
    a1 = sum(r if m else 0.0 for r, m in zip(rev, promo))   # sequential fold, as the aggregator adds
    a2 = sum(rev)
    return 100.0 * a1 / a2
