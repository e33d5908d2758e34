"""TPC-H Q12 as end_to_end_test/tpch/q12_test.cc plans it: Select(lineitem:
(l_shipmode = 'RAIL' OR 'MAIL') AND l_commitdate < l_receiptdate AND
l_shipdate < l_commitdate AND l_receiptdate in [1994-01-01, 1995-01-01))
JOIN orders ON l_orderkey = o_orderkey, GROUP BY l_shipmode with
SUM(CASE o_orderpriority IN ('1-URGENT', '2-HIGH') THEN 1 ELSE 0) and
SUM(CASE both NEQ THEN 1 ELSE 0), ORDER BY l_shipmode. Rows: (l_shipmode,
high_line_count, low_line_count)."""
import numpy as np

import kdb
from oracle import pyoracle as po

LO, HI = (1994, 1, 1), (1995, 1, 1)


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("l_orderkey", "int"), ("l_commitdate", "date"), ("l_receiptdate", "date"), ("l_shipdate", "date"),
        ("o_orderkey", "int")]}
    c["l_shipmode"] = kdb.text("l_shipmode")[0]
    c["o_orderpriority"] = kdb.text("o_orderpriority")[0]
    return c


def oracle_q12(c):
    lo, hi = po.build_date(*LO), po.build_date(*HI)
    sm = c["l_shipmode"]
    keep = [i for i in range(len(sm)) if sm[i] in ("RAIL", "MAIL") and c["l_commitdate"][i] < c["l_receiptdate"][i]
            and c["l_shipdate"][i] < c["l_commitdate"][i] and lo <= c["l_receiptdate"][i] < hi]
    rows_l = np.array(keep, dtype=np.int32)
    # HashJoin(build: selected lineitem, probe: orders); output in probe order
    pr, br = po.hash_join([po.Col(po.INT, c["l_orderkey"][rows_l])], [po.Col(po.INT, c["o_orderkey"])])
    groups = {}
    for o, b in zip(pr, br):
        s = sm[rows_l[b]]
        p = c["o_orderpriority"][o]
        g = groups.setdefault(s, [0, 0])
        g[0] += 1 if p in ("1-URGENT", "2-HIGH") else 0
        g[1] += 1 if p != "1-URGENT" and p != "2-HIGH" else 0
    return [(s, g[0], g[1]) for s, g in sorted(groups.items(), key=lambda kv: kv[0].encode())]


def lines(rows):
    return [f"{s}|{h}|{lo}" for s, h, lo in rows]
