"""TPC-H Q19 as end_to_end_test/tpch/q19_test.cc plans it: ScanSelect(part:
one of three (brand, container set, size range) conjunctions) JOIN
ScanSelect(lineitem: shipinstruct = 'DELIVER IN PERSON' AND shipmode IN
('AIR', 'AIR REG') AND quantity in one of three ranges) ON p_partkey =
l_partkey, then Select(joined rows: one of the three part conjunctions
together with its own quantity range), ungrouped SUM(l_extendedprice *
(1.0 - l_discount)). On the fixture data no pair qualifies, so the result is
the NULL of an aggregate over no rows (aggregate_translator.cc:82-88)."""
import numpy as np

import kdb
from oracle import pyoracle as po

PART = [("Brand#31", ("SM CASE", "SM BOX", "SM PACK", "SM PKG"), 5),
        ("Brand#52", ("MED BAG", "MED BOX", "MED PKG", "MED PACK"), 10),
        ("Brand#42", ("LG CASE", "LG BOX", "LG PACK", "LG PKG"), 15)]
QTY = [(3.0, 13.0), (12.0, 22.0), (25.0, 35.0)]


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("p_partkey", "int"), ("p_size", "int"), ("l_partkey", "int"), ("l_quantity", "real"),
        ("l_extendedprice", "real"), ("l_discount", "real")]}
    for t in ("p_brand", "p_container", "l_shipinstruct", "l_shipmode"):
        c[t] = kdb.text(t)[0]
    return c


def oracle_q19(c):
    """Returns the SUM, or None for the NULL of no qualifying rows."""
    part = [i for i in range(len(c["p_partkey"]))
            if any(c["p_brand"][i] == b and c["p_container"][i] in ct and 1 <= c["p_size"][i] <= hi
                   for b, ct, hi in PART)]
    li = [i for i in range(len(c["l_partkey"]))
          if c["l_shipinstruct"][i] == "DELIVER IN PERSON" and c["l_shipmode"][i] in ("AIR", "AIR REG")
          and any(lo <= c["l_quantity"][i] <= hi for lo, hi in QTY)]
    pr, br = po.hash_join([po.Col(po.INT, np.ascontiguousarray(c["p_partkey"][part]))],
                          [po.Col(po.INT, np.ascontiguousarray(c["l_partkey"][li]))])
    prow, lrow = np.array(part, dtype=np.int64)[br], np.array(li, dtype=np.int64)[pr]
    keep = [k for k in range(len(pr)) if any(
        c["p_brand"][prow[k]] == b and c["p_container"][prow[k]] in ct and 1 <= c["p_size"][prow[k]] <= hi
        and lo <= c["l_quantity"][lrow[k]] <= qhi for (b, ct, hi), (lo, qhi) in zip(PART, QTY))]
    if not keep:
        return None
    rows = lrow[keep]
    return float(sum(c["l_extendedprice"][rows] * (1.0 - c["l_discount"][rows])))


def line(v):
    return "|" if v is None else f"{v}|"
