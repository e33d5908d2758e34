"""TPC-H Q18 as end_to_end_test/tpch/q18_test.cc plans it: the subquery
GROUP BY l_orderkey SUM(l_quantity) with Select(sum > 315.0), joined with
orders (l_orderkey = o_orderkey), customer (o_custkey = c_custkey) and
lineitem again (o_orderkey = l_orderkey), GROUP BY (c_name, c_custkey,
o_orderkey, o_orderdate, o_totalprice) SUM(l_quantity), ORDER BY
o_totalprice DESC, o_orderdate. On the fixture data no order passes the
HAVING bar, so every later operator sees no rows and the result is empty;
oracle_rows restates the whole plan for any bar, so that a lower bar sends
rows through the joins, the five-column GROUP BY and the ORDER BY."""
import numpy as np

import kdb
from oracle import pyoracle as po

BAR = 315.0


def columns():
    return {name: kdb.col(name, typ)[0] for name, typ in [("l_orderkey", "int"), ("l_quantity", "real")]}


def oracle_subquery(c):
    """(order keys, SUM(l_quantity)) per group in insertion order."""
    kv, _, vals, _ = po.group_by([po.Col(po.INT, c["l_orderkey"])], [(po.SUM, po.Col(po.REAL, c["l_quantity"]))],
                                 max_groups=len(c["l_orderkey"]) + 1)
    return kv[:, 0], vals[:, 0].view(np.float64)


def oracle_q18(c):
    keys, sums = oracle_subquery(c)
    if not np.any(sums > BAR):
        return []
    raise AssertionError("the fixture data has orders above the bar; the rest of the plan is not restated")


def full_columns():
    """Every column the plan reads."""
    spec = [("l_orderkey", "int"), ("l_quantity", "real"), ("o_orderkey", "int"), ("o_orderdate", "date"),
            ("o_totalprice", "real"), ("o_custkey", "int"), ("c_custkey", "int")]
    c = {name: kdb.col(name, typ)[0] for name, typ in spec}
    c["c_name"] = kdb.text("c_name")[0]
    return c


def oracle_rows(c, bar):
    """The plan's result rows (c_name, c_custkey, o_orderkey, o_orderdate,
    o_totalprice, SUM(l_quantity)) in ORDER BY o_totalprice DESC, o_orderdate
    order, restated with numpy: the subquery's groups above the bar (insertion
    order), the three equi-joins (every join key is unique on one side except
    lineitem's l_orderkey), the GROUP BY over the joined rows, the sort."""
    keys, sums = oracle_subquery(c)
    sel = set(int(k) for k in keys[sums > bar])
    orow = [i for i, k in enumerate(c["o_orderkey"]) if int(k) in sel]
    cpos = {int(k): i for i, k in enumerate(c["c_custkey"])}
    groups = {}
    for o in orow:
        ci = cpos.get(int(c["o_custkey"][o]))
        if ci is None:
            continue
        lrows = np.nonzero(c["l_orderkey"] == c["o_orderkey"][o])[0]
        if len(lrows) == 0:
            continue
        key = (c["c_name"][ci], int(c["c_custkey"][ci]), int(c["o_orderkey"][o]), int(c["o_orderdate"][o]),
               float(c["o_totalprice"][o]))
        groups[key] = groups.get(key, 0.0) + float(c["l_quantity"][lrows].sum())
    rows = [k + (v,) for k, v in groups.items()]
    assert len(set((r[4], r[3]) for r in rows)) == len(rows), "ties in the ORDER BY keys"
    return sorted(rows, key=lambda r: (-r[4], r[3]))
