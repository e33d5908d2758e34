"""TPC-H Q18 as end_to_end_test/tpch/q18_test.cc plans it: the subquery
GROUP BY l_orderkey SUM(l_quantity) with Select(sum > 315.0), joined with
orders (l_orderkey = o_orderkey), customer (o_custkey = c_custkey) and
lineitem again (o_orderkey = l_orderkey), GROUP BY (c_name, c_custkey,
o_orderkey, o_orderdate, o_totalprice) SUM(l_quantity), ORDER BY
o_totalprice DESC, o_orderdate. On the fixture data no order passes the
HAVING bar, so every later operator sees no rows and the result is empty."""
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
