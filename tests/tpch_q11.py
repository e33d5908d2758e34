"""TPC-H Q11 as end_to_end_test/tpch/q11_test.cc plans it: Select(nation:
n_name = 'GERMANY') JOIN supplier ON n_nationkey = s_nationkey, JOIN
partsupp ON s_suppkey = ps_suppkey; the subquery SUM(ps_supplycost *
Float(ps_availqty)) * 0.0001 over that join; the query GROUP BY ps_partkey of
the same SUM; cross product of the two and Select(value > value1), ORDER BY
value DESC. Rows: (ps_partkey, value)."""
import numpy as np

import kdb
from oracle import pyoracle as po


def columns():
    c = {name: kdb.col(name, typ)[0] for name, typ in [
        ("n_nationkey", "int"), ("s_suppkey", "int"), ("s_nationkey", "int"), ("ps_partkey", "int"),
        ("ps_suppkey", "int"), ("ps_supplycost", "real"), ("ps_availqty", "int")]}
    c["n_name"] = kdb.text("n_name")[0]
    return c


def _join(b, p):
    return po.hash_join([po.Col(po.INT, np.ascontiguousarray(b))], [po.Col(po.INT, np.ascontiguousarray(p))])


def joined(c):
    nat = np.array([i for i, s in enumerate(c["n_name"]) if s == "GERMANY"], dtype=np.int64)
    pr, _ = _join(c["n_nationkey"][nat], c["s_nationkey"])
    sup = pr
    pr, _ = _join(c["s_suppkey"][sup], c["ps_suppkey"])
    return pr                                                      # partsupp rows


def oracle_q11(c):
    ps = joined(c)
    v = c["ps_supplycost"][ps] * c["ps_availqty"][ps].astype(np.float64)
    total = 0.0
    for x in v:                                                    # the aggregator's sequential fold
        total += x
    bar = total * 0.0001
    kv, _, vals, _ = po.group_by([po.Col(po.INT, np.ascontiguousarray(c["ps_partkey"][ps]))],
                                 [(po.SUM, po.Col(po.REAL, v))], max_groups=len(ps) + 1)
    value = vals[:, 0].view(np.float64)
    keep = [g for g in range(len(kv)) if value[g] > bar]
    return sorted([(int(kv[g, 0]), float(value[g])) for g in keep], key=lambda t: -t[1])


def check(rows, expected_lines, tol=1e-5):
    assert len(rows) == len(expected_lines)
    for (k, v), line in zip(rows, expected_lines):
        e = line.split("|")
        assert int(e[0]) == k and abs(float(e[1]) - v) < tol, (line, k, v)
