"""The oracle's answer for a batched/served query spec (kushdb_amd.workloads
serve_queries): ScanSelect -> [HashJoin] -> GroupByAggregate composed from the
oracle's restated operators (oracle/pyoracle.py), and the comparison bar:
counts and integer sums exact, FP64 sums and averages within 1e-12 relative,
group order (insertion order) exact."""
import numpy as np

from oracle import pyoracle as po
from kushdb_amd import abi
from kushdb_amd.workloads import FACT_COLS

REL = 1e-12
PO_T = {"day": po.DATE, "g1": po.ENUM, "g2": po.ENUM, "fk": po.INT, "m1": po.REAL, "m2": po.REAL, "m3": po.REAL,
        "m4": po.REAL, "i1": po.INT, "i2": po.INT, "i3": po.INT, "k64": po.BIGINT, "k64b": po.BIGINT}


def host_tables(tables):
    def h(d):
        return {k: v.cpu().numpy() for k, v in d.items()}
    H = {"facts": [h(t) for t in tables["facts"]], "dim": h(tables["dim"]), "build": h(tables["build"]),
         "big": h(tables["big"])}
    if "chain" in tables:
        H["chain"] = tables["chain"]["host"]
    return H


def _agg_col(f, a, at=None):
    if a[1] < 0:
        return None
    x = f[FACT_COLS[a[1]]]
    typ = PO_T[FACT_COLS[a[1]]]
    if len(a) > 2 and a[2] >= 0:
        # price * Float(qty): the product in double
        x = x.astype(np.float64) * f[FACT_COLS[a[2]]].astype(np.float64)
        typ = po.REAL
    if at is not None:
        x = x[at]
    return po.Col(typ, x)


def oracle_spec(spec, H):
    """(keys int64 [G, nk], vals int64 [G, na], nulls [G, na]) in insertion order.
    A chain spec: (MIN(t0.name) bytes or None, [[tuples, MIN(t4.v), -1]],
    [[0, MIN null, name null]]) from the oracle's chain restatement."""
    if spec["kind"] == "chain":
        n, name, vmin = po.chain_query(H["chain"], spec["p0"].encode(), spec["p1"].encode(), spec["vmax"])
        vals = np.array([[n, vmin if vmin is not None else 0, -1]], dtype=np.int64)
        nulls = np.array([[0, vmin is None, name is None]], dtype=np.int8)
        return name, vals, nulls
    f = H["facts"][spec["table"]]
    n = len(f["day"])
    fcols = [po.Col(PO_T[c], f[c]) for c in FACT_COLS]
    rows = po.scan_select(fcols, spec["preds"]) if spec["preds"] else np.arange(n, dtype=np.int32)
    na = len(spec["aggs"])
    if spec["kind"] == "scan":
        aggs = [(a[0], _agg_col(f, a)) for a in spec["aggs"]]
        if spec["keys"]:
            keys = [po.Col(PO_T[FACT_COLS[k]], f[FACT_COLS[k]]) for k, _ in spec["keys"]]
            kv, _, vals, nulls = po.group_by(keys, aggs, rows=rows, max_groups=64)
            return kv, vals, nulls
        got, v, nl = po.aggregate(aggs, n, rows)
        if not got:
            return np.zeros((0, 0), np.int64), np.zeros((0, na), np.int64), np.zeros((0, na), np.int8)
        return np.zeros((1, 0), np.int64), v[None, :], nl[None, :]
    if spec["build"] == "dim":
        B = H["dim"]
        bcols = [po.Col(po.INT, B["pk"]), po.Col(po.ENUM, B["cat"])]
    else:
        B = H[spec["build"]]
        bcols = [po.Col(po.BIGINT, B["key"]), po.Col(po.INT, B["payload"])]
    brows = po.scan_select(bcols, spec["build_preds"]) if spec["build_preds"] else None
    pk = FACT_COLS[spec["probe_key"]]
    op, ob = po.hash_join([bcols[spec["build_key"]]], [po.Col(PO_T[pk], f[pk])], build_rows=brows, probe_rows=rows)
    aggs = [(a[0], _agg_col(f, a, at=op)) for a in spec["aggs"]]
    if spec["build_codes"] >= 0:
        codes = bcols[spec["build_codes"]].data[ob]
        if len(op) == 0:
            return np.zeros((0, 1), np.int64), np.zeros((0, na), np.int64), np.zeros((0, na), np.int8)
        kv, _, vals, nulls = po.group_by([po.Col(po.INT, codes)], aggs, max_groups=64)
        return kv, vals, nulls
    got, v, nl = po.aggregate(aggs, len(op))
    if not got or len(op) == 0:
        return np.zeros((0, 0), np.int64), np.zeros((0, na), np.int64), np.zeros((0, na), np.int8)
    return np.zeros((1, 0), np.int64), v[None, :], nl[None, :]


def is_real_agg(a):
    if a[0] == abi.KG_AGG_AVG:
        return True
    if a[0] == abi.KG_AGG_COUNT:
        return False
    return (len(a) > 2 and a[2] >= 0) or FACT_COLS[a[1]].startswith("m")


def chain_name(H, vals, nulls):
    """MIN(t0.name) bytes of a chain query result (its row in vals[0, 2])."""
    if nulls[0, 2]:
        return None
    t0 = H["chain"][0]
    r = int(vals[0, 2])
    return bytes(t0["name_heap"][t0["name_off"][r]:t0["name_off"][r + 1]])


def assert_same(spec, got, want, what="", H=None):
    """got = (status, codes, vals, nulls) from the GPU; want = oracle_spec(...)
    or another GPU result in the same 3-tuple form. Chain specs compare the
    MIN string's bytes (host tables H), not its row: equal strings may sit in
    several rows."""
    status, codes, vals, nulls = got
    wk, wv, wn = want
    assert status == abi.KG_OK, (what, spec, status)
    if spec["kind"] == "chain":
        assert len(vals) == 1 and len(wv) == 1, (what, spec, len(vals), len(wv))
        wv, wn = np.asarray(wv), np.asarray(wn)
        assert np.array_equal(nulls, wn), (what, spec, nulls, wn)
        assert vals[0, 0] == wv[0, 0] and (nulls[0, 1] or vals[0, 1] == wv[0, 1]), (what, spec, vals, wv)
        want_name = wk if (wk is None or isinstance(wk, bytes)) else chain_name(H, wv, wn)
        assert chain_name(H, vals, nulls) == want_name, (what, spec)
        return
    assert len(vals) == len(wv), (what, spec, len(vals), len(wv))
    if codes.shape[1] and len(wv):
        assert np.array_equal(codes.astype(np.int64), np.asarray(wk, dtype=np.int64)), (what, spec)
    assert np.array_equal(nulls, wn), (what, spec)
    for j, a in enumerate(spec["aggs"]):
        g, w = vals[:, j], np.asarray(wv)[:, j]
        if is_real_agg(a):
            gf, wf = g.view(np.float64), w.view(np.float64)
            assert np.allclose(gf, wf, rtol=REL, atol=0), (what, spec, j, gf, wf)
        else:
            assert np.array_equal(g, w), (what, spec, j, g, w)
