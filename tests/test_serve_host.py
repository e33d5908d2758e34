"""Host-side pieces of batched serving, on the CPU: the composed oracle of the
served query templates against numpy closed forms (row counts, integer sums,
float sums), and the batch arena footprint (kg_batch_footprint, host code in
libkushgpu.so that needs no GPU)."""
import ctypes
import re

import numpy as np
import pytest

from kushdb_amd import abi
from kushdb_amd.workloads import FACT_COLS, SERVE_TEMPLATES, serve_queries, serve_tables
from serve_oracle import host_tables, oracle_spec

SIZES = (2_000, 30_000)
NBUILD = 20_000


@pytest.fixture(scope="module")
def H():
    # chain tables of 1k / 5k / 20k / 125k / 250k rows
    return host_tables(serve_tables(SIZES, nbuild=NBUILD, device="cpu", chain_scale=0.005))


def _chain_closed_form(T, p0, p1, vmax):
    """(tuples, MIN(t4.v), MIN(t0.name) bytes) of the chain query: ids are the
    row numbers, so each qualifying t4 row fixes its whole chain through the
    fk columns."""
    like = re.compile(p1.replace("%", ".*").replace("_", "."), re.S)
    ok0 = np.array([x.startswith(p0) for x in T[0]["name_list"]])
    ok1 = np.array([like.fullmatch(x) is not None for x in T[1]["s_list"]])
    r4 = np.nonzero(T[4]["v"] < vmax)[0]
    r3 = T[4]["fk"][r4]
    r1 = T[2]["fk"][T[3]["fk"][r3]]
    r0 = T[1]["fk"][r1]
    m = ok1[r1] & ok0[r0]
    if not m.any():
        return 0, None, None
    return int(m.sum()), int(T[4]["v"][r4[m]].min()), min(T[0]["name_list"][i].encode() for i in r0[m])


def _mask(f, preds):
    m = np.ones(len(f["day"]), dtype=bool)
    ops = {abi.KG_EQ: np.equal, abi.KG_NEQ: np.not_equal, abi.KG_LT: np.less, abi.KG_LEQ: np.less_equal,
           abi.KG_GT: np.greater, abi.KG_GEQ: np.greater_equal}
    for c, op, lit in preds:
        m &= ops[op](f[FACT_COLS[c]], lit)
    return m


@pytest.mark.parametrize("tpl", SERVE_TEMPLATES)
def test_composed_oracle_closed_forms(H, tpl):
    for s in serve_queries(6, SIZES, NBUILD, seed=3, templates=[tpl]):
        if s["kind"] == "chain":
            name, vals, nulls = oracle_spec(s, H)
            n, vmin, wname = _chain_closed_form(H["chain"], s["p0"], s["p1"], s["vmax"])
            assert vals[0, 0] == n and name == wname
            assert (vmin is None and nulls[0, 1]) or vals[0, 1] == vmin
            continue
        f = H["facts"][s["table"]]
        kv, vals, nulls = oracle_spec(s, H)
        m = _mask(f, s["preds"])
        if s["kind"] == "join":
            B = H["dim"] if s["build"] == "dim" else H[s["build"]]
            bkey = B["pk"] if s["build"] == "dim" else B["key"]
            bm = np.ones(len(bkey), dtype=bool)
            for c, op, lit in s["build_preds"]:
                col = (B["pk"], B["cat"])[c] if s["build"] == "dim" else (B["key"], B["payload"])[c]
                bm &= {abi.KG_LT: np.less, abi.KG_GEQ: np.greater_equal}[op](col, lit)
            # build keys are unique: a probe row joins when its key is a kept build key
            keep = np.zeros(int(max(bkey.max(), f[FACT_COLS[s["probe_key"]]].max())) + 1, dtype=bool)
            keep[bkey[bm]] = True
            m &= keep[f[FACT_COLS[s["probe_key"]]]]
        ci = [j for j, a in enumerate(s["aggs"]) if a[0] == abi.KG_AGG_COUNT][0]
        assert vals[:, ci].sum() == m.sum()
        for j, a in enumerate(s["aggs"]):
            if a[0] != abi.KG_AGG_SUM or len(a) > 2:
                continue
            x = f[FACT_COLS[a[1]]][m]
            if FACT_COLS[a[1]].startswith("m"):
                got = vals[:, j].view(np.float64).sum()
                assert abs(got - x.sum()) <= 1e-9 * max(abs(x.sum()), 1.0)
            else:
                # SUM(int32) wraps in int32 per group; these sums stay far below 2^31
                assert vals[:, j].sum() == x.astype(np.int64).sum()


def test_footprint_counts_join_table():
    lib = abi.load()
    from kushdb_amd.runtime import Query  # noqa: F401  (ctypes layout of kg_query)
    q = abi.KgQuery()
    q.kind = abi.KG_Q_JOIN_AGG
    q.build_n = 1_000_000
    q.build_code_col = -1
    q.group_card = 1
    q.naggs = 2
    q.load_factor = 0.5
    b = ctypes.c_int64()
    assert lib.kg_batch_footprint(ctypes.byref(q), ctypes.byref(b)) == abi.KG_OK
    nbuckets = -(-2_000_000 // 8)
    table = nbuckets * 8 * 12                    # 8-byte keys + 4-byte row ids per slot
    dmap = (1 << 21) * 4                         # direct map: 2 x rows to a power of two, <= 2^21 entries
    assert table + dmap <= b.value < table + dmap + (1 << 16)
    q.build_n = 3_000_000                        # beyond a table-backed map: the map alone
    assert lib.kg_batch_footprint(ctypes.byref(q), ctypes.byref(b)) == abi.KG_OK
    dmap = (1 << 23) * 4                         # 2 x rows to a power of two, <= 2^25 entries
    assert dmap <= b.value < dmap + (1 << 16)
    q.build_n = 20_000_000                       # beyond any map (2^25 entries): the table
    assert lib.kg_batch_footprint(ctypes.byref(q), ctypes.byref(b)) == abi.KG_OK
    table = -(-40_000_000 // 8)
    table = table * 8 * 12
    assert table <= b.value < table * 1.01 + (1 << 16)
    q.build_n = 0
    q.kind = abi.KG_Q_SCAN_AGG
    assert lib.kg_batch_footprint(ctypes.byref(q), ctypes.byref(b)) == abi.KG_OK
    assert 0 < b.value < (1 << 16)
