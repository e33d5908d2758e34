"""Batched execution (kg_batch_*) and the serving loop (kg_serve) on the HIP
path: every query of a mixed batch against the oracle's composed operators,
each query's result alone equal to its result in the mix, repeated submits on
one slot, unsupported queries reported per query, and both serving modes
returning the same results."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Batch, Context, Query, serve
from kushdb_amd.workloads import (BATCH_TEMPLATES, chain_indexes, make_query, serve_columns, serve_queries,
                                  serve_tables)
from serve_oracle import assert_same, host_tables, oracle_spec

pytestmark = pytest.mark.gpu

SIZES = (3_000, 40_000, 250_000)
NBUILD = 100_000
CHAIN_SCALE = 0.01          # chain tables of 2k / 10k / 40k / 250k / 500k rows


@pytest.fixture(scope="module")
def env():
    ctx = Context(0)
    tables = serve_tables(SIZES, nbuild=NBUILD, device="cuda:0", chain_scale=CHAIN_SCALE)
    torch.cuda.synchronize()
    cols = serve_columns(tables)
    cols[2]["chain_idx"] = chain_indexes(ctx, cols[2]["chain"])
    return ctx, tables, cols, host_tables(tables)


def _queries(cols, specs):
    facts, dim, build = cols
    return [make_query(s, facts, dim, build) for s in specs]


def test_batch_mix_vs_oracle(env):
    ctx, _, cols, H = env
    specs = serve_queries(48, SIZES, NBUILD, seed=11, templates=BATCH_TEMPLATES)
    print("batch mix of", len(specs), flush=True)
    b = Batch(ctx, arena_bytes=1 << 30)
    got = b.run(_queries(cols, specs), max_groups=64)
    for s, g in zip(specs, got):
        assert_same(s, g, oracle_spec(s, H), what="mix")
    b.free()


def test_alone_equals_mix(env):
    ctx, _, cols, _ = env
    specs = serve_queries(32, SIZES, NBUILD, seed=12, templates=BATCH_TEMPLATES)
    print("alone vs mix", flush=True)
    b = Batch(ctx, arena_bytes=1 << 30)
    mix = b.run(_queries(cols, specs), max_groups=64)
    for s, m in zip(specs, mix):
        alone = b.run(_queries(cols, [s]), max_groups=64)[0]
        assert_same(s, alone, (m[1], m[2], m[3]), what="alone")
    b.free()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 7, 9])
def test_shared_scan_group_sizes(env, k):
    """k register-path scans over one table: groups of one to four queries
    (one to four tile slices per query), several groups per table, and the
    tail tile (250k rows is not a multiple of the tile)."""
    ctx, _, cols, H = env
    specs = serve_queries(64, SIZES, NBUILD, seed=31,
                          templates=["c1_small", "two_filter_ungrouped", "two_filter", "c1_scan_agg"])[:k]
    for s in specs:
        s["table"] = 2
    print("shared scan group of", k, flush=True)
    b = Batch(ctx, arena_bytes=64 << 20)
    got = b.run(_queries(cols, specs), max_groups=64)
    for s, g in zip(specs, got):
        assert_same(s, g, oracle_spec(s, H), what=f"group of {k}")
    b.free()


@pytest.mark.parametrize("case", ["dup_keys", "sparse_keys", "no_build_rows", "negative_keys"])
def test_small_build_direct_map(env, case):
    """Star joins over a small dimension use a direct map from key offset to
    (group, build row) when the qualifying keys are unique and dense; repeated
    keys or a span beyond the map fall back to the hash table, an empty build
    side matches nothing. Every case against the oracle, alone and in a mix."""
    from kushdb_amd.runtime import Column
    ctx, tables, cols, H = env
    facts, dim, build = cols
    pk = tables["dim"]["pk"]
    if case == "dup_keys":
        pk = pk // 2                      # every key twice
    elif case == "sparse_keys":
        pk = pk * 1000                    # span 10^7: beyond the map
    elif case == "negative_keys":
        pk = pk - 5000                    # dense, negative base
    pk = pk.contiguous()
    ddim = [Column(abi.KG_INT, pk), dim[1]]
    Hd = dict(H, dim={"pk": pk.cpu().numpy(), "cat": H["dim"]["cat"]})
    specs = serve_queries(40, SIZES, NBUILD, seed=41, templates=["c0_star", "star_dim_filter", "selective_join"])[:6]
    if case == "no_build_rows":
        for s_ in specs:
            s_["template"] = "star_dim_filter"
            s_["build_preds"] = [(1, abi.KG_LT, 0)]
            s_["aggs"] = [(abi.KG_AGG_SUM, 5), (abi.KG_AGG_COUNT, -1)]
    if case in ("sparse_keys", "negative_keys"):
        # fact keys that hit the transformed dimension keys
        fk = [Column(c.type, c.data) for c in facts[2]]
        fkv = tables["facts"][2]["fk"]
        fk[3] = Column(abi.KG_INT, (fkv * 1000 if case == "sparse_keys" else fkv - 5000).contiguous())
        fcols = [f if i != 2 else fk for i, f in enumerate(facts)]
        Hf = [dict(h) for h in Hd["facts"]]
        Hf[2]["fk"] = fk[3].data.cpu().numpy()
        Hd = dict(Hd, facts=Hf)
    else:
        fcols = facts
    for s_ in specs:
        s_["table"] = 2
    print("direct map case", case, flush=True)
    b = Batch(ctx, arena_bytes=64 << 20)
    qs = [make_query(s_, fcols, ddim, build) for s_ in specs]
    mix = b.run(qs, max_groups=64)
    for s_, g in zip(specs, mix):
        assert_same(s_, g, oracle_spec(s_, Hd), what=f"{case} mix")
    alone = b.run(qs[:1], max_groups=64)[0]
    assert_same(specs[0], alone, oracle_spec(specs[0], Hd), what=f"{case} alone")
    b.free()


def test_repeated_submits_and_empty(env):
    ctx, _, cols, H = env
    b = Batch(ctx, arena_bytes=64 << 20)
    assert b.run([], max_groups=8) == []
    specs = serve_queries(8, SIZES, NBUILD, seed=13, templates=BATCH_TEMPLATES)
    first = b.run(_queries(cols, specs), max_groups=64)
    for _ in range(3):
        again = b.run(_queries(cols, specs), max_groups=64)
        for s, x, y in zip(specs, first, again):
            assert_same(s, y, (x[1], x[2], x[3]), what="repeat")
    # all rows filtered out; an empty probe side
    facts, dim, _ = cols
    q_none = Query.scan_agg(facts[0], [(0, abi.KG_LT, 0)], [(1, 3)], [(abi.KG_AGG_COUNT, -1)], table_id=0)
    q_nojoin = Query.join_agg(facts[1], [(0, abi.KG_LT, 0)], 3, dim, 0, [(abi.KG_AGG_COUNT, -1)], table_id=1)
    r = b.run([q_none, q_nojoin], max_groups=8)
    assert r[0][0] == abi.KG_OK and len(r[0][2]) == 0
    assert r[1][0] == abi.KG_OK and len(r[1][2]) == 0
    b.free()


def test_unsupported_and_bad_codes_per_query(env):
    ctx, _, cols, H = env
    facts, dim, _ = cols
    b = Batch(ctx, arena_bytes=64 << 20)
    ok_spec = serve_queries(1, SIZES, NBUILD, seed=14, templates=["c1_small"])[0]
    q_ok = _queries(cols, [ok_spec])[0]
    q_min = Query.scan_agg(facts[0], [], [(1, 3)], [(abi.KG_AGG_MIN, 4)], table_id=0)
    # fk codes reach 9999: a group cardinality of 10 is violated
    q_bad = Query.scan_agg(facts[0], [], [(3, 10)], [(abi.KG_AGG_COUNT, -1)], table_id=0)
    # the same on the register path (at most 8 groups)
    q_bad8 = Query.scan_agg(facts[2], [(0, abi.KG_GEQ, 0)], [(3, 8)], [(abi.KG_AGG_SUM, 4)], table_id=2)
    q_chain = make_query({"kind": "chain", "p0": "a", "p1": "%e_a%", "vmax": 50}, *cols)
    r = b.run([q_min, q_ok, q_bad, q_bad8, q_chain], max_groups=64)
    assert r[0][0] == abi.KG_ERR_UNSUPPORTED
    assert r[2][0] == abi.KG_ERR_INVALID
    assert r[3][0] == abi.KG_ERR_INVALID
    assert r[4][0] == abi.KG_ERR_UNSUPPORTED
    assert_same(ok_spec, r[1], oracle_spec(ok_spec, H), what="ok next to failures")
    b.free()


def test_serve_plain_and_batched_agree(env):
    ctx, _, cols, H = env
    specs = serve_queries(64, SIZES, NBUILD, seed=15)
    rng = np.random.RandomState(5)
    arrivals = np.cumsum(rng.exponential(1.0 / 4000.0, size=len(specs)))
    print("serve plain", flush=True)
    plain, lat_p, st_p = serve(ctx, _queries(cols, specs), arrivals, abi.KG_SERVE_PLAIN, max_groups=64)
    print("serve batched", st_p, flush=True)
    batched, lat_b, st_b = serve(ctx, _queries(cols, specs), arrivals, abi.KG_SERVE_BATCHED, slots=2, max_batch=32,
                                 arena_bytes=128 << 20, max_groups=64)
    print("served", st_b, flush=True)
    for s, p, q in zip(specs, plain, batched):
        assert_same(s, p, oracle_spec(s, H), what="plain", H=H)
        assert_same(s, q, (p[1], p[2], p[3]), what="batched vs plain", H=H)
    assert (lat_p > 0).all() and (lat_b > 0).all()
    # chain queries are not batched: serving runs them on the plain path
    nchain = sum(s["kind"] == "chain" for s in specs)
    assert nchain >= 1
    assert st_b["batches"] >= 1 and st_b["fallbacks"] == nchain
    assert st_p["p99_s"] >= st_p["p50_s"] > 0


def test_unaligned_columns_take_the_generic_path(env):
    """Columns that start off a 16-byte boundary cannot use vector loads; the
    same query over row-shifted views must still match the oracle."""
    ctx, tables, cols, H = env
    from kushdb_amd.runtime import Column
    from kushdb_amd.workloads import FACT_COLS
    facts, dim, build = cols
    t = tables["facts"][2]
    shifted = [Column(c.type, c.data[1:]) for c in facts[2]]
    spec = serve_queries(1, SIZES, NBUILD, seed=16, templates=["c1_scan_agg"])[0]
    spec["table"] = 2
    q = Query.scan_agg(shifted, spec["preds"], spec["keys"], spec["aggs"], table_id=7)
    b = Batch(ctx, arena_bytes=64 << 20)
    got = b.run([q], max_groups=64)[0]
    Hs = {"facts": [{k: v[1:] for k, v in H["facts"][2].items()}] * 3, "dim": H["dim"], "build": H["build"]}
    assert_same(spec, got, oracle_spec(spec, Hs), what="unaligned")
    b.free()


def test_queries_complete_individually(env):
    """Each query's result is published by the device when its own last item is
    done; collecting them one by one while the batch runs gives the oracle's
    answers."""
    import time
    ctx, _, cols, H = env
    specs = serve_queries(24, SIZES, NBUILD, seed=17, templates=BATCH_TEMPLATES)
    b = Batch(ctx, arena_bytes=128 << 20)
    b.submit(_queries(cols, specs), max_groups=64)
    got = [None] * len(specs)
    deadline = time.time() + 60
    while any(g is None for g in got):
        assert time.time() < deadline, "queries not published"
        for i in range(len(specs)):
            if got[i] is None and b.ready(i):
                got[i] = b.collect_query(i)
    b.release()
    for s, g in zip(specs, got):
        assert_same(s, g, oracle_spec(s, H), what="streamed")
    b.free()


def test_alone_equals_mix_join_templates(env):
    """The join templates, including the big-build join (a 10-50% key range of
    the 16x larger build table, probed by k64b), each alone equal to its
    result in a mix with scans, and both against the oracle."""
    ctx, _, cols, H = env
    specs = serve_queries(60, SIZES, NBUILD, seed=19, templates=["c2_join_16m", "c2_join", "c1_scan_agg",
                                                                 "selective_join"])[:16]
    assert any(s["template"] == "c2_join_16m" for s in specs)
    print("alone vs mix (join templates)", flush=True)
    b = Batch(ctx, arena_bytes=1 << 30)
    mix = b.run(_queries(cols, specs), max_groups=64)
    for s, m in zip(specs, mix):
        assert_same(s, m, oracle_spec(s, H), what="join mix")
        alone = b.run(_queries(cols, [s]), max_groups=64)[0]
        assert_same(s, alone, (m[1], m[2], m[3]), what="join alone")
    b.free()


def test_map_only_build_sources(env):
    """Joins whose build side exceeds the row limit of a table-backed direct map
    (here a 3M-key build table) run map-only: one map per build source per
    batch, built over all build rows when several queries share the source,
    each query re-checking its own key range on the probe key. Every query
    against the oracle in a mix with table-backed joins and scans, and alone
    (the map then holds only its own range); a source whose keys repeat cannot
    use a map and is reported unsupported, and serving runs such queries on
    the plain path with the same results."""
    from kushdb_amd.runtime import Column
    ctx, _, _, _ = env
    nbig = 3_000_000
    tables = serve_tables(SIZES, nbuild=NBUILD, nbig=nbig, device="cuda:0", chain_scale=CHAIN_SCALE)
    torch.cuda.synchronize()
    cols = serve_columns(tables)
    H = host_tables(tables)
    specs = serve_queries(80, SIZES, NBUILD, seed=23, templates=["c2_join_16m", "c2_join", "c0_star", "c1_small"],
                          nbig=nbig)[:24]
    assert sum(s["template"] == "c2_join_16m" for s in specs) >= 3
    print("map-only sources: mix of", len(specs), flush=True)
    b = Batch(ctx, arena_bytes=1 << 30)
    mix = b.run(_queries(cols, specs), max_groups=64)
    for s, m in zip(specs, mix):
        assert_same(s, m, oracle_spec(s, H), what="map-only mix")
    big = [s for s in specs if s["template"] == "c2_join_16m"]
    for s in big[:2]:
        alone = b.run(_queries(cols, [s]), max_groups=64)[0]
        assert_same(s, alone, oracle_spec(s, H), what="map-only alone")
    # Repeated build keys: the map is off and there is no table to fall back to.
    facts, dim, builds = cols
    dup = dict(builds)
    dup["big"] = [Column(abi.KG_BIGINT, (tables["big"]["key"] // 2).contiguous()), builds["big"][1]]
    qd = [make_query(s, facts, dim, dup) for s in big[:2]]
    got = b.run(qd, max_groups=64)
    assert all(g[0] == abi.KG_ERR_UNSUPPORTED for g in got), [g[0] for g in got]
    b.free()
    Hd = dict(H, big={"key": H["big"]["key"] // 2, "payload": H["big"]["payload"]})
    arrivals = np.zeros(len(qd))
    res, _, st = serve(ctx, qd, arrivals, abi.KG_SERVE_BATCHED, slots=1, max_batch=8, arena_bytes=1 << 30,
                       max_groups=64)
    assert st["fallbacks"] == len(qd)
    for s, g in zip(big[:2], res):
        assert_same(s, g, oracle_spec(s, Hd), what="repeated keys served")


def test_large_batch_counter_dealing(env):
    """A batch whose join queue holds more items than an XCD has waves (about
    18 joins over a 16M-row table: some 2k items per queue), so waves take the
    items past their first from the queue counters. Every query equals its
    result alone, where each item is some wave's first."""
    ctx = env[0]
    n = 16_000_000
    tables = serve_tables((n,), nbuild=NBUILD, device="cuda:0", chain_scale=CHAIN_SCALE)
    torch.cuda.synchronize()
    cols = serve_columns(tables)
    specs = serve_queries(64, (n,), NBUILD, seed=41,
                          templates=["c2_join", "c0_star", "selective_join", "star_dim_filter", "two_filter"])[:24]
    assert sum(s["kind"] == "join" for s in specs) >= 16
    print("large batch of", len(specs), "over", n, "rows", flush=True)
    b = Batch(ctx, arena_bytes=1 << 30)
    mix = b.run(_queries(cols, specs), max_groups=64)
    for s, m in zip(specs, mix):
        alone = b.run(_queries(cols, [s]), max_groups=64)[0]
        assert_same(s, alone, (m[1], m[2], m[3]), what="counter dealing")
    b.free()
    del tables, cols
    torch.cuda.empty_cache()


def test_shared_source_key_intervals(env):
    """Joins over one build source whose build predicates compare the build key
    with literals: the source's map is built over the union of their key
    intervals (merged; each query's own predicates instead when the union has
    more than eight pieces or a query uses <>), and every query re-checks its
    own range at a match. Each query against the oracle: overlapping and
    adjacent ranges, a point, open ends, an empty range, a query without build
    predicates, a <> query and twelve separate points."""
    ctx, _, cols, H = env
    base = serve_queries(64, SIZES, NBUILD, seed=31, templates=["c2_join"])
    EQ, NE, LT, LE, GT, GE = abi.KG_EQ, abi.KG_NEQ, abi.KG_LT, abi.KG_LEQ, abi.KG_GT, abi.KG_GEQ
    cases = [
        [[(0, GE, 100), (0, LT, 5000)], [(0, GE, 5000), (0, LE, 9000)], [(0, EQ, 20000)],
         [(0, GT, NBUILD - 50)], [(0, LT, 10)], [(0, GT, 500), (0, LT, 400)]],
        [[(0, GE, 100), (0, LT, 5000)], []],
        [[(0, NE, 77)], [(0, LT, 3000)]],
        [[(0, EQ, 1000 * i)] for i in range(12)],
    ]
    b = Batch(ctx, arena_bytes=1 << 30)
    for preds in cases:
        specs = []
        for j, bp in enumerate(preds):
            s = dict(base[j])
            s["build_preds"] = bp
            specs.append(s)
        print("shared source, build predicates", preds, flush=True)
        got = b.run(_queries(cols, specs), max_groups=64)
        for s, g in zip(specs, got):
            assert_same(s, g, oracle_spec(s, H), what="key intervals")
    b.free()


CHAIN_CASES = [("a", "%e_a%", 50), ("t", "%o_i%", 10), ("q", "%a_e%", 100), ("s", "%i_e%", 37),
               ("zz", "%e_a%", 50), ("a", "%e_a%", 0)]


@pytest.mark.parametrize("budget,stored", [(0, False), (0, True), (4096, False)])
def test_chain_query_vs_oracle(env, budget, stored):
    """The chain kind through kg_chain_run: five tables, t0.name STARTS_WITH,
    t1.s LIKE, t4.v < x, joined in the fixed order (budget 0; indexes built
    per query, or the stored fk indexes of the unfiltered t2 and t3) or by
    the Skinner driver over several episodes; tuple count, MIN(t4.v) and the
    MIN(t0.name) string equal to the oracle's chain restatement, including
    empty results (no prefix match, no v below 0)."""
    from kushdb_amd.runtime import chain_run
    from kushdb_amd.workloads import chain_query_for
    from oracle import pyoracle as po
    from serve_oracle import chain_name
    ctx, _, cols, H = env
    tabs = cols[2]["chain"]
    for p0, p1, vmax in CHAIN_CASES:
        print("chain", p0, p1, vmax, "budget", budget, flush=True)
        q = chain_query_for(tabs, p0, p1, vmax, skinner_budget=budget, seed=3,
                            indexes=cols[2]["chain_idx"] if stored else None)
        n, vmin, row = chain_run(ctx, q)
        wn, wname, wvmin = po.chain_query(H["chain"], p0.encode(), p1.encode(), vmax)
        assert n == wn, (p0, p1, vmax, n, wn)
        assert vmin == wvmin, (p0, p1, vmax, vmin, wvmin)
        vals = np.array([[n, 0, -1 if row is None else row]], dtype=np.int64)
        nulls = np.array([[0, 0, row is None]], dtype=np.int8)
        assert chain_name(H, vals, nulls) == wname, (p0, p1, vmax)
    assert any(po.chain_query(H["chain"], p0.encode(), p1.encode(), v)[0] > 0 for p0, p1, v in CHAIN_CASES)
