"""Device expressions (kg_expr_eval / kg_expr_select) against the reference's
fixtures and the oracle: the select/* filters over fixture columns loaded with
kg_file_* (sorted rows equal to the expected .tbl), TPC-H q06 (filter, then
SUM of a product, within the reference's 1e-5) and q01 with its derived
measures computed on the device (then GROUP BY and ORDER BY on the device),
and random typed expression trees with NULLs in every column, value- and
NULL-exact against kor_expr_eval (REAL bit-exact: every node is one IEEE
operation), including row subsets and the rejected forms."""
import ctypes
import os

import numpy as np
import pytest
import torch

import kdb
import queries
from kushdb_amd import abi
from kushdb_amd.expr import (Add, And, Case, ColRef, Contains, Div, EndsWith, Eq, ExtractYear, Float, Geq, Gt, In,
                             IsNull, Leq, Like, Literal, Lt, Mul, Neq, Not, Or, Program, StartsWith, Sub, expr_eval,
                             expr_select)
from kushdb_amd.runtime import Column, Context, KgError, TextColumn, order_by
from oracle import pyoracle as po
from test_oracle_expr import SCAN, SELECT, T, oracle_col

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.file_release()


def dev_col(ctx, name, typ):
    path = os.path.join(kdb.DATA, name + ".kdb")
    if typ == "text":
        return ctx.file_text(path), None
    d = ctx.file_enum(os.path.join(kdb.DATA, name + ".kdbenum")) if typ == "enum" else None
    return ctx.file_column(path, T[typ]), d


@pytest.mark.parametrize("name", sorted(SELECT))
def test_select_fixture(ctx, name):
    print("select fixture", name, flush=True)
    f, typ, flt = SELECT[name]
    col, d = dev_col(ctx, f, typ)
    _, printed = oracle_col(f, typ)
    rows = expr_select(ctx, len(col), [col], flt(ColRef(0, T[typ])), dictionaries={0: d} if d else None)
    got = sorted(printed[r] + "|" for r in rows.cpu().numpy())
    assert got == sorted(kdb.expected(f"select/{name}_expected.tbl"))


@pytest.mark.parametrize("name", sorted(SCAN))
def test_scan_fixture(ctx, name):
    """ScanSelect / SimdScanSelect fixtures: the general expression filter
    (kg_expr_select) and, where the predicate is column-op-literal, the scan
    fast path (kg_scan_select; ENUM literals mapped to their code through the
    dictionary as the SIMD translator does) or kg_text_select."""
    from kushdb_amd.runtime import text_select
    print("scan fixture", name, flush=True)
    f, typ, flt, fast = SCAN[name]
    col, d = dev_col(ctx, f, typ)
    _, printed = oracle_col(f, typ)
    want = sorted(kdb.expected(f"{name}_expected.tbl"))
    rows = expr_select(ctx, len(col), [col], flt(ColRef(0, T[typ])), dictionaries={0: d} if d else None)
    assert sorted(printed[r] + "|" for r in rows.cpu().numpy()) == want
    if fast is None:
        return
    if fast[0] == "text":
        r2 = text_select(ctx, col, abi.KG_TEXT_GT, fast[2])
    else:
        op, lit = fast
        if isinstance(lit, tuple) and lit[0] == "date":
            lit = po.build_date(*lit[1:])
        elif isinstance(lit, tuple) and lit[0] == "code":
            lit = kdb.enum_dict(f).index(lit[1])
        r2 = ctx.scan_select([col], [(0, op, lit)])
    assert np.array_equal(np.sort(r2.cpu().numpy()), np.sort(rows.cpu().numpy()))


def test_tpch_q06(ctx):
    from test_oracle_expr import q06_filter
    print("tpch q06", flush=True)
    cols = [dev_col(ctx, "l_shipdate", "date")[0], dev_col(ctx, "l_discount", "real")[0],
            dev_col(ctx, "l_quantity", "real")[0], dev_col(ctx, "l_extendedprice", "real")[0]]
    n = len(cols[0])
    rows = expr_select(ctx, n, cols, q06_filter())
    rev = expr_eval(ctx, n, cols, Mul(ColRef(3, abi.KG_REAL), ColRef(1, abi.KG_REAL)), rows=rows)
    _, vals, _ = ctx.scan_group_agg([rev], [], [], [(abi.KG_AGG_SUM, 0)])
    total = kdb.bits_f64(vals[0, 0])
    exp = float(kdb.expected("tpch/q06_expected.tbl")[0])
    assert abs(total - exp) < 1e-5
    # the same rows and products as the oracle
    from test_oracle_expr import q06_columns, selected
    ocols = q06_columns()
    want = selected(q06_filter(), ocols, n)
    assert np.array_equal(rows.cpu().numpy(), want)
    assert np.array_equal(rev.data.cpu().numpy(), ocols[3].data[want] * ocols[1].data[want])


def test_tpch_q01_device_expressions(ctx):
    print("tpch q01", flush=True)
    names = [("l_returnflag_enum", "enum"), ("l_linestatus_enum", "enum"), ("l_quantity", "real"),
             ("l_extendedprice", "real"), ("l_discount", "real"), ("l_tax", "real"), ("l_shipdate", "date")]
    loaded = [dev_col(ctx, n_, t) for n_, t in names]
    cols = [c for c, _ in loaded]
    rfd, lsd = loaded[0][1], loaded[1][1]
    n = len(cols[0])
    rows = expr_select(ctx, n, cols, Leq(ColRef(6, abi.KG_DATE), Literal((1998, 9, 13))))
    E, R = abi.KG_ENUM, abi.KG_REAL
    p, d, t = ColRef(3, R), ColRef(4, R), ColRef(5, R)
    derived = [ColRef(0, E), ColRef(1, E), ColRef(2, R), p, Mul(p, Sub(Literal(1.0), d)),
               Mul(Mul(p, Sub(Literal(1.0), d)), Add(Literal(1.0), t)), d]
    dc = [expr_eval(ctx, n, cols, e, rows=rows) for e in derived]
    dc = [Column(c.type, c.data) for c in dc]          # no NULLs in lineitem
    S, A, C = abi.KG_AGG_SUM, abi.KG_AGG_AVG, abi.KG_AGG_COUNT
    aggs = [(S, 2), (S, 3), (S, 4), (S, 5), (A, 2), (A, 3), (A, 6), (C, -1)]
    codes, vals, _ = ctx.scan_group_agg(dc, [], [(0, len(rfd)), (1, len(lsd))], aggs)
    # ORDER BY l_returnflag, l_linestatus on the device (decoded strings)
    kf = Column(E, torch.as_tensor(codes[:, 0].copy()).to("cuda:0"))
    ks = Column(E, torch.as_tensor(codes[:, 1].copy()).to("cuda:0"))
    perm = order_by(ctx, [kf, ks], [True, True], {0: rfd, 1: lsd}).cpu().numpy()
    got = []
    for g in perm:
        nums = [kdb.bits_f64(v) for v in vals[g, :7]] + [int(vals[g, 7])]
        got.append((rfd[codes[g, 0]], lsd[codes[g, 1]], nums))
    queries.check_q01(got, kdb.expected("tpch/q01_expected.tbl"))


# ---- random expression trees vs the oracle ----------------------------------------
TYPES = [abi.KG_SMALLINT, abi.KG_INT, abi.KG_BIGINT, abi.KG_REAL, abi.KG_DATE, abi.KG_BOOLEAN, abi.KG_ENUM,
         abi.KG_TEXT]
DICT = ["pear", "apple", "fig", "apple pie", "Fig", "", "zebra"]
WORDS = ["", "a", "ab", "abc", "b", "banana", "bandana", "xyz", "abab", "Zed", "a_c"]


def make_columns(rng, n):
    host, dev = [], []
    for typ in TYPES:
        nulls = (rng.random(n) < 0.15).astype(np.int8)
        if typ == abi.KG_TEXT:
            strs = [WORDS[k] for k in rng.integers(0, len(WORDS), n)]
            host.append(("text", strs, nulls))
            dev.append(TextColumn.from_strings(strs, nulls=nulls))
            continue
        if typ == abi.KG_REAL:
            v = np.round(rng.standard_normal(n) * 100, 2)
            v[rng.random(n) < 0.05] = 0.0
        elif typ == abi.KG_SMALLINT:
            v = rng.integers(-300, 300, n)
        elif typ == abi.KG_INT:
            v = rng.integers(-2**31, 2**31, n, dtype=np.int64) if rng.random() < 0.5 else rng.integers(-50, 50, n)
        elif typ == abi.KG_BIGINT:
            v = rng.integers(-2**62, 2**62, n, dtype=np.int64)
        elif typ == abi.KG_DATE:
            v = rng.integers(po.build_date(1990, 1, 1), po.build_date(2030, 1, 1), n)
        elif typ == abi.KG_BOOLEAN:
            v = rng.integers(0, 2, n)
        else:
            v = rng.integers(0, len(DICT), n)
        c = po.Col(typ, v, nulls, dictionary=DICT if typ == abi.KG_ENUM else None)
        host.append(c)
        dev.append(Column(typ, torch.as_tensor(c.data).to("cuda:0"), torch.as_tensor(nulls).to("cuda:0")))
    return host, dev


def col(typ):
    return ColRef(TYPES.index(typ), typ)


def gen(rng, typ, depth):
    r = rng.random()
    leaf = depth <= 0 or r < 0.25
    if typ == abi.KG_BOOLEAN:
        if leaf:
            return col(typ) if rng.random() < 0.5 else Literal(bool(rng.integers(2)), null=rng.random() < 0.1)
        k = int(rng.integers(10))
        if k == 0:
            return And([gen(rng, typ, depth - 1), gen(rng, typ, depth - 1)])
        if k == 1:
            return Or([gen(rng, typ, depth - 1), gen(rng, typ, depth - 1)])
        if k == 2:
            return Not(gen(rng, typ, depth - 1))
        if k == 3:
            return IsNull(gen(rng, TYPES[int(rng.integers(6))], depth - 1))
        if k == 4:
            t = [abi.KG_SMALLINT, abi.KG_INT, abi.KG_BIGINT, abi.KG_REAL, abi.KG_DATE][int(rng.integers(5))]
            mk = [Eq, Neq, Lt, Leq, Gt, Geq][int(rng.integers(6))]
            return mk(gen(rng, t, depth - 1), gen(rng, t, depth - 1))
        if k == 5:
            pat = ["a%", "%an%", "_", "%", "b%a", "a_c", "Z%"][int(rng.integers(7))]
            return Like(col([abi.KG_TEXT, abi.KG_ENUM][int(rng.integers(2))]), pat)
        if k == 6:
            mk = [StartsWith, EndsWith, Contains, Lt, Geq][int(rng.integers(5))]
            return mk(col([abi.KG_TEXT, abi.KG_ENUM][int(rng.integers(2))]), Literal(WORDS[int(rng.integers(len(WORDS)))]))
        if k == 7:
            return In(col(abi.KG_ENUM), list(rng.choice(len(DICT), int(rng.integers(0, 4)), replace=False)))
        if k == 8:
            mk = [Eq, Neq, Lt, Gt][int(rng.integers(4))]
            return mk(col(abi.KG_ENUM), Literal(int(rng.integers(len(DICT))), abi.KG_ENUM) if rng.random() < 0.5
                      else col(abi.KG_ENUM))
        return Case(gen(rng, typ, depth - 1), gen(rng, typ, depth - 1), gen(rng, typ, depth - 1))
    if typ in (abi.KG_SMALLINT, abi.KG_INT, abi.KG_BIGINT, abi.KG_REAL):
        if leaf:
            if rng.random() < 0.6:
                return col(typ)
            v = float(np.round(rng.standard_normal() * 10, 2)) if typ == abi.KG_REAL else int(rng.integers(-9, 9))
            return Literal(v, typ, null=rng.random() < 0.05)
        k = int(rng.integers(6))
        if k <= 2:
            return [Add, Sub, Mul][k](gen(rng, typ, depth - 1), gen(rng, typ, depth - 1))
        if k == 3:
            return Case(gen(rng, abi.KG_BOOLEAN, depth - 1), gen(rng, typ, depth - 1), gen(rng, typ, depth - 1))
        if typ == abi.KG_REAL:
            if k == 4:
                return Div(gen(rng, typ, depth - 1), gen(rng, typ, depth - 1))
            return Float(gen(rng, [abi.KG_SMALLINT, abi.KG_INT, abi.KG_BIGINT][int(rng.integers(3))], depth - 1))
        if typ == abi.KG_INT and k == 4:
            return ExtractYear(gen(rng, abi.KG_DATE, depth - 1))
        return gen(rng, typ, 0)
    if typ == abi.KG_DATE:
        if leaf or rng.random() < 0.5:
            return col(typ) if rng.random() < 0.7 else Literal((2000 + int(rng.integers(30)), 1, 1))
        return Case(gen(rng, abi.KG_BOOLEAN, depth - 1), gen(rng, typ, depth - 1), gen(rng, typ, depth - 1))
    return col(typ)


def as_bits(c):
    """Device output column -> (int64 value bits, int8 nulls) like the oracle's."""
    v = c.data.cpu().numpy()
    if v.dtype == np.float64:
        v = v.view(np.int64)
    return v.astype(np.int64), c.nulls.cpu().numpy()


def same(typ, got, want):
    gv, gn = got
    wv, wn = want
    assert np.array_equal(gn, wn)
    m = wn == 0
    if typ == abi.KG_REAL:
        a, b = gv[m].view(np.float64), wv[m].view(np.float64)
        assert np.array_equal(a, b, equal_nan=True)
        assert np.array_equal(np.signbit(a) | np.isnan(a), np.signbit(b) | np.isnan(b))
    else:
        assert np.array_equal(gv[m], wv[m])


@pytest.mark.parametrize("seed", range(8))
def test_random_expressions_vs_oracle(ctx, seed):
    rng = np.random.default_rng(1000 + seed)
    n = [1, 70, 1000, 4097, 20000, 65536, 100003, 250000][seed]
    print("random expressions", seed, n, flush=True)
    host, dev = make_columns(rng, n)
    dicts = {TYPES.index(abi.KG_ENUM): DICT}
    sub = np.sort(rng.choice(n, max(n // 3, 1), replace=False)).astype(np.int32)
    dsub = torch.as_tensor(sub).to("cuda:0")
    for trial in range(12):
        typ = [abi.KG_BOOLEAN, abi.KG_INT, abi.KG_REAL, abi.KG_BIGINT, abi.KG_SMALLINT, abi.KG_DATE][trial % 6]
        e = gen(rng, typ, 4)
        p = Program(e)
        want = po.expr_eval(ctypes.addressof(p.nodes), p.n, host, n)
        assert want is not None, (seed, trial)
        same(e.type, as_bits(expr_eval(ctx, n, dev, e, dictionaries=dicts)), want)
        want_s = po.expr_eval(ctypes.addressof(p.nodes), p.n, host, n, sub)
        same(e.type, as_bits(expr_eval(ctx, n, dev, e, rows=dsub, dictionaries=dicts)), want_s)
        if e.type == abi.KG_BOOLEAN:
            rows = expr_select(ctx, n, dev, e, dictionaries=dicts).cpu().numpy()
            assert np.array_equal(rows, np.nonzero((want[0] != 0) & (want[1] == 0))[0])
            rows = expr_select(ctx, n, dev, e, rows=dsub, dictionaries=dicts).cpu().numpy()
            assert np.array_equal(rows, sub[(want_s[0] != 0) & (want_s[1] == 0)])


def test_rejected_expressions(ctx):
    n = 10
    i = Column(abi.KG_INT, torch.arange(n, dtype=torch.int32, device="cuda:0"))
    r = Column(abi.KG_REAL, torch.ones(n, dtype=torch.float64, device="cuda:0"))
    I, Rr = ColRef(0, abi.KG_INT), ColRef(1, abi.KG_REAL)
    with pytest.raises(KgError):
        expr_eval(ctx, n, [i, r], Div(I, I))            # DIV takes REAL
    with pytest.raises(KgError):
        expr_eval(ctx, n, [i, r], Add(I, Rr))           # operand types differ
    with pytest.raises(KgError):
        expr_select(ctx, n, [i, r], Add(I, I))          # a filter must be BOOLEAN
    with pytest.raises(KgError):
        expr_eval(ctx, n, [i, r], Lt(ColRef(0, abi.KG_BIGINT), I))   # column type mismatch
    assert expr_select(ctx, 0, [i, r], Gt(I, Literal(3))).numel() == 0


def test_group_by_enum_fixture(ctx):
    """group_by_aggregate/enum: GROUP BY an ENUM column with NULLs, MIN/MAX by
    string order, COUNT(*), COUNT(name), all on the device: the key is the code
    with NULL as one more code (CASE(IS NULL) on the device), MIN/MAX run over
    each code's string rank (kg_enum_ranks, gathered on the device) and map back
    through the sorted dictionary."""
    from kushdb_amd.runtime import enum_ranks, gather
    from test_oracle_expr import groupby_enum_rows
    col, d = dev_col(ctx, "people_name_enum", "enum")
    strings = kdb.enum_dict("people_name_enum")
    card = len(strings)
    n = len(col)
    key = expr_eval(ctx, n, [col], Case(IsNull(ColRef(0, abi.KG_ENUM)), Literal(card, abi.KG_ENUM),
                                        ColRef(0, abi.KG_ENUM)))
    code0 = expr_eval(ctx, n, [col], Case(IsNull(ColRef(0, abi.KG_ENUM)), Literal(0, abi.KG_ENUM),
                                          ColRef(0, abi.KG_ENUM)))
    ranks = np.asarray(enum_ranks(strings), dtype=np.int32)
    rcol = gather(ctx, Column(abi.KG_INT, torch.from_numpy(ranks).to("cuda:0")), code0.data.contiguous())
    rank_col = Column(abi.KG_INT, rcol.data, col.nulls)
    kc = Column(abi.KG_ENUM, key.data, None)
    # 943 groups x 4 aggregates exceed one workgroup's LDS table budget: two
    # passes (the group order is the first appearance in both).
    codes, v1, n1 = ctx.scan_group_agg([kc, rank_col], [], [(0, card + 1)],
                                       [(abi.KG_AGG_MIN, 1), (abi.KG_AGG_MAX, 1)], max_groups=card + 1)
    codes2, v2, n2 = ctx.scan_group_agg([kc, col], [], [(0, card + 1)],
                                        [(abi.KG_AGG_COUNT, -1), (abi.KG_AGG_COUNT, 1)], max_groups=card + 1)
    assert np.array_equal(codes, codes2)
    vals, nulls = np.concatenate([v1, v2], axis=1), np.concatenate([n1, n2], axis=1)
    # One call with all four: the dense table leaves LDS and runs as the hashed GROUP BY.
    c4, v4, n4 = ctx.scan_group_agg([kc, rank_col, col], [], [(0, card + 1)],
                                    [(abi.KG_AGG_MIN, 1), (abi.KG_AGG_MAX, 1), (abi.KG_AGG_COUNT, -1),
                                     (abi.KG_AGG_COUNT, 2)], max_groups=card + 1)
    assert np.array_equal(c4, codes) and np.array_equal(v4, vals) and np.array_equal(n4, nulls)
    by_rank = [None] * card
    for c, r in enumerate(ranks):
        by_rank[r] = strings[c]
    got = []
    for g in range(len(codes)):
        mn = "" if nulls[g, 0] else by_rank[int(vals[g, 0])]
        mx = "" if nulls[g, 1] else by_rank[int(vals[g, 1])]
        got.append(f"{mn}|{mx}|{int(vals[g, 2])}|{int(vals[g, 3])}|")
    assert sorted(got) == sorted(kdb.expected("group_by_aggregate/enum_expected.tbl"))
    assert got == groupby_enum_rows()   # the same first-appearance group order


def test_group_by_text_fixture(ctx):
    """group_by_aggregate/text: GROUP BY CASE(id > 500, 'hello', '') (the two
    branches as codes 1 / 0 from a device CASE; a NULL id takes the ELSE
    branch), MIN/MAX(name) per group with kg_text_minmax over the group's rows,
    COUNT(*) and COUNT(name) from the device group-by (COUNT(name) over a column
    carrying name's NULL flags)."""
    from kushdb_amd.runtime import text_minmax
    from test_oracle_expr import groupby_text_rows
    ids, _ = dev_col(ctx, "people_id", "int")
    names, _ = dev_col(ctx, "people_name", "text")
    n = len(ids)
    code = expr_eval(ctx, n, [ids], Case(Gt(ColRef(0, abi.KG_INT), Literal(500)), Literal(1), Literal(0)))
    assert code.nulls is None or int(code.nulls.sum()) == 0
    kc = Column(abi.KG_INT, code.data, None)
    name_flags = Column(abi.KG_INT, ids.data, names.nulls)
    codes, vals, _ = ctx.scan_group_agg([kc, name_flags], [], [(0, 2)], [(abi.KG_AGG_COUNT, -1),
                                                                          (abi.KG_AGG_COUNT, 1)])
    got = []
    for g in range(len(codes)):
        rows = ctx.scan_select([kc], [(0, abi.KG_EQ, int(codes[g, 0]))])
        _, mn = text_minmax(ctx, names, rows)
        _, mx = text_minmax(ctx, names, rows, is_max=True)
        got.append(f"{(mn or b'').decode()}|{(mx or b'').decode()}|{int(vals[g, 0])}|{int(vals[g, 1])}|")
    assert sorted(got) == sorted(kdb.expected("group_by_aggregate/text_expected.tbl"))
    assert got == groupby_text_rows()
