"""Expression oracle (kor_expr_eval) pinned on the reference's fixtures and on
closed forms: the select/* tests (end_to_end_test/select/*_test.cc: one
filter per type, sorted output rows), tpch/q06 (a five-way AND filter, then
SUM(l_extendedprice * l_discount)), tpch/q01's derived measures (through the
GROUP BY of the same fixture); numpy's wrapping integer arithmetic and IEEE
REAL arithmetic and comparisons on random values; every combination of
TRUE/FALSE/NULL through AND, OR, NOT and CASE against the rules the
translator emits; LIKE against Python's re.fullmatch of the planner's regex;
EXTRACT(YEAR) against datetime. The expression nodes are the product's own
ctypes layout (kushdb_amd.expr.Program), which the oracle reads as its
kor_expr_node."""
import ctypes
import datetime
import re

import numpy as np
import pytest

import kdb
from kushdb_amd import abi
from kushdb_amd.expr import (Add, And, Case, ColRef, Contains, Div, EndsWith, Eq, ExtractYear, Float, Geq, Gt, In,
                             IsNull, Leq, Like, Literal, Lt, Mul, Neq, Not, Or, Program, StartsWith, Sub)
from oracle import pyoracle as po

T = {"int": abi.KG_INT, "bigint": abi.KG_BIGINT, "smallint": abi.KG_SMALLINT, "date": abi.KG_DATE,
     "real": abi.KG_REAL, "enum": abi.KG_ENUM, "boolean": abi.KG_BOOLEAN, "text": abi.KG_TEXT}

# select fixture: (column file, type, filter on ColRef(0))
SELECT = {
    "int": ("info_id", "int", lambda c: Gt(c, Literal(29))),
    "bigint": ("info_num2", "bigint", lambda c: Gt(c, Literal(0, abi.KG_BIGINT))),
    "smallint": ("info_num1", "smallint", lambda c: Lt(c, Literal(0, abi.KG_SMALLINT))),
    "date": ("info_date", "date", lambda c: Geq(c, Literal((2021, 1, 1)))),
    "real": ("info_zscore", "real", lambda c: Gt(c, Literal(-0.5))),
    "boolean": ("info_cheated", "boolean", lambda c: c),
    "enum": ("people_name_enum", "enum", lambda c: Gt(c, Literal("M"))),
    "text": ("people_name", "text", lambda c: Gt(c, Literal("M"))),
}

# scan_select / simd_scan_select fixtures with one printed column: (column file, type, filter on
# ColRef(0), fast-path predicate (op, literal) for kg_scan_select / kg_text_select or None)
# (end_to_end_test/scan_select/{boolean_eq,boolean,enum,int_eq,text}_test.cc,
#  end_to_end_test/simd_scan_select/{date,enum}_test.cc)
SCAN = {
    "scan_select/boolean_eq": ("info_cheated", "boolean", lambda c: Eq(c, Literal(True)), (po.EQ, 1)),
    "scan_select/boolean": ("info_cheated", "boolean", lambda c: c, (po.EQ, 1)),
    "scan_select/enum": ("people_name_enum", "enum", lambda c: Gt(c, Literal("M")), None),
    "scan_select/text": ("people_name", "text", lambda c: Gt(c, Literal("M")), ("text", po.GT, "M")),
    "scan_select/int_eq": ("c_custkey", "int", lambda c: Eq(c, Literal(555)), (po.EQ, 555)),
    "simd_scan_select/date": ("l_shipdate", "date", lambda c: Leq(c, Literal((1998, 9, 13))),
                              (po.LEQ, ("date", 1998, 9, 13))),
    # ENUM equality compares codes: the literal is the dictionary code of "R" (an enum literal)
    "simd_scan_select/enum": ("l_returnflag_enum", "enum",
                              lambda c: Eq(c, Literal(kdb.enum_dict("l_returnflag_enum").index("R"), abi.KG_ENUM)),
                              (po.EQ, ("code", "R"))),
}


def oracle_col(name, typ):
    if typ == "text":
        strs, nulls = kdb.text(name)
        return ("text", strs, nulls), ["" if nulls is not None and nulls[i] else s for i, s in enumerate(strs)]
    vals, nulls = kdb.col(name, typ)
    d = kdb.enum_dict(name) if typ == "enum" else None
    printed = ["" if nulls is not None and nulls[i] else (d[v] if d else kdb.fmt(typ, v)) for i, v in enumerate(vals)]
    return po.Col(T[typ], vals, nulls, dictionary=d), printed


def oracle_eval(expr, cols, n, rows=None):
    p = Program(expr)
    return po.expr_eval(ctypes.addressof(p.nodes), p.n, cols, n, rows)


def selected(expr, cols, n):
    res = oracle_eval(expr, cols, n)
    assert res is not None
    v, nl = res
    return np.nonzero((v != 0) & (nl == 0))[0]


@pytest.mark.parametrize("name", sorted(SELECT))
def test_oracle_select_fixture(name):
    f, typ, flt = SELECT[name]
    col, printed = oracle_col(f, typ)
    n = len(printed)
    rows = selected(flt(ColRef(0, T[typ])), [col], n)
    got = sorted(printed[r] + "|" for r in rows)
    assert got == sorted(kdb.expected(f"select/{name}_expected.tbl"))


def q06_filter():
    ship, disc, qty = ColRef(0, abi.KG_DATE), ColRef(1, abi.KG_REAL), ColRef(2, abi.KG_REAL)
    return And([Geq(ship, Literal((1993, 1, 1))), Lt(ship, Literal((1994, 1, 1))), Geq(disc, Literal(0.02)),
                Leq(disc, Literal(0.04)), Lt(qty, Literal(25.0))])


def q06_columns():
    return [po.Col(po.DATE, *kdb.col("l_shipdate", "date")), po.Col(po.REAL, *kdb.col("l_discount", "real")),
            po.Col(po.REAL, *kdb.col("l_quantity", "real")), po.Col(po.REAL, *kdb.col("l_extendedprice", "real"))]


def test_oracle_tpch_q06():
    cols = q06_columns()
    n = len(cols[0].data)
    rows = selected(q06_filter(), cols, n)
    rev = Mul(ColRef(3, abi.KG_REAL), ColRef(1, abi.KG_REAL))
    v, nl = oracle_eval(rev, cols, n, rows.astype(np.int32))
    assert not nl.any()
    total = 0.0
    for x in v.view(np.float64):
        total += x                      # the reference's sequential SUM
    exp = float(kdb.expected("tpch/q06_expected.tbl")[0])
    assert abs(total - exp) < 1e-5      # util/test_util.h REAL tolerance
    assert np.array_equal(v.view(np.float64), cols[3].data[rows] * cols[1].data[rows])


def test_oracle_tpch_q01_derived():
    price = po.Col(po.REAL, *kdb.col("l_extendedprice", "real"))
    disc = po.Col(po.REAL, *kdb.col("l_discount", "real"))
    tax = po.Col(po.REAL, *kdb.col("l_tax", "real"))
    ship = po.Col(po.DATE, *kdb.col("l_shipdate", "date"))
    rf = kdb.col("l_returnflag_enum", "enum")[0]
    ls = kdb.col("l_linestatus_enum", "enum")[0]
    n = len(price.data)
    cols = [price, disc, tax, ship]
    rows = selected(Leq(ColRef(3, abi.KG_DATE), Literal((1998, 9, 13))), cols, n).astype(np.int32)
    p, d, t = ColRef(0, abi.KG_REAL), ColRef(1, abi.KG_REAL), ColRef(2, abi.KG_REAL)
    disc_price = Mul(p, Sub(Literal(1.0), d))
    charge = Mul(Mul(p, Sub(Literal(1.0), d)), Add(Literal(1.0), t))
    dp = oracle_eval(disc_price, cols, n, rows)[0].view(np.float64)
    ch = oracle_eval(charge, cols, n, rows)[0].view(np.float64)
    rfn, lsn = kdb.enum_dict("l_returnflag_enum"), kdb.enum_dict("l_linestatus_enum")
    exp = {tuple(ln.split("|")[:2]): ln.split("|") for ln in kdb.expected("tpch/q01_expected.tbl")}
    code = rf[rows] * len(lsn) + ls[rows]
    for g in np.unique(code):
        m = code == g
        key = (rfn[g // len(lsn)], lsn[g % len(lsn)])
        s_dp = s_ch = 0.0
        for a, b in zip(dp[m], ch[m]):
            s_dp += a
            s_ch += b
        assert abs(s_dp - float(exp[key][4])) < 1e-5 and abs(s_ch - float(exp[key][5])) < 1e-5, key


def _rand_col(rng, typ, n, null_frac=0.2):
    nulls = (rng.random(n) < null_frac).astype(np.int8)
    if typ == po.REAL:
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 6, n)
        v[rng.random(n) < 0.05] = 0.0
    elif typ == po.SMALLINT:
        v = rng.integers(-32768, 32768, n)
    elif typ == po.INT:
        v = rng.integers(-2**31, 2**31, n, dtype=np.int64)
    elif typ == po.BIGINT:
        v = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    else:
        v = rng.integers(0, 2, n)
    return po.Col(typ, v, nulls)


@pytest.mark.parametrize("typ", [po.SMALLINT, po.INT, po.BIGINT, po.REAL])
def test_oracle_arithmetic_vs_numpy(typ):
    rng = np.random.default_rng(typ)
    n = 2000
    a, b = _rand_col(rng, typ, n), _rand_col(rng, typ, n)
    A, B = ColRef(0, typ), ColRef(1, typ)
    nulls = (a.nulls | b.nulls).astype(bool)
    npt = po._NP[typ]
    for mk, fn in ((Add, np.add), (Sub, np.subtract), (Mul, np.multiply)):
        v, nl = oracle_eval(mk(A, B), [a, b], n)
        assert np.array_equal(nl.astype(bool), nulls)
        with np.errstate(over="ignore"):
            want = fn(a.data.astype(npt), b.data.astype(npt))
        got = v.view(np.float64) if typ == po.REAL else v.astype(npt)
        assert np.array_equal(got[~nulls], want[~nulls])
    for mk, fn in ((Eq, np.equal), (Neq, np.not_equal), (Lt, np.less), (Leq, np.less_equal), (Gt, np.greater),
                   (Geq, np.greater_equal)):
        v, nl = oracle_eval(mk(A, B), [a, b], n)
        assert np.array_equal(nl.astype(bool), nulls)
        assert np.array_equal(v[~nulls].astype(bool), fn(a.data, b.data)[~nulls])
    if typ == po.REAL:
        v, nl = oracle_eval(Div(A, B), [a, b], n)
        with np.errstate(divide="ignore", invalid="ignore"):
            want = a.data / b.data
        got = v.view(np.float64)
        assert np.array_equal(got[~nulls], want[~nulls], equal_nan=True)
    else:
        assert oracle_eval(Div(A, B), [a, b], n) is None      # DIV takes REAL only


def test_oracle_boolean_null_rules():
    # All 9 (left, right) combinations of TRUE, FALSE, NULL.
    vals = [(1, 0), (0, 0), (0, 1)]          # (value, null)
    L = np.array([x for x in vals for _ in vals], np.int8)
    R = np.array([y for _ in vals for y in vals], np.int8)
    a = po.Col(po.BOOLEAN, L[:, 0], L[:, 1])
    b = po.Col(po.BOOLEAN, R[:, 0], R[:, 1])
    A, B = ColRef(0, abi.KG_BOOLEAN), ColRef(1, abi.KG_BOOLEAN)
    name = {(1, 0): "T", (0, 0): "F", (0, 1): "N"}

    def run(e):
        v, nl = oracle_eval(e, [a, b], 9)
        return ["N" if nl[i] else ("T" if v[i] else "F") for i in range(9)]
    pairs = [(name[tuple(L[i])], name[tuple(R[i])]) for i in range(9)]
    # AND: NULL when the left is NULL, else FALSE when the left is FALSE, else the right.
    assert run(And([A, B])) == [("N" if l == "N" else ("F" if l == "F" else r)) for l, r in pairs]
    # OR: the right when the left is NULL or FALSE, else TRUE.
    assert run(Or([A, B])) == [(r if l in "NF" else "T") for l, r in pairs]
    assert run(Not(A)) == [{"T": "F", "F": "T", "N": "N"}[l] for l, _ in pairs]
    assert run(IsNull(A)) == [("T" if l == "N" else "F") for l, _ in pairs]
    # CASE WHEN a THEN b ELSE NOT b: a NULL condition takes ELSE.
    assert run(Case(A, B, Not(B))) == [(r if l == "T" else {"T": "F", "F": "T", "N": "N"}[r]) for l, r in pairs]
    # Comparisons of BOOLEANs are NULL when either side is.
    assert run(Eq(A, B)) == [("N" if "N" in (l, r) else ("T" if l == r else "F")) for l, r in pairs]


def test_oracle_conversions_and_year():
    rng = np.random.default_rng(3)
    n = 1000
    i = _rand_col(rng, po.INT, n)
    v, nl = oracle_eval(Float(ColRef(0, abi.KG_INT)), [i], n)
    assert not nl.any()                              # NULL converts to a non-NULL 0.0
    want = np.where(i.nulls.astype(bool), 0.0, i.data.astype(np.float64))
    assert np.array_equal(v.view(np.float64), want)
    days = rng.integers(po.build_date(1600, 1, 1), po.build_date(2400, 12, 31), n).astype(np.int32)
    dn = (rng.random(n) < 0.2).astype(np.int8)
    v, nl = oracle_eval(ExtractYear(ColRef(0, abi.KG_DATE)), [po.Col(po.DATE, days, dn)], n)
    assert np.array_equal(nl, dn)
    years = [datetime.date.fromordinal(int(d) - 1721425).year for d in days]
    assert v[dn == 0].tolist() == [y for y, z in zip(years, dn) if not z]


def test_oracle_strings_like_in():
    rng = np.random.default_rng(5)
    words = ["", "a", "ab", "abc", "b", "ba", "banana", "bandana", "a_c", "a%c", "xyz", "abab",
             "a\nc", "\u00e9a", "a\u20acc", "ban\nana"]
    n = 600
    strs = [words[k] for k in rng.integers(0, len(words), n)]
    tn = (rng.random(n) < 0.15).astype(np.int8)
    dic = ["apple", "banana", "cherry", "date", "apricot", "a\u00e9c", "a\nc"]
    codes = rng.integers(0, len(dic), n).astype(np.int32)
    en = (rng.random(n) < 0.15).astype(np.int8)
    cols = [("text", strs, tn), po.Col(po.ENUM, codes, en, dictionary=dic)]
    S, E = ColRef(0, abi.KG_TEXT), ColRef(1, abi.KG_ENUM)
    for pat in ["a%", "%a", "%an%", "a_c", "_", "", "%", "b%a", "a\\%c", "_a", "a%c", "ban%ana"]:
        # re2's default "." (one code point, never a newline) = Python's without DOTALL
        rx = r"\A" + "".join(".*" if ch == "%" else "." if ch == "_" else re.escape(ch) for ch in pat) + r"\Z"
        v, nl = oracle_eval(Like(S, pat), cols, n)
        assert not nl.any()                          # LIKE on NULL is a non-NULL FALSE
        want = [bool(re.fullmatch(rx, s)) and not tn[i] for i, s in enumerate(strs)]
        assert v.astype(bool).tolist() == want, pat
        v, nl = oracle_eval(Like(E, pat), cols, n)
        want = [bool(re.fullmatch(rx, dic[c])) and not en[i] for i, c in enumerate(codes)]
        assert v.astype(bool).tolist() == want, pat
    for mk, fn in ((StartsWith, str.startswith), (EndsWith, str.endswith), (Contains, lambda s, p: p in s)):
        for p in ["a", "an", "", "ana"]:
            v, nl = oracle_eval(mk(S, Literal(p)), cols, n)
            assert np.array_equal(nl, tn)
            assert [bool(x) for x, z in zip(v, tn) if not z] == [fn(s, p) for s, z in zip(strs, tn) if not z]
            v, nl = oracle_eval(mk(E, Literal(p)), cols, n)
            assert [bool(x) for x, z in zip(v, en) if not z] == [fn(dic[c], p) for c, z in zip(codes, en) if not z]
    for lit in ["b", "banana", "c", ""]:
        v, nl = oracle_eval(Lt(E, Literal(lit)), cols, n)
        assert [bool(x) for x, z in zip(v, en) if not z] == [dic[c].encode() < lit.encode() for c, z in zip(codes, en)
                                                             if not z]
    v, nl = oracle_eval(In(E, [1, 3]), cols, n)
    assert not nl.any()
    assert v.astype(bool).tolist() == [c in (1, 3) and not z for c, z in zip(codes, en)]


@pytest.mark.parametrize("name", sorted(SCAN))
def test_oracle_scan_fixture(name):
    f, typ, flt, _ = SCAN[name]
    col, printed = oracle_col(f, typ)
    rows = selected(flt(ColRef(0, T[typ])), [col], len(printed))
    assert sorted(printed[r] + "|" for r in rows) == sorted(kdb.expected(f"{name}_expected.tbl"))


def group_minmax_count(keys, values, value_nulls, less=None):
    """GROUP BY key with MIN(v), MAX(v), COUNT(*), COUNT(v), groups in first-appearance
    order: NULL values are skipped by MIN/MAX (compile/proxy/aggregator.cc:14-171, ENUM and
    TEXT by string order); COUNT(v) starts at 1 on the group's first row whatever its NULL
    flag and then counts non-NULL values (aggregator.cc:270-298). keys: hashable per row
    (None = the NULL key). Returns printed rows min|max|count1|count2|."""
    g = {}
    for i, k in enumerate(keys):
        v = None if value_nulls is not None and value_nulls[i] else values[i]
        if k not in g:
            g[k] = [v, v, 1, 1]
            continue
        st = g[k]
        if v is not None:
            st[0] = v if st[0] is None or v < st[0] else st[0]
            st[1] = v if st[1] is None or v > st[1] else st[1]
            st[3] += 1
        st[2] += 1
    return [f"{'' if a is None else a}|{'' if b is None else b}|{c}|{d}|" for a, b, c, d in g.values()]


def groupby_enum_rows():
    """group_by_aggregate/enum_test.cc (EnumSchema): GROUP BY people.name with MIN/MAX(name),
    COUNT(*), COUNT(name)."""
    codes, nulls = kdb.col("people_name_enum", "enum")
    d = kdb.enum_dict("people_name_enum")
    keys = [None if nulls is not None and nulls[i] else int(c) for i, c in enumerate(codes)]
    return group_minmax_count(keys, [d[int(c)] for c in codes], nulls)


def groupby_text_rows():
    """group_by_aggregate/text_test.cc: GROUP BY CASE(id > 500, 'hello', '') with
    MIN/MAX(name), COUNT(*), COUNT(name); a NULL id takes the ELSE branch."""
    ids, idn = kdb.col("people_id", "int")
    names, nn = kdb.text("people_name")
    keys = ["hello" if (idn is None or not idn[i]) and ids[i] > 500 else "" for i in range(len(ids))]
    return group_minmax_count(keys, names, nn)


@pytest.mark.parametrize("t", ["enum", "text"])
def test_oracle_group_by_string_fixture(t):
    got = groupby_enum_rows() if t == "enum" else groupby_text_rows()
    assert sorted(got) == sorted(kdb.expected(f"group_by_aggregate/{t}_expected.tbl"))
