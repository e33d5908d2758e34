"""The reference's end-to-end fixture queries, restated over an engine.

Each query mirrors one end_to_end_test/<op>/<type>_test.cc plan and prints
rows in the reference's text format; the tests sort lines and compare with the
checked-in *_expected.tbl exactly as the reference tests do
(std::sort + EXPECT_EQ), or with CHECK_EQ_TBL's absolute 1e-5 tolerance on
REAL columns for TPC-H (util/test_util.h:82-175).

An engine provides:
  select(cols, preds, conj)      -> ascending row ids   (cols: [(typ, values, nulls)])
  join(build_col, probe_col)     -> (probe_rows, build_rows) pairs
  group_agg(key_codes, card, aggs, rows) -> [(vals, nulls)] per group, insertion order
       aggs: [(kind, (typ, values, nulls)) or (COUNT, None)]
  aggregate(aggs, n, rows)       -> (vals, nulls) or None for empty input
"""
import numpy as np

import kdb

SUM, MIN, MAX, AVG, COUNT = range(5)
EQ, NEQ, LT, LEQ, GT, GEQ = range(6)


def _c(name, typ):
    v, n = kdb.col(name, typ)
    return (typ, v, n)


def _date(y, m, d):
    from oracle import pyoracle as po
    return po.build_date(y, m, d)


def _print_val(typ, v, null):
    if typ in ("real",) or typ == "avg":
        return kdb.fmt("real", kdb.bits_f64(v) if isinstance(v, (int, np.integer)) else v, null)
    return kdb.fmt(typ, v, null)


# ---------------------------------------------------------------- hash join
HASH_JOIN = {
    # test -> (build column, probe column, type): end_to_end_test/hash_join/<t>_test.cc:40-80
    "int": ("people_id", "info_id", "int"),
    "bigint": ("info_num2", "info_num2", "bigint"),
    "smallint": ("info_num1", "info_num1", "smallint"),
    "date": ("info_date", "info_date", "date"),
    "real": ("info_zscore", "info_zscore", "real"),
    "boolean": ("info_cheated", "info_cheated", "boolean"),
    "enum": ("people_name_enum", "people_name_enum", "enum"),   # EnumSchema people.name
    "text": ("people_name", "people_name", "text"),
}


def hash_join(engine, t):
    b, p, typ = HASH_JOIN[t]
    if typ == "text":
        bs, ps = kdb.text(b), kdb.text(p)
        pr, br = engine.join_text(bs, ps)
        return [kdb.row([bs[0][br[i]], ps[0][pr[i]]]) for i in range(len(pr))]
    bc, pc = _c(b, typ), _c(p, typ)
    pr, br = engine.join(bc, pc)
    if typ == "enum":
        d = kdb.enum_dict(b)
        return [kdb.row([d[int(bc[1][br[i]])], d[int(pc[1][pr[i]])]]) for i in range(len(pr))]
    lines = []
    for i in range(len(pr)):
        lines.append(kdb.row([kdb.fmt(typ, bc[1][br[i]]), kdb.fmt(typ, pc[1][pr[i]])]))
    return lines


# -------------------------------------------------------------- scan select
SCAN_SELECT = {
    # end_to_end_test/scan_select/<t>_test.cc
    "int": ([("info_id", "int")], [(0, GT, 29)]),
    "bigint": ([("info_num2", "bigint")], [(0, GT, 0)]),
    "date": ([("info_date", "date")], [(0, GEQ, ("date", 2021, 1, 1))]),
    "real": ([("info_zscore", "real")], [(0, GT, -0.5)]),
    "smallint": ([("info_num1", "smallint")], [(0, LT, 0)]),
    "multiple_eq": ([("info_cheated", "boolean"), ("info_date", "date")],
                    [(0, EQ, 1), (1, EQ, ("date", 2021, 1, 29))]),
}


def _lit(x):
    if isinstance(x, tuple):
        return _date(*x[1:])
    return x


def scan_select(engine, t):
    spec, preds = SCAN_SELECT[t]
    cols = [_c(n, typ) for n, typ in spec]
    rows = engine.select(cols, [(c, op, _lit(l)) for c, op, l in preds], True)
    out_cols = cols[-1:] if t != "multiple_eq" else cols
    return [kdb.row([kdb.fmt(c[0], c[1][r]) for c in out_cols]) for r in rows]


# ------------------------------------------------------ group-by aggregate
GROUP_BY = {
    # end_to_end_test/group_by_aggregate/<t>_test.cc: key = CASE(cheated, a, b),
    # aggregates over the measure column; output omits the key.
    "int": ("info_id", "int", [SUM, MIN, MAX, AVG, "count*", COUNT]),
    "bigint": ("info_num2", "bigint", [SUM, MIN, MAX, AVG, "count*", COUNT]),
    "smallint": ("info_num1", "smallint", [SUM, MIN, MAX, AVG, "count*", COUNT]),
    "real": ("info_zscore", "real", [SUM, MIN, MAX, AVG, "count*", COUNT]),
    "date": ("info_date", "date", [MIN, MAX, "count*", COUNT]),
}


def case_codes():
    """CASE(cheated, then, else): NULL or false -> else (code 1), true -> then (code 0)
    (compile/translators/expression_translator.cc:170-181)."""
    ch, chn = kdb.col("info_cheated", "boolean")
    code = np.where((chn == 0) & (ch != 0), 0, 1).astype(np.int32)
    return code


def _print_aggs(kinds, mtyp, vals, nulls):
    out = []
    for k, v, nl in zip(kinds, vals, nulls):
        if k in ("count*", COUNT):
            out.append(kdb.fmt("bigint", v, nl))
        elif k == AVG:
            out.append(kdb.fmt("real", kdb.bits_f64(v), nl))
        elif mtyp == "real":
            out.append(kdb.fmt("real", kdb.bits_f64(v), nl))
        else:
            out.append(kdb.fmt(mtyp, v, nl))
    return kdb.row(out)


def _agg_list(kinds, m):
    return [(COUNT, None) if k == "count*" else (k, m) for k in kinds]


def group_by(engine, t):
    name, mtyp, kinds = GROUP_BY[t]
    m = _c(name, mtyp)
    groups = engine.group_agg(case_codes(), 2, _agg_list(kinds, m), None)
    return [_print_aggs(kinds, mtyp, v, nl) for v, nl in groups]


def aggregate(engine, t):
    """end_to_end_test/aggregate/{int,empty}_test.cc: ungrouped SUM/MIN/MAX/AVG/COUNT(*)/COUNT(id);
    `empty` filters every row out (Select(Literal(false)))."""
    m = _c("info_id", "int")
    kinds = [SUM, MIN, MAX, AVG, "count*", COUNT]
    rows = np.zeros(0, dtype=np.int32) if t == "empty" else None
    res = engine.aggregate(_agg_list(kinds, m), len(m[1]), rows)
    if res is None:
        return [kdb.row([""] * len(kinds))]
    vals, nulls = res
    return [_print_aggs(kinds, "int", vals, nulls)]


# -------------------------------------------------------- SIMD scan select
def simd_scan_select(engine, t):
    """end_to_end_test/simd_scan_select/{int,date_range,multiple_col}_test.cc."""
    if t == "int":
        c = _c("n_nationkey", "int")
        rows = engine.select([c], [(0, GT, 5)], True)
        return [kdb.row([kdb.fmt("int", c[1][r])]) for r in rows]
    if t == "date_range":
        c = _c("l_shipdate", "date")
        rows = engine.select([c], [(0, LEQ, _date(1998, 12, 31)), (0, GEQ, _date(1998, 1, 1))], True)
        return [kdb.row([kdb.fmt("date", c[1][r])]) for r in rows]
    if t == "multiple_col":
        d = _c("l_shipdate", "date")
        rf = _c("l_returnflag_enum", "enum")
        names = kdb.enum_dict("l_returnflag_enum")
        code_r = names.index("R")
        rows = engine.select([d, rf], [(0, LEQ, _date(1994, 12, 31)), (0, GEQ, _date(1994, 1, 1)),
                                      (1, EQ, code_r)], True)
        return [kdb.row([kdb.fmt("date", d[1][r]), names[rf[1][r]]]) for r in rows]
    raise KeyError(t)


# ------------------------------------------------------------------ TPC-H Q1
def tpch_q01(engine):
    """end_to_end_test/tpch/q01_test.cc with the dictionary-encoded flags:
    ScanSelect(l_shipdate <= 1998-09-13) -> GroupBy(l_returnflag, l_linestatus;
    SUM(qty), SUM(price), SUM(price*(1-disc)), SUM(price*(1-disc)*(1+tax)),
    AVG(qty), AVG(price), AVG(disc), COUNT(*)) -> OrderBy(flag, status).
    The two derived measures are materialized as REAL columns first."""
    ship = _c("l_shipdate", "date")
    rf = _c("l_returnflag_enum", "enum")
    ls = _c("l_linestatus_enum", "enum")
    qty = _c("l_quantity", "real")
    price = _c("l_extendedprice", "real")
    disc = _c("l_discount", "real")
    tax = _c("l_tax", "real")
    disc_price = ("real", price[1] * (1.0 - disc[1]), None)
    charge = ("real", disc_price[1] * (1.0 + tax[1]), None)
    rows = engine.select([ship], [(0, LEQ, _date(1998, 9, 13))], True)
    rfn, lsn = kdb.enum_dict("l_returnflag_enum"), kdb.enum_dict("l_linestatus_enum")
    codes = rf[1] * len(lsn) + ls[1]
    aggs = [(SUM, qty), (SUM, price), (SUM, disc_price), (SUM, charge), (AVG, qty), (AVG, price), (AVG, disc),
            (COUNT, None)]
    groups = engine.group_agg(codes.astype(np.int32), len(rfn) * len(lsn), aggs, rows, keys_out=True)
    out = []
    for code, vals, nulls in groups:
        f, s = rfn[code // len(lsn)], lsn[code % len(lsn)]
        nums = [kdb.bits_f64(v) for v in vals[:7]] + [int(vals[7])]
        out.append((f, s, nums))
    out.sort(key=lambda x: (x[0], x[1]))  # ORDER BY on ENUM compares decoded strings
    return out


def check_q01(got, expected_lines, tol=1e-5):
    assert len(got) == len(expected_lines)
    for (f, s, nums), line in zip(got, expected_lines):
        e = line.split("|")
        assert e[0] == f and e[1] == s
        for j in range(7):
            assert abs(float(e[2 + j]) - nums[j]) < tol, (line, nums)
        assert int(e[9]) == nums[7]


# ------------------------------------------------------------- skinner join
SKINNER = {
    # end_to_end_test/skinner_join/<t>_test.cc: two scans of one column joined on equality,
    # output both columns (enum: the EnumSchema people.name, printed as strings)
    "int": ("info_id", "int"),
    "smallint": ("info_num1", "smallint"),
    "date": ("info_date", "date"),
    "bigint": ("info_num2", "bigint"),
    "real": ("info_zscore", "real"),
    "boolean": ("info_cheated", "boolean"),
    "enum": ("people_name_enum", "enum"),
    "text": ("people_name", "text"),
}


def skinner(engine, t, **kw):
    """Result lines of a skinner_join fixture query on `engine`."""
    if t == "multiple_tables_join":
        # multiple_tables_join_test.cc: four scans of info(id, num2);
        # s1.id = s2.id, s2.num2 = s3.num2, s2.id = s4.id; output the four ids.
        ident, num2 = _c("info_id", "int"), _c("info_num2", "bigint")
        n = len(ident[1])
        tuples = engine.skinner([n] * 4, [(0, ident, 1, ident), (1, num2, 2, num2), (1, ident, 3, ident)], **kw)
        return [kdb.row([kdb.fmt("int", ident[1][tp[i]], ident[2] is not None and ident[2][tp[i]])
                         for i in range(4)]) for tp in tuples]
    name, typ = SKINNER[t]
    c = ("text", *kdb.text(name)) if typ == "text" else _c(name, typ)
    n = len(c[1])
    tuples = engine.skinner([n, n], [(0, c, 1, c)], **kw)
    if typ == "enum":
        d = kdb.enum_dict(name)
        return [kdb.row([d[int(c[1][tp[0]])], d[int(c[1][tp[1]])]]) for tp in tuples]
    if typ == "text":
        return [kdb.row([c[1][tp[0]], c[1][tp[1]]]) for tp in tuples]
    return [kdb.row([kdb.fmt(typ, c[1][tp[i]]) for i in range(2)]) for tp in tuples]


def cross_product(engine):
    """cross_product_test.cc: region.r_name x region.r_name (CrossProductOperator), as a
    join with no conditions."""
    names, nulls = kdb.text("r_name")
    n = len(names)
    tuples = engine.skinner([n, n], [])
    return [kdb.row([names[a], names[b]]) for a, b in tuples]
