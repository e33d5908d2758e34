"""ORDER BY oracle (kor_order_by) pinned three ways: the reference's own
order_by fixtures (end_to_end_test/order_by/*_test.cc: expected .tbl ascending,
the same lines reversed descending); the reference's Vector::Sort
(runtime/vector.cc) compiled into oracle/_ref, on random keys with ties, NULLs
and -0.0; and Python's stable sorted() of the reversed input (rows equal on
every key come out in reverse input order) on small inputs with TEXT / ENUM
keys. Also the result-row formatting of the product library (kg_format_rows,
host code) against the printed fixture values, and kg_enum_ranks."""
import numpy as np
import pytest

import kdb
from oracle import pyoracle as po

T = {"int": po.INT, "bigint": po.BIGINT, "smallint": po.SMALLINT, "date": po.DATE, "real": po.REAL,
     "enum": po.ENUM}

# fixture: (ORDER BY keys, printed columns), columns as (file, type)
FIXTURES = {
    "int": ([("info_id", "int")], [("info_id", "int")]),
    "bigint": ([("info_num2", "bigint"), ("info_id", "int")], [("info_num2", "bigint"), ("info_id", "int")]),
    "smallint": ([("info_num1", "smallint"), ("info_id", "int")], [("info_num1", "smallint"), ("info_id", "int")]),
    "date": ([("info_date", "date"), ("info_id", "int")], [("info_date", "date"), ("info_id", "int")]),
    "real": ([("info_zscore", "real"), ("info_id", "int")], [("info_zscore", "real"), ("info_id", "int")]),
    "enum": ([("people_name_enum", "enum"), ("people_id", "int")],
             [("people_id", "int"), ("people_name_enum", "enum")]),
    "text": ([("people_name", "text"), ("people_id", "int")], [("people_id", "int"), ("people_name", "text")]),
}


def load(name, typ):
    """(oracle key, printed strings per row)."""
    if typ == "text":
        strs, nulls = kdb.text(name)
        printed = ["" if nulls is not None and nulls[i] else s for i, s in enumerate(strs)]
        return ("text", strs, nulls), printed
    vals, nulls = kdb.col(name, typ)
    d = kdb.enum_dict(name) if typ == "enum" else None
    col = po.Col(T[typ], vals, nulls, dictionary=d)
    printed = []
    for i, v in enumerate(vals):
        null = nulls is not None and bool(nulls[i])
        printed.append("" if null else (d[v] if typ == "enum" else kdb.fmt(typ, v)))
    return col, printed


@pytest.mark.parametrize("asc", [True, False])
@pytest.mark.parametrize("fixture", sorted(FIXTURES))
def test_oracle_order_by_fixture(fixture, asc):
    keys, outs = FIXTURES[fixture]
    order = po.order_by([(load(n, t)[0], asc) for n, t in keys])
    printed = [load(n, t)[1] for n, t in outs]
    got = ["".join(p[r] + "|" for p in printed) for r in order]
    exp = kdb.expected(f"order_by/{fixture}_expected.tbl")
    if not asc:
        exp = exp[::-1]
    assert got == exp


def _random_cols(rng, n, card):
    cols = []
    for typ in (po.SMALLINT, po.INT, po.BIGINT, po.REAL, po.DATE):
        if typ == po.REAL:
            v = rng.integers(-card, card, n).astype(np.float64) / 4
            v[rng.random(n) < 0.1] = -0.0
            v[rng.random(n) < 0.1] = 0.0
        else:
            v = rng.integers(-card, card, n)
        nulls = (rng.random(n) < 0.15).astype(np.int8)
        cols.append(po.Col(typ, v, nulls))
    return cols


@pytest.mark.skipif(not po.ref_built(), reason="oracle/_ref not built (reference checkout absent)")
@pytest.mark.parametrize("seed", range(6))
def test_oracle_order_by_vs_reference_sort(seed):
    rng = np.random.default_rng(seed)
    n = [1, 2, 7, 100, 1000, 5000][seed]
    cols = _random_cols(rng, n, 3 + seed)
    for trial in range(8):
        k = 1 + trial % 3
        pick = rng.choice(len(cols), k, replace=False)
        keys = [(cols[i], bool(rng.integers(2))) for i in pick]
        assert np.array_equal(po.order_by(keys), po.ref_sort_rows(keys)), (seed, trial)


def _py_key(kind, vals, nulls, dictionary, asc):
    """Sort key of one column for sorted(): NULL greatest; desc by negation of
    the rank among the distinct non-NULL values (works for strings too)."""
    conv = (lambda v: dictionary[v].encode()) if kind == "enum" else (
        (lambda v: v.encode()) if kind == "text" else (lambda v: v))
    distinct = sorted({conv(v) for i, v in enumerate(vals) if not (nulls is not None and nulls[i])})
    rank = {v: i for i, v in enumerate(distinct)}

    def key(r):
        null = nulls is not None and bool(nulls[r])
        if null:
            return (1 if asc else -1, 0)
        x = rank[conv(vals[r])]
        return (0, x if asc else -x)
    return key


@pytest.mark.parametrize("seed", range(10))
def test_oracle_order_by_vs_stable_sorted(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(0, 60))
    words = ["", "a", "ab", "b", "abc", "B", "é", "aé", "zz"]
    dic = ["pear", "apple", "fig", "apple pie", "Fig"]
    tv = [words[i] for i in rng.integers(0, len(words), n)]
    tn = (rng.random(n) < 0.2).astype(np.int8)
    ev = rng.integers(0, len(dic), n).astype(np.int32)
    en = (rng.random(n) < 0.2).astype(np.int8)
    iv = rng.integers(-2, 3, n).astype(np.int32)
    rv = rng.choice([-1.5, -0.0, 0.0, 2.25], n)
    cols = {"text": (("text", tv, tn), tv, tn, None), "enum": (po.Col(po.ENUM, ev, en, dictionary=dic), ev, en, dic),
            "int": (po.Col(po.INT, iv), iv, None, None), "real": (po.Col(po.REAL, rv), rv, None, None)}
    for trial in range(6):
        names = list(rng.permutation(list(cols))[: 1 + trial % 3])
        ascs = [bool(rng.integers(2)) for _ in names]
        keys = [(cols[c][0], a) for c, a in zip(names, ascs)]
        fns = [_py_key(c, cols[c][1], cols[c][2], cols[c][3], a) for c, a in zip(names, ascs)]
        want = sorted(range(n - 1, -1, -1), key=lambda r: tuple(f(r) for f in fns))
        assert po.order_by(keys).tolist() == want, (seed, trial, names, ascs)


def test_oracle_order_by_boolean_key_rejected():
    assert po.order_by([(po.Col(po.BOOLEAN, np.array([1, 0], np.int8)), True)]) is None


def test_oracle_order_by_empty():
    assert po.order_by([(po.Col(po.INT, np.array([], np.int32)), True)]).size == 0
