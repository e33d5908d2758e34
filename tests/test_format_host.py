"""Result-row printing (kg_format_rows, host code of the product library) and
ENUM ranking (kg_enum_ranks): the reference's order_by fixtures printed in the
oracle's order must equal the expected .tbl text; random values of every type
print as the reference's printer does (runtime/printer.cc:13-35: REAL with
std::fixed six decimals, DATE as YYYY-MM-DD from the Julian day number, NULL as
the empty string); ENUM ranks follow the byte order of the decoded strings."""
import ctypes

import numpy as np
import pytest

import kdb
from kushdb_amd import abi
from kushdb_amd.runtime import enum_ranks, format_rows
from oracle import pyoracle as po
from test_oracle_order import FIXTURES, T, load

KG = {"int": abi.KG_INT, "bigint": abi.KG_BIGINT, "smallint": abi.KG_SMALLINT, "date": abi.KG_DATE,
      "real": abi.KG_REAL, "enum": abi.KG_ENUM, "text": abi.KG_TEXT, "boolean": abi.KG_BOOLEAN}


def fmt_column(name, typ):
    if typ == "text":
        strs, nulls = kdb.text(name)
        return (abi.KG_TEXT, strs, nulls, None)
    vals, nulls = kdb.col(name, typ)
    return (KG[typ], vals, nulls, kdb.enum_dict(name) if typ == "enum" else None)


@pytest.mark.parametrize("fixture", sorted(FIXTURES))
def test_format_fixture_rows(fixture):
    keys, outs = FIXTURES[fixture]
    order = po.order_by([(load(n, t)[0], True) for n, t in keys])
    text = format_rows([fmt_column(n, t) for n, t in outs], order)
    assert text.endswith("\n")
    assert text.split("\n")[:-1] == kdb.expected(f"order_by/{fixture}_expected.tbl")


def test_format_random_values():
    rng = np.random.default_rng(7)
    n = 3000
    cols = {
        "smallint": rng.integers(-32768, 32768, n).astype(np.int16),
        "int": rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32),
        "bigint": rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64),
        "real": np.concatenate([rng.standard_normal(n - 6) * 10.0 ** rng.integers(-8, 12, n - 6),
                                [-0.0, 0.0, 0.0000005, -0.0000005, 1e15, -2.5]]),
        "date": rng.integers(po.build_date(1700, 1, 1), po.build_date(2400, 12, 31), n).astype(np.int32),
        "boolean": rng.integers(0, 2, n).astype(np.int8),
    }
    nulls = {k: (rng.random(n) < 0.1).astype(np.int8) for k in cols}
    spec = [(KG[k], v, nulls[k], None) for k, v in cols.items()]
    lines = format_rows(spec).split("\n")[:-1]
    assert len(lines) == n
    for i, ln in enumerate(lines):
        want = "".join(kdb.fmt(k, v[i], bool(nulls[k][i])) + "|" for k, v in cols.items())
        assert ln == want, i


def test_format_enum_text_and_order():
    dic = ["x", "", "yy"]
    codes = np.array([2, 0, 1, 2], np.int32)
    txt = ["a|b", "", "héllo", "z"]
    got = format_rows([(abi.KG_ENUM, codes, None, dic), (abi.KG_TEXT, txt, np.array([0, 0, 0, 1], np.int8), None)],
                      order=np.array([3, 1, 0, 2], np.int32))
    assert got == "yy||\nx||\nyy|a|b|\n|héllo|\n"
    assert format_rows([(abi.KG_INT, np.zeros(0, np.int32), None, None)]) == ""


def test_format_capacity_and_bad_order():
    lib = abi.load()
    vals = np.arange(5, dtype=np.int32)
    fc = (abi.KgFmtCol * 1)(abi.KgFmtCol(abi.KG_INT, 0, vals.ctypes.data, 0, 0, 0))
    ln = ctypes.c_int64()
    buf = ctypes.create_string_buffer(4)
    assert lib.kg_format_rows(fc, 1, 5, None, buf, 4, ctypes.byref(ln)) == abi.KG_ERR_CAPACITY
    assert ln.value == 15
    bad = np.array([0, 1, 2, 3, 9], np.int32)
    assert lib.kg_format_rows(fc, 1, 5, bad.ctypes.data, None, 0, ctypes.byref(ln)) == abi.KG_ERR_INVALID


def test_enum_ranks_byte_order():
    d = kdb.enum_dict("people_name_enum")
    r = enum_ranks(d)
    assert sorted(r.tolist()) == list(range(len(d)))
    by_rank = [None] * len(d)
    for code, k in enumerate(r):
        by_rank[k] = d[code].encode()
    assert by_rank == sorted(x.encode() for x in d)
    small = enum_ranks(["b", "a", "B", "ab", ""])
    assert small.tolist() == [4, 2, 1, 3, 0]
