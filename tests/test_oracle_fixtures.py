"""The oracle (CPU restatement) against the reference's own end-to-end
fixtures: the checked-in column files and *_expected.tbl outputs."""
import pytest

import kdb
import queries
from engines import OracleEngine

E = OracleEngine()


@pytest.mark.parametrize("t", sorted(queries.HASH_JOIN))
def test_hash_join(t):
    assert sorted(queries.hash_join(E, t)) == sorted(kdb.expected(f"hash_join/{t}_expected.tbl"))


@pytest.mark.parametrize("t", sorted(queries.SCAN_SELECT))
def test_scan_select(t):
    assert sorted(queries.scan_select(E, t)) == sorted(kdb.expected(f"scan_select/{t}_expected.tbl"))


@pytest.mark.parametrize("t", sorted(queries.GROUP_BY))
def test_group_by(t):
    assert sorted(queries.group_by(E, t)) == sorted(kdb.expected(f"group_by_aggregate/{t}_expected.tbl"))


@pytest.mark.parametrize("t", ["int", "empty"])
def test_aggregate(t):
    assert sorted(queries.aggregate(E, t)) == sorted(kdb.expected(f"aggregate/{t}_expected.tbl"))


@pytest.mark.parametrize("t", ["int", "date_range", "multiple_col"])
def test_simd_scan_select(t):
    assert sorted(queries.simd_scan_select(E, t)) == sorted(kdb.expected(f"simd_scan_select/{t}_expected.tbl"))


def test_tpch_q01():
    queries.check_q01(queries.tpch_q01(E), kdb.expected("tpch/q01_expected.tbl"))


def test_tpch_q03():
    import tpch_q03
    tpch_q03.check(tpch_q03.oracle_q03(tpch_q03.columns()), kdb.expected("tpch/q03_expected.tbl"))


def test_tpch_q12():
    import tpch_q12
    assert tpch_q12.lines(tpch_q12.oracle_q12(tpch_q12.columns())) == kdb.expected("tpch/q12_expected.tbl")


def test_tpch_q05():
    import tpch_q05
    tpch_q05.check(tpch_q05.oracle_q05(tpch_q05.columns()), kdb.expected("tpch/q05_expected.tbl"))


def test_tpch_q14():
    import tpch_q14
    assert abs(tpch_q14.oracle_q14(tpch_q14.columns()) - float(kdb.expected("tpch/q14_expected.tbl")[0])) < 1e-5


def test_tpch_q19():
    import tpch_q19
    assert tpch_q19.line(tpch_q19.oracle_q19(tpch_q19.columns())) == kdb.expected("tpch/q19_expected.tbl")[0]


def test_tpch_q18():
    import tpch_q18
    assert tpch_q18.oracle_q18(tpch_q18.columns()) == kdb.expected("tpch/q18_expected.tbl") == []


def test_tpch_q10():
    import tpch_q10
    tpch_q10.check(tpch_q10.oracle_q10(tpch_q10.columns()), kdb.expected("tpch/q10_expected.tbl"))


def test_tpch_q11():
    import tpch_q11
    tpch_q11.check(tpch_q11.oracle_q11(tpch_q11.columns()), kdb.expected("tpch/q11_expected.tbl"))


def test_tpch_q07():
    import tpch_q07
    tpch_q07.check(tpch_q07.oracle_q07(tpch_q07.columns()), kdb.expected("tpch/q07_expected.tbl"))


def test_tpch_q09():
    import tpch_q09
    tpch_q09.check(tpch_q09.oracle_q09(tpch_q09.columns()), kdb.expected("tpch/q09_expected.tbl"))
