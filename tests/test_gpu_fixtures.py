"""The HIP path (libkushgpu.so via its C ABI) on the reference's own
end-to-end fixtures, compared with the checked-in *_expected.tbl files."""
import pytest

import kdb
import queries

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from engines import GpuEngine
    return GpuEngine()


@pytest.mark.parametrize("t", sorted(queries.HASH_JOIN))
def test_hash_join(E, t):
    assert sorted(queries.hash_join(E, t)) == sorted(kdb.expected(f"hash_join/{t}_expected.tbl"))


@pytest.mark.parametrize("t", sorted(queries.SCAN_SELECT))
def test_scan_select(E, t):
    assert sorted(queries.scan_select(E, t)) == sorted(kdb.expected(f"scan_select/{t}_expected.tbl"))


@pytest.mark.parametrize("t", sorted(queries.GROUP_BY))
def test_group_by(E, t):
    assert sorted(queries.group_by(E, t)) == sorted(kdb.expected(f"group_by_aggregate/{t}_expected.tbl"))


@pytest.mark.parametrize("t", ["int", "empty"])
def test_aggregate(E, t):
    assert sorted(queries.aggregate(E, t)) == sorted(kdb.expected(f"aggregate/{t}_expected.tbl"))


@pytest.mark.parametrize("t", ["int", "date_range", "multiple_col"])
def test_simd_scan_select(E, t):
    assert sorted(queries.simd_scan_select(E, t)) == sorted(kdb.expected(f"simd_scan_select/{t}_expected.tbl"))


def test_tpch_q01(E):
    queries.check_q01(queries.tpch_q01(E), kdb.expected("tpch/q01_expected.tbl"))
