"""Pins the oracle restatement independently of its own formulas:
  * against the reference's runtime sources compiled into oracle/_ref
    (HashCombine, dates, FastForwardBucket, sorted intersection, ART dedup and
    iteration order, the chained join table, the aggregate table's group
    order) — skipped when that library was not built;
  * against the reference's unit-test known answers
    (runtime/column_index_bucket_test.cc:8-27, tuple_idx_table_test.cc:8-71,
    date_test.cc);
  * against closed forms / numpy (np.searchsorted, np.intersect1d, bincount,
    a Python left fold for FP64 SUM and the running-mean AVG).
"""
import ctypes
import datetime
import math

import numpy as np
import pytest

from oracle import pyoracle as po

# The pin library is opened inside the tests that use it, never at import, so
# collecting this module loads no reference code.
needs_ref = pytest.mark.skipif(not po.ref_built(), reason="oracle/_ref not built (reference checkout absent)")


class _Ref:
    def __getattr__(self, name):
        return getattr(po.ref(), name)


ref = _Ref()


# ---------------------------------------------------------------- hashing
def test_hash_combine_known_answers():
    # h = 0: 0 ^ (v + 0x9e3779b9) truncated to int32.
    assert po.hash_combine(0, 5) == np.int64(5 + 0x9e3779b9).astype(np.uint32).astype(np.int32)
    # Two rounds by hand, mod 2^32 with an arithmetic shift of the signed hash.
    h1 = po.hash_combine(0, 7)
    u = h1 & 0xFFFFFFFF
    mix = (11 + 0x9e3779b9 + ((u << 6) & 0xFFFFFFFF) + ((h1 >> 2) & 0xFFFFFFFF)) & 0xFFFFFFFF
    want = np.uint32(u ^ mix).astype(np.int32)
    assert po.hash_combine(h1, 11) == want


@needs_ref
def test_hash_combine_vs_reference():
    rng = np.random.default_rng(1)
    for h0, v in zip(rng.integers(-2**31, 2**31, 2000), rng.integers(-2**63, 2**63 - 1, 2000, dtype=np.int64)):
        x = ctypes.c_int32(int(h0))
        ref.ref_hash_combine(ctypes.byref(x), int(v))
        assert po.hash_combine(int(h0), int(v)) == x.value


def test_agg_hash_murmur_constants():
    # Single key: k*m, k ^= k >> 47 (logical), k*m; h = (0 ^ k) * m + 0xe6546b64.
    m = 0xc6a4a7935bd1e995
    M = (1 << 64) - 1
    for key in [0, 1, 12345, 2**40 + 3]:
        k = (key * m) & M
        k ^= k >> 47
        k = (k * m) & M
        h = ((k * m) + 0xe6546b64) & M
        arr = np.array([key], dtype=np.int64)
        got = po.lib().kor_agg_hash(arr.ctypes.data, None, 1) & M
        assert got == h


# ------------------------------------------------------------------ dates
def test_dates_known_answers_and_round_trip():
    # date_test.cc: ordering, year extraction, round trip.
    assert po.build_date(2021, 1, 1) < po.build_date(2021, 1, 2)
    assert po.split_date(po.build_date(1998, 9, 13)) == (1998, 9, 13)
    # Julian day numbers are proleptic Gregorian ordinals + 1721425.
    d = datetime.date(1900, 1, 1)
    while d.year < 2101:
        jd = po.build_date(d.year, d.month, d.day)
        assert jd == d.toordinal() + 1721425
        assert po.split_date(jd) == (d.year, d.month, d.day)
        d += datetime.timedelta(days=37)


@needs_ref
def test_dates_vs_reference():
    for y in range(1900, 2100, 7):
        for m in range(1, 13):
            for dd in (1, 15, 28):
                assert po.build_date(y, m, dd) == ref.ref_build_date(y, m, dd)


# --------------------------------------------------- column-index buckets
def test_fast_forward_known_answers():
    data = [0, 4, 5, 8, 9]
    want = {0: 0, 1: 1, 2: 1, 3: 1, 4: 1, 5: 2, 6: 3, 7: 3, 8: 3, 9: 4, 10: 2**31 - 1, 11: 2**31 - 1}
    for prev, w in want.items():
        assert po.fast_forward(data, prev) == w
    assert po.fast_forward([], 3) == 2**31 - 1


def test_fast_forward_is_lower_bound():
    rng = np.random.default_rng(2)
    a = np.unique(rng.integers(0, 1000, 300)).astype(np.int32)
    for prev in range(-5, 1010, 3):
        i = int(np.searchsorted(a, prev, side="left"))
        assert po.fast_forward(a, prev) == (i if i < len(a) else 2**31 - 1)


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_intersection_closed_form(k):
    rng = np.random.default_rng(10 + k)
    lists = [np.unique(rng.integers(0, 5000, 1500)).astype(np.int32) for _ in range(k)]
    want = lists[0]
    for x in lists[1:]:
        want = np.intersect1d(want, x)
    for nt in (0, 700):
        # One chunk: exactly the intersection.
        got = po.intersect_all(lists, next_tuple=nt, chunk=1 << 20)
        assert np.array_equal(got, want[want >= nt])
        # 64-id chunks: with 3+ lists the reference re-emits the last id of a
        # chunk when the shortest list changes between chunks (the other lists
        # were only fast-forwarded to it); the set is unchanged and the
        # Skinner result dedup removes the repeat.
        got = po.intersect_all(lists, next_tuple=nt, chunk=64)
        assert np.array_equal(np.unique(got), want[want >= nt])
        if k <= 2:
            assert np.array_equal(got, want[want >= nt])


@needs_ref
@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_intersection_vs_reference(k):
    rng = np.random.default_rng(20 + k)
    for trial in range(20):
        lists = [np.unique(rng.integers(0, 300, rng.integers(0, 200))).astype(np.int32) for _ in range(k)]
        nt = int(rng.integers(0, 50))
        a = po.intersect_all(lists, next_tuple=nt, chunk=int(rng.integers(1, 70)))
        b = po.intersect_all(lists, next_tuple=nt, chunk=64, fn=ref.ref_intersect_all)
        assert np.array_equal(a, b)


# -------------------------------------------------------- tuple-id dedup
def test_tuple_dedup_known_answers():
    t = np.array([[1, 0, 2], [0, 1, 2], [1, 2, 0], [1, 2, 0], [0, 1, 2], [1, 0, 2]], dtype=np.int32)
    is_new, order = po.tuple_dedup(t)
    assert is_new.tolist() == [True, True, True, False, False, False]
    assert order.tolist() == [[0, 1, 2], [1, 0, 2], [1, 2, 0]]


@needs_ref
def test_tuple_dedup_vs_reference():
    rng = np.random.default_rng(5)
    t = rng.integers(0, 600, (4000, 5)).astype(np.int32)
    t[2000:] = t[rng.integers(0, 2000, 2000)]
    a_new, a_ord = po.tuple_dedup(t)
    b_new, b_ord = po.tuple_dedup(t, fn=ref.ref_tuple_dedup)
    assert np.array_equal(a_new, b_new)
    assert np.array_equal(a_ord, b_ord)


# ---------------------------------------------------------------- joins
def test_hash_join_nested_loop():
    rng = np.random.default_rng(3)
    b = rng.integers(-20, 20, 300).astype(np.int64)
    p = rng.integers(-25, 25, 400).astype(np.int64)
    pr, br = po.hash_join([po.Col(po.BIGINT, b)], [po.Col(po.BIGINT, p)])
    want = [(j, i) for j in range(len(p)) for i in range(len(b)) if p[j] == b[i]]
    assert list(zip(pr.tolist(), br.tolist())) == want  # probe order x insertion order


@needs_ref
def test_hash_join_vs_reference():
    rng = np.random.default_rng(4)
    b = rng.integers(-1000, 1000, 3000).astype(np.int64)
    p = rng.integers(-1200, 1200, 5000).astype(np.int64)
    pr, br = po.hash_join([po.Col(po.BIGINT, b)], [po.Col(po.BIGINT, p)])
    cap = len(pr) + 10
    rp = np.empty(cap, dtype=np.int32)
    rb = np.empty(cap, dtype=np.int32)
    k = ref.ref_hash_join_i64(len(b), b.ctypes.data, len(p), p.ctypes.data, cap, rp.ctypes.data, rb.ctypes.data)
    assert k == len(pr)
    assert np.array_equal(pr, rp[:k]) and np.array_equal(br, rb[:k])


# ------------------------------------------------------------- group-by
def test_group_by_integer_closed_form():
    rng = np.random.default_rng(6)
    keys = rng.integers(0, 500, 20000).astype(np.int64)
    vals = rng.integers(-2**40, 2**40, 20000).astype(np.int64)
    kv, _, v, _ = po.group_by([po.Col(po.BIGINT, keys)], [(po.SUM, po.Col(po.BIGINT, vals)), (po.COUNT, None)])
    # insertion order = first appearance
    _, first = np.unique(keys, return_index=True)
    order = keys[np.sort(first)]
    assert np.array_equal(kv[:, 0], order)
    sums = np.zeros(500, dtype=np.int64)
    np.add.at(sums, keys, vals)
    assert np.array_equal(v[:, 0], sums[order])
    assert np.array_equal(v[:, 1], np.bincount(keys, minlength=500)[order])


@needs_ref
def test_group_by_order_vs_reference():
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 40000, 100000).astype(np.int64)  # forces several resizes
    vals = rng.integers(-100, 100, 100000).astype(np.int64)
    kv, _, v, _ = po.group_by([po.Col(po.BIGINT, keys)], [(po.SUM, po.Col(po.BIGINT, vals)), (po.COUNT, None)],
                              max_groups=1 << 17)
    g = len(kv)
    rk = np.empty(g, dtype=np.int64)
    rs = np.empty(g, dtype=np.int64)
    rc = np.empty(g, dtype=np.int64)
    assert ref.ref_group_by_i64(len(keys), keys.ctypes.data, vals.ctypes.data, g, rk.ctypes.data, rs.ctypes.data,
                                rc.ctypes.data) == g
    assert np.array_equal(kv[:, 0], rk) and np.array_equal(v[:, 0], rs) and np.array_equal(v[:, 1], rc)


def test_sum_and_running_mean_left_fold():
    rng = np.random.default_rng(8)
    keys = rng.integers(0, 3, 3000).astype(np.int32)
    x = rng.uniform(0, 1e5, 3000)
    kv, _, v, _ = po.group_by([po.Col(po.INT, keys)], [(po.SUM, po.Col(po.REAL, x)), (po.AVG, po.Col(po.REAL, x))])
    for g in range(len(kv)):
        sel = x[keys == kv[g, 0]]
        s = 0.0
        first = True
        cma, cnt = 0.0, 1.0
        for val in sel:
            if first:
                s, cma, first = val, val, False
            else:
                s = s + val
                cnt = cnt + 1
                cma = cma + (val - cma) / cnt
        assert v[g, 0] == np.array([s]).view(np.int64)[0]
        assert v[g, 1] == np.array([cma]).view(np.int64)[0]
        assert abs(cma - math.fsum(sel) / len(sel)) <= 1e-12 * abs(cma)


def test_dtree_order_vs_fsum():
    """The GPU deterministic order is a valid summation: within n*eps of fsum."""
    n = 50000
    rng = np.random.default_rng(9)
    day = rng.integers(0, 100, n).astype(np.int32)
    g1 = rng.integers(0, 3, n).astype(np.int32)
    g2 = rng.integers(0, 2, n).astype(np.int32)
    m = [rng.uniform(0, 1e5, n) for _ in range(2)]
    ptrs = (ctypes.c_void_p * 2)(*[a.ctypes.data for a in m])
    out = np.zeros(12)
    po.lib().kor_dtree_scan_agg(n, day.ctypes.data, 97, g1.ctypes.data, g2.ctypes.data, 2, 6, 2, ptrs, 7, 256, 4,
                                out.ctypes.data)
    sel = day <= 97
    for g in range(6):
        for f in range(2):
            mask = sel & (g1 * 2 + g2 == g)
            exact = math.fsum(m[f][mask])
            assert abs(out[g * 2 + f] - exact) <= 1e-12 * exact
