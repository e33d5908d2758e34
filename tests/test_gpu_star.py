"""The star query (1M-row fact, 10k-row dim): ScanSelect(30% day range) ->
HashJoin(fact.fk = dim.pk) -> GroupBy(cat; SUM(price * Float(qty)), COUNT(*))
-> OrderBy(cat by decoded string), on the HIP path against the oracle.

Bars: group keys, order and COUNT exact; FP64 SUM within 1e-12 relative."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Column, Context, HashTable
from kushdb_amd.workloads import star_tables
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
REL = 1e-12
# A dictionary whose string order differs from code order, so ORDER BY on the
# ENUM (which compares decoded strings, compile/proxy/evaluate.cc:390-393) is
# visible.
CAT_NAMES = [f"{chr(ord('Z') - (7 * i) % 26)}-cat{i:02d}" for i in range(25)]


def run_star(ctx, f, d, lo, hi):
    ht = HashTable(ctx, len(d["pk"]))
    ht.insert(Column(abi.KG_INT, d["pk"]))
    cols = [Column(abi.KG_DATE, f["day"]), Column(abi.KG_INT, f["fk"]), Column(abi.KG_REAL, f["price"]),
            Column(abi.KG_INT, f["qty"])]
    preds = [(0, abi.KG_GEQ, lo), (0, abi.KG_LT, hi)]
    codes, vals, nulls = ht.scan_probe_group_agg(cols, preds, 1, Column(abi.KG_ENUM, d["cat"]), 25,
                                                 [(abi.KG_AGG_SUM, 2, 3), (abi.KG_AGG_COUNT, -1)])
    ht.free()
    return codes, vals, nulls


def oracle_star(f, d, lo, hi):
    h = {k: v.cpu().numpy() for k, v in f.items()}
    hd = {k: v.cpu().numpy() for k, v in d.items()}
    oc = np.zeros(64, dtype=np.int32)
    osum = np.zeros(64)
    ocnt = np.zeros(64, dtype=np.int64)
    g = po.lib().kor_q_star(len(h["fk"]), h["fk"].ctypes.data, h["qty"].ctypes.data, h["price"].ctypes.data,
                            h["day"].ctypes.data, lo, hi, len(hd["pk"]), hd["pk"].ctypes.data,
                            hd["cat"].ctypes.data, 64, oc.ctypes.data, osum.ctypes.data, ocnt.ctypes.data)
    return oc[:g], osum[:g], ocnt[:g]


@pytest.mark.parametrize("nfact", [1, 5000, 1_000_000])
def test_star_vs_oracle(nfact):
    ctx = Context(0)
    f, d, (lo, hi) = star_tables(nfact=nfact, device="cuda:0", seed=31 + nfact)
    codes, vals, nulls = run_star(ctx, f, d, lo, hi)
    oc, osum, ocnt = oracle_star(f, d, lo, hi)
    assert np.array_equal(codes, oc)  # insertion order
    assert np.array_equal(vals[:, 1], ocnt)
    assert np.allclose(vals[:, 0].view(np.float64), osum, rtol=REL, atol=0)
    assert not nulls.any()
    # ORDER BY cat: decoded strings
    got = sorted(zip([CAT_NAMES[c] for c in codes], vals[:, 0].view(np.float64), vals[:, 1]))
    want = sorted(zip([CAT_NAMES[c] for c in oc], osum, ocnt))
    assert [g[0] for g in got] == [w[0] for w in want]


def test_star_join_duplicates_and_misses():
    """Build keys with duplicates and probe keys that miss: every match counts."""
    ctx = Context(0)
    rng = np.random.default_rng(3)
    pk = rng.integers(0, 300, 1000).astype(np.int32)
    cat = rng.integers(0, 25, 1000).astype(np.int32)
    fk = rng.integers(-10, 400, 20000).astype(np.int32)
    day = rng.integers(0, 10, 20000).astype(np.int32)
    price = rng.uniform(1, 1e5, 20000)
    qty = rng.integers(1, 51, 20000).astype(np.int32)
    t = lambda a: torch.as_tensor(a).to("cuda:0")
    f = {"fk": t(fk), "qty": t(qty), "price": t(price), "day": t(day)}
    d = {"pk": t(pk), "cat": t(cat)}
    codes, vals, _ = run_star(ctx, f, d, 2, 8)
    oc, osum, ocnt = oracle_star(f, d, 2, 8)
    assert np.array_equal(codes, oc)
    assert np.array_equal(vals[:, 1], ocnt)
    assert np.allclose(vals[:, 0].view(np.float64), osum, rtol=REL, atol=0)


def test_group_order_at_scale():
    """A 12M-row fact with a measure (the star kernel's LDS table holds the
    sums ahead of the first-row words, so the first-row slots start in a later
    wave's range than the threads that set them): group order and every value
    against numpy, for the star path and for the dense LDS GROUP BY
    (50 groups, SUM, MIN and MAX). Regression for a race between the zeroing
    and the first-row / MIN / MAX initialisation of the LDS table."""
    ctx = Context(0)
    rng = np.random.default_rng(8)
    n = 12_000_000
    fk = rng.integers(0, 10_000, n).astype(np.int32)
    day = rng.integers(0, 2557, n).astype(np.int32)
    qty = rng.integers(1, 51, n).astype(np.int32)
    cat = rng.integers(0, 25, 10_000).astype(np.int32)
    dv = lambda x: torch.as_tensor(x).to("cuda:0")
    cols = [Column(abi.KG_INT, dv(fk)), Column(abi.KG_INT, dv(qty)), Column(abi.KG_DATE, dv(day))]
    ht = HashTable(ctx, 10_000)
    ht.insert(Column(abi.KG_INT, dv(np.arange(10_000, dtype=np.int32))))
    codes, vals, _ = ht.scan_probe_group_agg(cols, [(2, abi.KG_GEQ, 100), (2, abi.KG_LT, 867)], 0,
                                             Column(abi.KG_ENUM, dv(cat)), 25,
                                             [(abi.KG_AGG_SUM, 1), (abi.KG_AGG_COUNT, -1)])
    ht.free()
    rows = np.nonzero((day >= 100) & (day < 867))[0]
    g = cat[fk[rows]]
    u, first = np.unique(g, return_index=True)
    order = u[np.argsort(first)]
    assert np.asarray(codes).reshape(-1).tolist() == order.tolist()
    for i, c in enumerate(order):
        assert vals[i, 0] == qty[rows][g == c].astype(np.int64).sum() and vals[i, 1] == (g == c).sum()
    # dense LDS GROUP BY over 50 groups with MIN / MAX
    key = rng.integers(0, 50, n).astype(np.int32)
    kc = [Column(abi.KG_INT, dv(key)), Column(abi.KG_INT, dv(qty)), Column(abi.KG_DATE, dv(day))]
    codes, vals, _ = ctx.scan_group_agg(kc, [(2, abi.KG_GEQ, 100)], [(0, 50)],
                                        [(abi.KG_AGG_SUM, 1), (abi.KG_AGG_MIN, 1), (abi.KG_AGG_MAX, 1),
                                         (abi.KG_AGG_COUNT, -1)], max_groups=64)
    assert ctx.last_path() == "lds"
    sel = np.nonzero(day >= 100)[0]
    kk = key[sel]
    u, first = np.unique(kk, return_index=True)
    order = u[np.argsort(first)]
    assert codes[:, 0].tolist() == order.tolist()
    for i, c in enumerate(order):
        q = qty[sel][kk == c].astype(np.int64)
        assert vals[i].tolist() == [q.sum(), q.min(), q.max(), len(q)]
    ctx.close()
