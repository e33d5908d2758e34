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
