"""Scan-select + low-cardinality group-by (the 60M-row workload's shape) on the
HIP path against the oracle.

Bars: counts and integer SUMs bit-exact; FP64 SUM/AVG within 1e-12 relative of
the oracle's sequential fold / running mean; the deterministic fixed order
bit-exact against the oracle's restatement of that order; group order =
insertion order."""
import ctypes

import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Column, Context
from kushdb_amd.workloads import scan_agg_table
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
REL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def run_gpu(ctx, t, n):
    cols = [Column(abi.KG_DATE, t["day"][:n]), Column(abi.KG_ENUM, t["g1"][:n]), Column(abi.KG_ENUM, t["g2"][:n])]
    cols += [Column(abi.KG_REAL, t[f"m{i}"][:n]) for i in range(1, 5)]
    cols += [Column(abi.KG_INT, t[f"i{i}"][:n]) for i in range(1, 4)]
    aggs = [(abi.KG_AGG_SUM, 3 + i) for i in range(7)] + [(abi.KG_AGG_AVG, 3 + i) for i in range(7)]
    aggs.append((abi.KG_AGG_COUNT, -1))
    return ctx.scan_group_agg(cols, [(0, abi.KG_LEQ, t["cutoff"])], [(1, 3), (2, 2)], aggs)


def host(t, n):
    return {k: (v[:n].cpu().numpy() if torch.is_tensor(v) else v) for k, v in t.items()}


def oracle_seq(h, n):
    fm = [np.ascontiguousarray(h[f"m{i}"]) for i in range(1, 5)]
    im = [np.ascontiguousarray(h[f"i{i}"]) for i in range(1, 4)]
    fp = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in fm])
    ip = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in im])
    w = 2 + 2 * 4 + 2 * 3 + 1
    out = np.zeros((16, w), dtype=np.int64)
    g = po.lib().kor_q_scan_agg(n, h["day"].ctypes.data, h["cutoff"], h["g1"].ctypes.data, h["g2"].ctypes.data, 4,
                                fp, 3, ip, 16, out.ctypes.data)
    return out[:g]


def oracle_dtree(ctx, h, n):
    fm = [np.ascontiguousarray(h[f"m{i}"]) for i in range(1, 5)]
    fp = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in fm])
    nb, vec = ctx.agg_order(n)
    out = np.zeros(6 * 4)
    po.lib().kor_dtree_scan_agg(n, h["day"].ctypes.data, h["cutoff"], h["g1"].ctypes.data, h["g2"].ctypes.data, 2,
                                6, 4, fp, nb, 256, vec, out.ctypes.data)
    return out.reshape(6, 4)


def f64(x):
    return np.asarray(x, dtype=np.int64).view(np.float64)


@pytest.mark.parametrize("n", [1, 1000, 257 * 256 + 3, 2_000_000])
def test_scan_agg_vs_oracle(ctx, n):
    print(f"scan_agg n={n}", flush=True)
    t = scan_agg_table(max(n, 1), device="cuda:0", seed=1234 + n)
    codes, vals, nulls = run_gpu(ctx, t, n)
    h = host(t, n)
    want = oracle_seq(h, n)
    assert len(codes) == len(want)
    # insertion order and keys
    assert np.array_equal(codes[:, 0], want[:, 0]) and np.array_equal(codes[:, 1], want[:, 1])
    assert not nulls.any()
    for g in range(len(want)):
        gv, wv = vals[g], want[g, 2:]
        # SUM(f64)
        assert np.allclose(f64(gv[0:4]), f64(wv[0:4]), rtol=REL, atol=0)
        # SUM(int32) wraps in int32: exact
        assert np.array_equal(gv[4:7], wv[8:11])
        # AVG(f64), AVG(int32): running mean vs sum/count
        assert np.allclose(f64(gv[7:11]), f64(wv[4:8]), rtol=REL, atol=0)
        assert np.allclose(f64(gv[11:14]), f64(wv[11:14]), rtol=REL, atol=0)
        # COUNT(*)
        assert gv[14] == wv[14]
    # Deterministic order: bit-exact against the oracle's restatement of it.
    # (These aligned, non-NULL INT/REAL columns take the 4-row vector path.)
    assert ctx.agg_order(n)[1] == 4
    dt = oracle_dtree(ctx, h, n)
    for g in range(len(want)):
        code = codes[g, 0] * 2 + codes[g, 1]
        assert np.array_equal(f64(vals[g, 0:4]), dt[code]), (g, f64(vals[g, 0:4]), dt[code])


def test_scan_agg_full_size_properties(ctx):
    """60M rows: counts and int sums against torch's exact int64 reductions,
    FP sums against torch float64 within 1e-12, two runs bit-identical."""
    n = 60_000_000
    print(f"scan_agg full n={n}", flush=True)
    t = scan_agg_table(n, device="cuda:0")
    codes, vals, _ = run_gpu(ctx, t, n)
    codes2, vals2, _ = run_gpu(ctx, t, n)
    assert np.array_equal(vals, vals2) and np.array_equal(codes, codes2)
    sel = t["day"] <= t["cutoff"]
    grp = t["g1"] * 2 + t["g2"]
    assert int(vals[:, 14].sum()) == int(sel.sum().item())
    for g in range(len(codes)):
        m = sel & (grp == int(codes[g, 0] * 2 + codes[g, 1]))
        assert vals[g, 14] == int(m.sum().item())
        for i in range(3):
            s = int(t[f"i{i + 1}"][m].to(torch.int64).sum().item())
            assert vals[g, 4 + i] == np.int64(s).astype(np.int32)  # wrapped like SUM(int32)
        for i in range(4):
            s = float(t[f"m{i + 1}"][m].sum().item())
            assert abs(f64(vals[g, i]) - s) <= REL * abs(s)


def test_scan_agg_empty_and_all_filtered(ctx):
    t = scan_agg_table(1000, device="cuda:0")
    t["cutoff"] = 0  # nothing qualifies
    codes, vals, _ = run_gpu(ctx, t, 1000)
    assert len(codes) == 0
    codes, vals, _ = run_gpu(ctx, t, 0)
    assert len(codes) == 0


def test_scan_select_vs_oracle(ctx):
    n = 3_000_001
    t = scan_agg_table(n, device="cuda:0", seed=77)
    lo, hi = t["cutoff"] - 900, t["cutoff"] - 100
    cols = [Column(abi.KG_DATE, t["day"]), Column(abi.KG_ENUM, t["g1"])]
    preds = [(0, abi.KG_GEQ, lo), (0, abi.KG_LT, hi), (1, abi.KG_NEQ, 1)]
    got = ctx.scan_select(cols, preds).cpu().numpy()
    h = host(t, n)
    want = po.scan_select([po.Col(po.DATE, h["day"]), po.Col(po.ENUM, h["g1"])], preds)
    assert np.array_equal(got, want)
    # disjunctive (SimdScanSelect with conjunction = false)
    got = ctx.scan_select(cols, preds, conjunction=False).cpu().numpy()
    want = po.scan_select([po.Col(po.DATE, h["day"]), po.Col(po.ENUM, h["g1"])], preds, conjunction=False)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [310, 5000, 700_003])
@pytest.mark.parametrize("ngroups", [1, 3, 12])
def test_minmax_extremes_anywhere(ctx, n, ngroups):
    """MIN / MAX over INT, BIGINT and REAL columns with NULLs, grouped by a
    dense code (1 group: ungrouped; 3: the register path; 12: the LDS path),
    with each group's extremes planted at positions that fall in every
    grid-stride step of a thread (the first row, row 256, the last rows, and
    random ones), against numpy per group in insertion order."""
    rng = np.random.default_rng(n * 31 + ngroups)
    g = rng.integers(0, ngroups, n).astype(np.int32)
    i32 = rng.integers(-1000, 1000, n).astype(np.int32)
    i64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    f = rng.uniform(-1e6, 1e6, n)
    nul = (rng.random(n) < 0.05).astype(np.int8)
    plant = sorted({0, min(256, n - 1), min(257, n - 1), n - 1, n - 2, *rng.integers(0, n, 8).tolist()})
    for k, p in enumerate(plant):
        sign = 1 if k % 2 else -1
        i32[p] = sign * (5000 + k)
        i64[p] = sign * ((1 << 50) + k)
        f[p] = sign * (1e9 + k)
        nul[p] = 0
    dev = "cuda:0"
    cols = [Column(abi.KG_ENUM, torch.from_numpy(g).to(dev)),
            Column(abi.KG_INT, torch.from_numpy(i32).to(dev), torch.from_numpy(nul).to(dev)),
            Column(abi.KG_BIGINT, torch.from_numpy(i64).to(dev)),
            Column(abi.KG_REAL, torch.from_numpy(f).to(dev), torch.from_numpy(nul).to(dev))]
    aggs = [(abi.KG_AGG_MIN, 1), (abi.KG_AGG_MAX, 1), (abi.KG_AGG_MIN, 2), (abi.KG_AGG_MAX, 2),
            (abi.KG_AGG_MIN, 3), (abi.KG_AGG_MAX, 3)]
    print("minmax", n, ngroups, flush=True)
    keys, vals, nulls = ctx.scan_group_agg(cols, [], [(0, ngroups)], aggs)
    order = [int(x) for x in dict.fromkeys(g.tolist())]  # insertion order
    assert [int(k) for k in np.asarray(keys).reshape(-1)] == order
    ok = nul == 0
    for row, code in enumerate(order):
        m = g == code
        want = [i32[m & ok].min(), i32[m & ok].max(), i64[m].min(), i64[m].max()]
        assert [int(x) for x in vals[row, :4]] == [int(x) for x in want], (n, ngroups, code)
        assert vals[row, 4:].view(np.float64).tolist() == [f[m & ok].min(), f[m & ok].max()], (n, ngroups, code)
        assert not nulls[row].any()
    # The hashed GROUP BY over the same key column.
    first, hvals, hnulls = ctx.hash_group_agg(cols, [], [0], aggs)
    hk = g[np.asarray(first)]
    for row, code in enumerate(hk.tolist()):
        m = g == code
        want = [i32[m & ok].min(), i32[m & ok].max(), i64[m].min(), i64[m].max()]
        assert [int(x) for x in hvals[row, :4]] == [int(x) for x in want], ("hashed", n, ngroups, code)
        assert hvals[row, 4:].view(np.float64).tolist() == [f[m & ok].min(), f[m & ok].max()], ("hashed", n, code)
