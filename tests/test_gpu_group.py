"""GROUP BY over hashed keys (kg_hash_group_agg) against the oracle's
restatement of the reference's aggregate hash table (salted directory,
insertion-order ForEach, the aggregators' type rules): multi-column keys with
NULL key values, REAL keys with -0.0 / 0.0, SUM / AVG / MIN / MAX / COUNT(*) /
COUNT(col) over INT, SMALLINT, BIGINT and REAL measures with NULLs, a
predicate, and group counts from a handful to 200k.

Bars: group order (first qualifying row) and key values exact; integer
results, COUNTs and MIN/MAX exact; FP64 SUM/AVG within 1e-12 of |result| +
the inputs' magnitude (atomic addition order; the reference's running-mean
AVG rounds relative to its inputs)."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Column, Context
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
REL = 1e-12

KEY_TYPES = {"int": (abi.KG_INT, np.int32), "bigint": (abi.KG_BIGINT, np.int64), "date": (abi.KG_DATE, np.int32),
             "real": (abi.KG_REAL, np.float64), "smallint": (abi.KG_SMALLINT, np.int16),
             "boolean": (abi.KG_BOOLEAN, np.int8)}


def make_keys(rng, typ, n, card):
    if typ == "real":
        v = rng.integers(0, card, n).astype(np.float64) - card // 2
        v[v == 0] = np.where(rng.random(int((v == 0).sum())) < 0.5, -0.0, 0.0)
        return v
    if typ == "boolean":
        return rng.integers(0, 2, n).astype(np.int8)
    return (rng.integers(0, card, n) - card // 3).astype(KEY_TYPES[typ][1])


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n,card,keys,seed", [
    (1, 3, ("int",), 0), (5000, 7, ("int", "bigint"), 1), (200_000, 600, ("date", "real"), 2),
    (1_000_000, 200_000, ("bigint",), 3), (300_000, 40, ("smallint", "boolean", "int"), 4),
    (400_000, 5000, ("real", "int", "date", "bigint"), 5)])
def test_hash_group_vs_oracle(ctx, n, card, keys, seed):
    print(f"hash group n={n} card={card} keys={keys}", flush=True)
    rng = np.random.default_rng(seed)
    hk, dk = [], []
    for i, t in enumerate(keys):
        v = make_keys(rng, t, n, card)
        nl = (rng.random(n) < 0.03).astype(np.int8) if i == 0 else None
        hk.append(po.Col(KEY_TYPES[t][0], v, nl))
        dk.append(Column(KEY_TYPES[t][0], torch.from_numpy(v).to(DEV),
                         torch.from_numpy(nl).to(DEV) if nl is not None else None))
    m_int = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    m_small = rng.integers(-2**15, 2**15 - 1, n).astype(np.int16)
    m_big = rng.integers(-2**40, 2**40, n).astype(np.int64)
    m_real = rng.random(n) * 1e5 - 5e4
    nl_real = (rng.random(n) < 0.1).astype(np.int8)
    flt = rng.integers(0, 100, n).astype(np.int32)
    measures = [(abi.KG_INT, m_int, None), (abi.KG_SMALLINT, m_small, None), (abi.KG_BIGINT, m_big, None),
                (abi.KG_REAL, m_real, nl_real)]
    dcols = dk + [Column(t, torch.from_numpy(v).to(DEV), torch.from_numpy(nl).to(DEV) if nl is not None else None)
                  for t, v, nl in measures] + [Column(abi.KG_INT, torch.from_numpy(flt).to(DEV))]
    nk = len(keys)
    mi, ms, mb, mr, fc = nk, nk + 1, nk + 2, nk + 3, nk + 4
    aggs = [(abi.KG_AGG_SUM, mi), (abi.KG_AGG_SUM, ms), (abi.KG_AGG_SUM, mb), (abi.KG_AGG_SUM, mr),
            (abi.KG_AGG_AVG, mr), (abi.KG_AGG_AVG, mi), (abi.KG_AGG_MIN, mr), (abi.KG_AGG_MAX, mb),
            (abi.KG_AGG_MIN, mi), (abi.KG_AGG_COUNT, -1), (abi.KG_AGG_COUNT, mr)]
    preds = [(fc, abi.KG_LT, 90)]
    first, vals, nulls = ctx.hash_group_agg(dcols, preds, list(range(nk)), aggs, max_groups=n + 1)

    rows = po.scan_select([po.Col(po.INT, flt)], [(0, po.LT, 90)])
    hm = {mi: po.Col(po.INT, m_int), ms: po.Col(po.SMALLINT, m_small), mb: po.Col(po.BIGINT, m_big),
          mr: po.Col(po.REAL, m_real, nl_real)}
    okeys, oknull, ovals, onulls = po.group_by(hk, [(k, hm[c] if c >= 0 else None) for k, c in aggs], rows=rows,
                                               max_groups=n + 1)
    assert len(first) == len(okeys)
    # group order and key values: the keys at each group's first row
    for j, c in enumerate(hk):
        got = c.data[first]
        gotn = c.nulls[first] if c.nulls is not None else np.zeros(len(first), np.int8)
        assert np.array_equal(gotn, oknull[:, j])
        want = okeys[:, j]
        if c.type == po.REAL:
            want = want.view(np.float64)
        m = gotn == 0
        assert np.array_equal(np.asarray(got, dtype=np.float64 if c.type == po.REAL else np.int64)[m], want[m])
    assert np.array_equal(nulls, onulls)
    for a, (kind, col) in enumerate(aggs):
        g, w = vals[:, a], ovals[:, a]
        real = col == mr and kind in (abi.KG_AGG_SUM, abi.KG_AGG_AVG, abi.KG_AGG_MIN) or kind == abi.KG_AGG_AVG
        if real and kind != abi.KG_AGG_MIN:
            # The reference's AVG is a running mean (cma + (x - cma) / n), rounded
            # relative to the inputs' magnitude, and FP64 SUMs add in atomic order:
            # both within 1e-12 of |result| + the inputs' scale.
            scale = float(np.abs(m_int).max()) if col == mi else float(np.abs(m_real).max())
            gf, wf = g.view(np.float64), w.view(np.float64)
            ok = onulls[:, a] == 0
            assert np.all(np.abs(gf[ok] - wf[ok]) <= REL * (np.abs(wf[ok]) + scale)), (a, kind)
        else:
            assert np.array_equal(g, w), (a, kind)


def test_hash_group_capacity_and_empty(ctx):
    k = Column(abi.KG_INT, torch.arange(1000, dtype=torch.int32, device=DEV))
    with pytest.raises(abi.KgError):
        ctx.hash_group_agg([k], [], [0], [(abi.KG_AGG_COUNT, -1)], max_groups=10)
    first, vals, _ = ctx.hash_group_agg([k], [(0, abi.KG_LT, -5)], [0], [(abi.KG_AGG_COUNT, -1)], max_groups=10)
    assert len(first) == 0


def test_hash_group_many_more_keys_than_max_groups(ctx):
    """16M distinct keys against max_groups = 1000: the table holds 2048 slots,
    and the call must end with KG_ERR_CAPACITY after about that many claims,
    not probe the full table for each of the remaining rows."""
    import time
    k = Column(abi.KG_BIGINT, torch.randperm(1 << 24, dtype=torch.int64, device=DEV))
    t0 = time.perf_counter()
    with pytest.raises(abi.KgError) as ei:
        ctx.hash_group_agg([k], [], [0], [(abi.KG_AGG_COUNT, -1)], max_groups=1000)
    assert ei.value.status == abi.KG_ERR_CAPACITY
    assert time.perf_counter() - t0 < 5.0
    first, vals, _ = ctx.hash_group_agg([k], [(0, abi.KG_LT, 1000)], [0], [(abi.KG_AGG_COUNT, -1)], max_groups=1000)
    assert len(first) == 1000 and np.all(vals[:, 0] == 1)


def test_dense_keys_beyond_on_chip_table(ctx):
    """kg_scan_group_agg over 3000 dense codes x 2 keys (6000 groups, more than
    an LDS table holds) takes the hashed path: the same groups in insertion
    order, codes and values as the oracle."""
    rng = np.random.default_rng(9)
    n = 500_000
    a = rng.integers(0, 3000, n).astype(np.int32)
    b = rng.integers(0, 2, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    cols = [Column(abi.KG_INT, torch.from_numpy(x).to(DEV)) for x in (a, b, v)]
    codes, vals, nulls = ctx.scan_group_agg(cols, [(2, abi.KG_GT, -900)], [(0, 3000), (1, 2)],
                                            [(abi.KG_AGG_SUM, 2), (abi.KG_AGG_MIN, 2), (abi.KG_AGG_COUNT, -1)],
                                            max_groups=6000)
    rows = po.scan_select([po.Col(po.INT, v)], [(0, po.GT, -900)])
    kv, _, ov, onl = po.group_by([po.Col(po.INT, a), po.Col(po.INT, b)],
                                 [(po.SUM, po.Col(po.INT, v)), (po.MIN, po.Col(po.INT, v)), (po.COUNT, None)],
                                 rows=rows, max_groups=6001)
    assert np.array_equal(codes.astype(np.int64), kv) and np.array_equal(vals, ov) and np.array_equal(nulls, onl)


def test_dense_1024_groups_7_measures_stay_in_lds(ctx):
    """1024 dense groups x 7 SUM measures need a 131 KB group table: it stays
    on the LDS path (gfx950 gives a workgroup up to 160 KB) and equals the
    oracle (integer SUMs exact, REAL SUMs within 1e-12)."""
    rng = np.random.default_rng(21)
    n = 2_000_000
    g = rng.integers(0, 1024, n).astype(np.int32)
    ints = [rng.integers(-1000, 1000, n).astype(np.int32) for _ in range(4)]
    reals = [rng.uniform(-1e5, 1e5, n) for _ in range(3)]
    cols = [Column(abi.KG_INT, torch.from_numpy(g).to(DEV))]
    cols += [Column(abi.KG_INT, torch.from_numpy(x).to(DEV)) for x in ints]
    cols += [Column(abi.KG_REAL, torch.from_numpy(x).to(DEV)) for x in reals]
    aggs = [(abi.KG_AGG_SUM, 1 + i) for i in range(7)] + [(abi.KG_AGG_COUNT, -1)]
    codes, vals, nulls = ctx.scan_group_agg(cols, [], [(0, 1024)], aggs, max_groups=1024)
    assert ctx.last_path() == "lds"
    ocols = [po.Col(po.INT, x) for x in ints] + [po.Col(po.REAL, x) for x in reals]
    kv, _, ov, _ = po.group_by([po.Col(po.INT, g)], [(po.SUM, c) for c in ocols] + [(po.COUNT, None)],
                               max_groups=1025)
    assert np.array_equal(codes[:, 0].astype(np.int64), kv[:, 0] if kv.ndim > 1 else kv)
    assert np.array_equal(vals[:, :4], ov[:, :4]) and np.array_equal(vals[:, 7], ov[:, 7])
    assert np.allclose(vals[:, 4:7].view(np.float64), ov[:, 4:7].view(np.float64), rtol=1e-12, atol=1e-6)
