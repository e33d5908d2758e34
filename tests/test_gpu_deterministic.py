"""The deterministic FP64 mode (kg_ctx_set_deterministic(1)) on every path
that otherwise adds REAL values through atomics: the dense LDS table (more than
8 groups), hashed GROUP BY, and the star pipeline (join output rows, build-side
duplicates included). Each REAL SUM must equal, bit for bit, the oracle's
restatement of the fixed order (po.det_sum over the group's values in input
order, NULL as -0.0), AVG must be that sum divided by the non-NULL count, and
two runs must agree bit for bit. Batches do not take REAL sums in this mode.
Reference order being replaced: the sequential fold of
compile/proxy/aggregator.cc:67-71 (AVG :239-268), which these sums match
within 1e-12 relative."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Batch, Column, Context, HashTable, Query
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def ctx():
    c = Context(0, deterministic=True)
    yield c
    c.close()


def t(a):
    return torch.as_tensor(a).to(DEV)


def expected_sums(groups_in_order, gid, vals, nulls):
    """{group: (det sum, non-NULL count)} over items in input order."""
    out = {}
    v = np.where(nulls.astype(bool), -0.0, vals)
    for g in groups_in_order:
        sel = gid == g
        out[g] = (po.det_sum(v[sel]), int((~nulls.astype(bool))[sel].sum()))
    return out


def check_group(got_sum_bits, got_avg_bits, got_nulls, want):
    s, nn = want
    if nn == 0:
        assert got_nulls[0] and got_nulls[1]
        return
    assert got_sum_bits == np.float64(s).view(np.int64), (np.int64(got_sum_bits).view(np.float64), s)
    assert got_avg_bits == np.float64(s / nn).view(np.int64)


def test_dense_lds_path(ctx):
    """50 dense groups (LDS table), REAL measure with NULLs and a REAL product
    measure; groups of ~10k rows (two chunk levels)."""
    rng = np.random.default_rng(1)
    n = 600_000
    key = rng.integers(0, 50, n).astype(np.int32)
    day = rng.integers(0, 100, n).astype(np.int32)
    m = rng.uniform(-1e5, 1e5, n)
    mn = (rng.random(n) < 0.1).astype(np.int8)
    price = rng.uniform(1, 1e5, n)
    qty = rng.integers(1, 51, n).astype(np.int32)
    cols = [Column(abi.KG_INT, t(key)), Column(abi.KG_DATE, t(day)), Column(abi.KG_REAL, t(m), t(mn)),
            Column(abi.KG_REAL, t(price)), Column(abi.KG_INT, t(qty))]
    aggs = [(abi.KG_AGG_SUM, 2), (abi.KG_AGG_AVG, 2), (abi.KG_AGG_SUM, 3, 4), (abi.KG_AGG_AVG, 3, 4),
            (abi.KG_AGG_COUNT, -1)]
    runs = [ctx.scan_group_agg(cols, [(1, abi.KG_LT, 90)], [(0, 50)], aggs, max_groups=64) for _ in range(2)]
    (codes, vals, nulls), (c2, v2, n2) = runs
    assert np.array_equal(codes, c2) and np.array_equal(vals, v2) and np.array_equal(nulls, n2)
    q = day < 90
    rows = np.nonzero(q)[0]
    first = {}
    for r in rows:
        first.setdefault(key[r], r)
        if len(first) == 50:
            break
    order = sorted(first, key=first.get)
    assert codes[:, 0].tolist() == order
    w1 = expected_sums(order, key[q], m[q], mn[q])
    w2 = expected_sums(order, key[q], price[q] * qty[q].astype(np.float64), np.zeros(q.sum(), np.int8))
    for i, g in enumerate(order):
        check_group(vals[i, 0], vals[i, 1], nulls[i, :2], w1[g])
        check_group(vals[i, 2], vals[i, 3], nulls[i, 2:4], w2[g])
        assert vals[i, 4] == int((key[q] == g).sum())


@pytest.mark.parametrize("ngroups,n", [(5000, 300_000), (2, 200_000)])
def test_hashed_path(ctx, ngroups, n):
    """BIGINT keys through kg_hash_group_agg; with 2 groups of 100k rows the
    sums take three chunk levels."""
    rng = np.random.default_rng(ngroups)
    domain = rng.integers(-2**62, 2**62, ngroups)
    key = domain[rng.integers(0, ngroups, n)]
    m = rng.normal(0, 1e6, n)
    mn = (rng.random(n) < 0.05).astype(np.int8)
    cols = [Column(abi.KG_BIGINT, t(key)), Column(abi.KG_REAL, t(m), t(mn))]
    aggs = [(abi.KG_AGG_SUM, 1), (abi.KG_AGG_AVG, 1), (abi.KG_AGG_COUNT, -1)]
    (f1, v1, n1), (f2, v2, n2) = [ctx.hash_group_agg(cols, [], [0], aggs, max_groups=ngroups) for _ in range(2)]
    assert np.array_equal(f1, f2) and np.array_equal(v1, v2) and np.array_equal(n1, n2)
    _, first_rows = np.unique(key, return_index=True)
    assert np.array_equal(np.asarray(f1), np.sort(first_rows))
    order = [key[r] for r in f1]
    want = expected_sums(order, key, m, mn)
    for i, g in enumerate(order):
        check_group(v1[i, 0], v1[i, 1], n1[i, :2], want[g])
        assert v1[i, 2] == int((key == g).sum())


def test_star_path_with_duplicates(ctx):
    """Star pipeline with duplicate build keys (a probe row feeds several join
    rows, possibly in several groups) and misses."""
    rng = np.random.default_rng(3)
    nb, nf = 2000, 400_000
    pk = rng.integers(0, 600, nb).astype(np.int32)
    cat = rng.integers(0, 25, nb).astype(np.int32)
    fk = rng.integers(-10, 700, nf).astype(np.int32)
    day = rng.integers(0, 10, nf).astype(np.int32)
    price = rng.uniform(1, 1e5, nf)
    qty = rng.integers(1, 51, nf).astype(np.int32)
    runs = []
    for _ in range(2):
        ht = HashTable(ctx, nb)
        ht.insert(Column(abi.KG_INT, t(pk)))
        cols = [Column(abi.KG_DATE, t(day)), Column(abi.KG_INT, t(fk)), Column(abi.KG_REAL, t(price)),
                Column(abi.KG_INT, t(qty))]
        runs.append(ht.scan_probe_group_agg(cols, [(0, abi.KG_GEQ, 2), (0, abi.KG_LT, 8)], 1,
                                            Column(abi.KG_ENUM, t(cat)), 25,
                                            [(abi.KG_AGG_SUM, 2, 3), (abi.KG_AGG_AVG, 2, 3),
                                             (abi.KG_AGG_COUNT, -1)]))
        ht.free()
    (codes, vals, nulls), (c2, v2, n2) = runs
    assert np.array_equal(codes, c2) and np.array_equal(vals, v2) and np.array_equal(nulls, n2)
    # Join output rows in probe order: (group, value) per match.
    by_pk = {}
    for b in range(nb):
        by_pk.setdefault(int(pk[b]), []).append(int(cat[b]))
    gid, val = [], []
    for r in np.nonzero((day >= 2) & (day < 8))[0]:
        for g in by_pk.get(int(fk[r]), ()):
            gid.append(g)
            val.append(price[r] * float(qty[r]))
    gid, val = np.array(gid), np.array(val)
    order = list(dict.fromkeys(gid.tolist()))
    assert np.asarray(codes).reshape(-1).tolist() == order
    want = expected_sums(order, gid, val, np.zeros(len(val), np.int8))
    for i, g in enumerate(order):
        check_group(vals[i, 0], vals[i, 1], nulls[i, :2], want[g])
        assert vals[i, 2] == int((gid == g).sum())


def test_batch_rejects_real_sums(ctx):
    """A batch does not take a REAL sum in the deterministic mode (reported per
    query); an integer-only query still runs."""
    rng = np.random.default_rng(4)
    n = 10_000
    cols = [Column(abi.KG_INT, t(rng.integers(0, 4, n).astype(np.int32))),
            Column(abi.KG_REAL, t(rng.uniform(0, 1, n))), Column(abi.KG_INT, t(rng.integers(0, 9, n).astype(np.int32)))]
    qs = [Query.scan_agg(cols, [], [(0, 4)], [(abi.KG_AGG_SUM, 1)]),
          Query.scan_agg(cols, [], [(0, 4)], [(abi.KG_AGG_SUM, 2), (abi.KG_AGG_COUNT, -1)])]
    b = Batch(ctx, arena_bytes=64 << 20)
    res = b.run(qs, max_groups=8)
    b.free()
    assert res[0][0] == abi.KG_ERR_UNSUPPORTED
    assert res[1][0] == abi.KG_OK
    codes, vals, _ = ctx.scan_group_agg(cols, [], [(0, 4)], [(abi.KG_AGG_SUM, 2), (abi.KG_AGG_COUNT, -1)])
    order = np.argsort(res[1][1][:, 0])
    assert np.array_equal(res[1][2][order], vals[np.argsort(codes[:, 0])])
