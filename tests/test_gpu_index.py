"""Column index build + batched lookups, batched sorted intersection and the
tuple-id set on the HIP path, against the oracle. All integer: bit-exact."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Column, ColumnIndex, Context, TupleSet, intersect_batch
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def t(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV)


def oracle_index(keys, rows=None):
    n = len(rows) if rows is not None else len(keys)
    ok = np.zeros(n + 1, dtype=np.int32)
    oo = np.zeros(n + 2, dtype=np.int64)
    ot = np.zeros(n + 1, dtype=np.int32)
    rr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nk = po.lib().kor_memory_index_build(n, keys.ctypes.data, rr.ctypes.data if rr is not None else None, n + 1,
                                         ok.ctypes.data, oo.ctypes.data, ot.ctypes.data)
    return {int(ok[i]): ot[oo[i]:oo[i + 1]] for i in range(nk)}


# Dense directories (keys spanning at most about 3x their count), hashed ones
# (sparse keys, the extremes of INT), and keys that arrive sorted (no sort).
@pytest.mark.parametrize("n,domain,with_nulls,with_sel,order", [
    (1, 5, False, False, "random"), (100_000, 1000, False, False, "random"),
    (300_000, 50_000, True, True, "random"), (2_000_000, 10, False, True, "random"),
    (200_000, 1 << 30, False, False, "random"), (200_000, 1 << 30, True, True, "random"),
    (50_000, 5, False, False, "extremes"), (400_000, 300_000, False, True, "sorted"),
    (400_000, 300_000, True, False, "sorted"), (300_000, 1 << 28, False, False, "sorted")])
def test_index_build_and_lookup(ctx, n, domain, with_nulls, with_sel, order):
    rng = np.random.default_rng(n + domain)
    keys = rng.integers(-domain, domain, n).astype(np.int32)
    if order == "sorted":
        keys = np.sort(keys)
    elif order == "extremes":
        keys = rng.choice(np.array([-2**31, 2**31 - 1, 0, 7], dtype=np.int64), n).astype(np.int32)
    nulls = (rng.random(n) < 0.05).astype(np.int8) if with_nulls else None
    sel = np.sort(rng.choice(n, size=max(1, n // 3), replace=False)).astype(np.int32) if with_sel else None
    col = Column(abi.KG_INT, t(keys), t(nulls) if nulls is not None else None)
    idx = ColumnIndex(ctx, col, t(sel) if sel is not None else None)
    # Oracle: tuple id = selection position; NULL keys are not indexed.
    pos = np.arange(len(sel) if sel is not None else n, dtype=np.int32)
    src = sel if sel is not None else pos
    keep = pos if nulls is None else pos[nulls[src] == 0]
    want = oracle_index(keys[src], keep)
    assert idx.nkeys == len(want)
    assert idx.ntids == sum(len(v) for v in want.values())
    probe = np.concatenate([np.array(list(want.keys()), dtype=np.int32),
                            rng.integers(-2 * domain - 5, 2 * domain + 5, 1000).astype(np.int64).clip(-2**31, 2**31 - 1)
                            .astype(np.int32), np.array([-2**31, 2**31 - 1], dtype=np.int32)])
    ptrs, cnts = idx.lookup(Column(abi.KG_INT, t(probe)))
    offs = idx.offsets_of(ptrs).cpu().numpy()
    cnts = cnts.cpu().numpy()
    tids = idx.tids().cpu().numpy()
    for j, k in enumerate(probe):
        got = tids[offs[j]:offs[j] + cnts[j]]
        exp = want.get(int(k), np.zeros(0, dtype=np.int32))
        assert np.array_equal(got, exp), (k, got[:5], exp[:5])
    idx.free()


@pytest.mark.parametrize("kind", ["bigint_sparse", "bigint_dense", "bigint_extremes", "real"])
def test_index_64bit_keys(ctx, kind):
    """BIGINT and REAL keys: runs grouped by value (REAL compares as double:
    -0.0 and 0.0 are one key, NaN is never indexed nor found), tuple ids
    ascending per key, NULLs skipped; the hashed directory with INT64_MIN as a
    key, and the dense directory for a small BIGINT span far from zero."""
    rng = np.random.default_rng(len(kind))
    n = 300_000
    if kind == "bigint_sparse":
        keys = rng.integers(-2**63, 2**63 - 1, 20_000, dtype=np.int64)[rng.integers(0, 20_000, n)]
        typ = abi.KG_BIGINT
    elif kind == "bigint_dense":
        keys = rng.integers(0, 50_000, n).astype(np.int64) + 7 * 10**15
        typ = abi.KG_BIGINT
    elif kind == "bigint_extremes":
        keys = rng.choice(np.array([-2**63, 2**63 - 1, 0, -1, 5], dtype=np.int64), n)
        typ = abi.KG_BIGINT
    else:
        keys = rng.integers(-500, 500, n).astype(np.float64) / 4
        keys[rng.random(n) < 0.02] = -0.0
        keys[rng.random(n) < 0.02] = np.nan
        keys[rng.random(n) < 0.01] = np.inf
        typ = abi.KG_REAL
    nulls = (rng.random(n) < 0.05).astype(np.int8)
    col = Column(typ, t(keys), t(nulls))
    idx = ColumnIndex(ctx, col)
    want = {}
    for i in range(n):
        v = keys[i]
        if nulls[i] or (typ == abi.KG_REAL and np.isnan(v)):
            continue
        kk = float(v) + 0.0 if typ == abi.KG_REAL else int(v)   # + 0.0 folds -0.0
        want.setdefault(kk, []).append(i)
    assert idx.nkeys == len(want) and idx.ntids == sum(len(v) for v in want.values())
    probe = list(want.keys())
    if typ == abi.KG_REAL:
        probe += [-0.0, np.nan, 1e300, -np.inf]
    else:
        probe += [int(x) for x in rng.integers(-2**63, 2**63 - 1, 500, dtype=np.int64)] + [-2**63, 2**63 - 1]
    pa = np.array(probe, dtype=np.float64 if typ == abi.KG_REAL else np.int64)
    ptrs, cnts = idx.lookup(Column(typ, t(pa)))
    offs = idx.offsets_of(ptrs).cpu().numpy()
    cnts = cnts.cpu().numpy()
    tids = idx.tids().cpu().numpy()
    for j, k in enumerate(pa.tolist()):
        kk = (k + 0.0) if typ == abi.KG_REAL else k
        exp = [] if (typ == abi.KG_REAL and k != k) else want.get(kk, [])
        assert tids[offs[j]:offs[j] + cnts[j]].tolist() == exp, (kind, k)
    idx.free()


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_intersect_batch(ctx, k):
    rng = np.random.default_rng(100 + k)
    ntasks = 300
    lists = []
    for _ in range(ntasks * k):
        sz = int(rng.choice([0, 1, 5, 64, 65, 500, 3000]))
        lists.append(np.unique(rng.integers(0, 6000, sz)).astype(np.int32))
    pool = np.concatenate(lists + [np.zeros(1, dtype=np.int32)])
    starts = np.cumsum([0] + [len(x) for x in lists])[:-1]
    dpool = t(pool)
    ptrs = (dpool.data_ptr() + 4 * torch.as_tensor(starts, dtype=torch.int64)).to(DEV)
    sizes = t(np.array([len(x) for x in lists], dtype=np.int32))
    nxt = rng.integers(0, 3000, ntasks).astype(np.int32)
    offs, out = intersect_batch(ctx, ptrs, sizes, k, next_tuple=t(nxt))
    offs, out = offs.cpu().numpy(), out.cpu().numpy()
    for ti in range(ntasks):
        want = po.intersect_all(lists[ti * k:(ti + 1) * k], next_tuple=int(nxt[ti]), chunk=1 << 20)
        assert np.array_equal(out[offs[ti]:offs[ti + 1]], want)


def test_tuple_set(ctx):
    rng = np.random.default_rng(9)
    a = rng.integers(0, 40, (50_000, 5)).astype(np.int32)
    b = np.concatenate([a[rng.integers(0, 50_000, 20_000)], rng.integers(0, 40, (30_000, 5)).astype(np.int32)])
    s = TupleSet(ctx, 5, 200_000)
    na = s.insert(t(a)).cpu().numpy().astype(bool)
    nb = s.insert(t(b)).cpu().numpy().astype(bool)
    want_new, want_sorted = po.tuple_dedup(np.concatenate([a, b]))
    assert np.array_equal(np.concatenate([na, nb]), want_new)
    got = s.extract().cpu().numpy()
    assert len(got) == len(want_sorted) == s.size
    # byte-lexicographic order of the little-endian int32 bytes
    order = sorted(range(len(got)), key=lambda i: got[i].tobytes())
    assert np.array_equal(got[order], want_sorted)
    s.free()


def test_tuple_set_capacity_error(ctx):
    s = TupleSet(ctx, 2, 10)
    with pytest.raises(abi.KgError):
        s.insert(t(np.arange(40, dtype=np.int32).reshape(20, 2)))
    s.free()
