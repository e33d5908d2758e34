"""ORDER BY on the device (kg_order_by) against the reference's order_by
fixtures (end_to_end_test/order_by/*_test.cc: columns loaded from the fixture
files with kg_file_*, ordered on the GPU, printed with kg_format_rows, equal to
the expected .tbl ascending and to its lines reversed descending), and against
the oracle's merge sort (kor_order_by) on random multi-key inputs with ties,
NULLs, -0.0, TEXT and ENUM keys — positions must match exactly, so rows equal
on every key come out in the reference's order (reverse input order)."""
import os

import numpy as np
import pytest
import torch

import kdb
from kushdb_amd import abi
from kushdb_amd.runtime import Column, Context, KgError, TextColumn, format_rows, order_by
from oracle import pyoracle as po
from test_oracle_order import FIXTURES, T

pytestmark = pytest.mark.gpu

KG = {"int": abi.KG_INT, "bigint": abi.KG_BIGINT, "smallint": abi.KG_SMALLINT, "date": abi.KG_DATE,
      "real": abi.KG_REAL, "enum": abi.KG_ENUM}
TORCH = {abi.KG_SMALLINT: torch.int16, abi.KG_INT: torch.int32, abi.KG_BIGINT: torch.int64,
         abi.KG_REAL: torch.float64, abi.KG_DATE: torch.int32, abi.KG_ENUM: torch.int32,
         abi.KG_BOOLEAN: torch.int8}


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.file_release()


def device_column(ctx, name, typ):
    """(device key, format spec) of a fixture column loaded by kg_file_*."""
    path = os.path.join(kdb.DATA, name + ".kdb")
    if typ == "text":
        col = ctx.file_text(path)
        off = col.offsets.cpu().numpy()
        heap = col.data.cpu().numpy().tobytes()
        strs = [heap[off[i]:off[i + 1]].decode() for i in range(len(off) - 1)]
        nulls = None if col.nulls is None else col.nulls.cpu().numpy()
        return col, None, (abi.KG_TEXT, strs, nulls, None)
    col = ctx.file_column(path, KG[typ])
    d = ctx.file_enum(os.path.join(kdb.DATA, name + ".kdbenum")) if typ == "enum" else None
    nulls = None if col.nulls is None else col.nulls.cpu().numpy()
    return col, d, (KG[typ], col.data.cpu().numpy(), nulls, d)


@pytest.mark.parametrize("asc", [True, False])
@pytest.mark.parametrize("fixture", sorted(FIXTURES))
def test_order_by_fixture(ctx, fixture, asc):
    print("order_by fixture", fixture, "asc" if asc else "desc", flush=True)
    keys, outs = FIXTURES[fixture]
    loaded = [device_column(ctx, n, t) for n, t in keys]
    dicts = {i: d for i, (_, d, _) in enumerate(loaded) if d is not None}
    perm = order_by(ctx, [c for c, _, _ in loaded], [asc] * len(keys), dicts)
    text = format_rows([device_column(ctx, n, t)[2] for n, t in outs], perm.cpu().numpy())
    exp = kdb.expected(f"order_by/{fixture}_expected.tbl")
    assert text.split("\n")[:-1] == (exp if asc else exp[::-1])


def _dev(typ, vals, nulls):
    return Column(typ, torch.as_tensor(vals).to("cuda:0"),
                  None if nulls is None else torch.as_tensor(nulls).to("cuda:0"))


def _random_keys(rng, n, card, with_nulls):
    out = {}
    for typ in (abi.KG_SMALLINT, abi.KG_INT, abi.KG_BIGINT, abi.KG_REAL, abi.KG_DATE):
        if typ == abi.KG_REAL:
            v = rng.integers(-card, card, n).astype(np.float64) / 8
            v[rng.random(n) < 0.05] = -0.0
        elif typ == abi.KG_BIGINT:
            v = rng.integers(-card, card, n, dtype=np.int64) * (1 << 40)
        else:
            v = rng.integers(-card, card, n).astype(po._NP[typ])
        nulls = (rng.random(n) < 0.1).astype(np.int8) if with_nulls else None
        out[typ] = (v, nulls)
    return out


@pytest.mark.parametrize("n,card,seed", [(1, 4, 0), (2, 1, 1), (37, 3, 2), (1000, 5, 3), (65536, 40, 4),
                                         (300001, 1000, 5), (2_000_000, 30000, 6)])
def test_order_by_random_vs_oracle(ctx, n, card, seed):
    print("order_by random", n, card, seed, flush=True)
    rng = np.random.default_rng(seed)
    keys = _random_keys(rng, n, card, with_nulls=seed % 2 == 0)
    types = list(keys)
    for trial in range(4):
        pick = [types[i] for i in rng.choice(len(types), 1 + trial % 3, replace=False)]
        ascs = [bool(rng.integers(2)) for _ in pick]
        want = po.order_by([(po.Col(t, *keys[t]), a) for t, a in zip(pick, ascs)])
        got = order_by(ctx, [_dev(t, *keys[t]) for t in pick], ascs).cpu().numpy()
        assert np.array_equal(got, want), (n, seed, trial, pick, ascs)


@pytest.mark.parametrize("n,seed", [(50, 0), (5000, 1), (200000, 2)])
def test_order_by_text_enum_vs_oracle(ctx, n, seed):
    print("order_by text/enum", n, seed, flush=True)
    rng = np.random.default_rng(50 + seed)
    words = ["", "a", "ab", "b", "abc", "B", "é", "aé", "zz", "abd", "aba"]
    words += [f"w{i:03d}" for i in range(200)]
    dic = ["pear", "apple", "fig", "apple pie", "Fig", "", "zebra", "ap"]
    tv = [words[i] for i in rng.integers(0, len(words), n)]
    tn = (rng.random(n) < 0.1).astype(np.int8)
    ev = rng.integers(0, len(dic), n).astype(np.int32)
    en = (rng.random(n) < 0.1).astype(np.int8)
    iv = rng.integers(-3, 4, n).astype(np.int32)
    text_dev = TextColumn.from_strings(tv, nulls=tn)
    cols = {"text": (("text", tv, tn), text_dev), "enum": (po.Col(po.ENUM, ev, en, dictionary=dic),
                                                           _dev(abi.KG_ENUM, ev, en)),
            "int": (po.Col(po.INT, iv), _dev(abi.KG_INT, iv, None))}
    for trial in range(6):
        names = list(rng.permutation(list(cols))[: 1 + trial % 3])
        ascs = [bool(rng.integers(2)) for _ in names]
        want = po.order_by([(cols[c][0], a) for c, a in zip(names, ascs)])
        dicts = {i: dic for i, c in enumerate(names) if c == "enum"}
        got = order_by(ctx, [cols[c][1] for c in names], ascs, dicts).cpu().numpy()
        assert np.array_equal(got, want), (n, seed, trial, names, ascs)


def test_order_by_empty_and_boolean(ctx):
    e = _dev(abi.KG_INT, np.zeros(0, np.int32), None)
    assert order_by(ctx, [e], [True]).numel() == 0
    b = _dev(abi.KG_BOOLEAN, np.array([1, 0, 1], np.int8), None)
    with pytest.raises(KgError):
        order_by(ctx, [b], [True])


def test_order_by_sorted_property_large(ctx):
    """64M BIGINT keys with NULLs: the permutation is a permutation, the keys it
    yields are non-decreasing with NULLs last, and equal keys appear in
    decreasing input position."""
    n = 64 << 20
    print("order_by large", n, flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(9)
    vals = torch.randint(-(1 << 20), 1 << 20, (n,), dtype=torch.int64, device="cuda:0", generator=g)
    nulls = (torch.rand(n, device="cuda:0", generator=g) < 0.01).to(torch.int8)
    perm = order_by(ctx, [Column(abi.KG_BIGINT, vals, nulls)], [True]).long()
    assert torch.equal(torch.sort(perm).values, torch.arange(n, device="cuda:0"))
    k = vals[perm]
    nl = nulls[perm].bool()
    nn = int((~nl).sum())
    assert not nl[:nn].any() and nl[nn:].all()
    k, p = k[:nn], perm[:nn]
    assert bool((k[1:] >= k[:-1]).all())
    ties = k[1:] == k[:-1]
    assert bool((p[1:][ties] < p[:-1][ties]).all())
    assert bool((perm[nn + 1:] < perm[nn:-1]).all())
