"""Hash join (build / probe / probe fused with GROUP BY the matched build key)
and the high-cardinality group-by on the HIP path, against the oracle.

Bars: join pairs, cardinalities, counts and integer sums bit-exact after
sorting; groups (key, SUM(BIGINT), COUNT) bit-exact as sets, insertion order
recovered from the first probe row."""
import ctypes

import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Column, Context, HashTable
from kushdb_amd.workloads import join_tables
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def col64(x):
    return Column(abi.KG_BIGINT, torch.as_tensor(np.ascontiguousarray(x, dtype=np.int64)).to(DEV))


def sorted_pairs(p, b):
    a = np.stack([np.asarray(p), np.asarray(b)], 1)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
def test_join_pairs_and_agg_vs_oracle(ctx, theta):
    nb, npr = 100_000, 1_000_000
    print(f"join theta={theta} nb={nb} np={npr}", flush=True)
    bk, pay, pk, pv = join_tables(nb, npr, theta=theta, device=DEV, seed=99 + int(theta * 10))
    ht = HashTable(ctx, nb)
    ht.insert(Column(abi.KG_BIGINT, bk))
    assert ht.size() == nb
    pc = Column(abi.KG_BIGINT, pk)
    count = ht.probe_count(pc)
    gp, gb = ht.probe(pc)
    hb, hp, hv = bk.cpu().numpy(), pk.cpu().numpy(), pv.cpu().numpy()
    op, ob = po.hash_join([po.Col(po.BIGINT, hb)], [po.Col(po.BIGINT, hp)])
    assert count == len(op) == len(gp)
    assert np.array_equal(sorted_pairs(gp.cpu().numpy(), gb.cpu().numpy()), sorted_pairs(op, ob))
    # closed form: build keys are 0..nb-1, so a probe matches iff its key < nb
    assert count == int((hp < nb).sum())
    # fused probe + GROUP BY matched key: SUM(v), COUNT(*)
    k, s, c, first = ht.probe_agg(pc, pv, max_groups=nb, with_first=True)
    k, s, c, first = (x.cpu().numpy() for x in (k, s, c, first))
    order = np.argsort(first, kind="stable")
    k, s, c = k[order], s[order], c[order]
    ok = np.empty(nb, dtype=np.int64)
    osum = np.empty(nb, dtype=np.int64)
    ocnt = np.empty(nb, dtype=np.int64)
    g = po.lib().kor_q_join_agg(nb, hb.ctypes.data, npr, hp.ctypes.data, hv.ctypes.data, nb, ok.ctypes.data,
                                osum.ctypes.data, ocnt.ctypes.data)
    assert len(k) == g
    assert np.array_equal(k, ok[:g]) and np.array_equal(s, osum[:g]) and np.array_equal(c, ocnt[:g])
    ht.free()


def test_join_duplicates_nulls_and_edge_keys(ctx):
    rng = np.random.default_rng(5)
    b = rng.integers(-50, 50, 5000).astype(np.int64)
    b[:7] = np.iinfo(np.int64).min  # the free-slot marker value is a legal key
    b[7:9] = np.iinfo(np.int64).max
    p = rng.integers(-60, 60, 20000).astype(np.int64)
    p[:3] = np.iinfo(np.int64).min
    p[3] = np.iinfo(np.int64).max
    bn = (rng.random(5000) < 0.1).astype(np.int8)
    pn = (rng.random(20000) < 0.1).astype(np.int8)
    bcol = Column(abi.KG_BIGINT, torch.as_tensor(b).to(DEV), torch.as_tensor(bn).to(DEV))
    pcol = Column(abi.KG_BIGINT, torch.as_tensor(p).to(DEV), torch.as_tensor(pn).to(DEV))
    ht = HashTable(ctx, 5000)
    ht.insert(bcol)
    assert ht.size() == int((bn == 0).sum())
    gp, gb = ht.probe(pcol)
    op, ob = po.hash_join([po.Col(po.BIGINT, b, bn)], [po.Col(po.BIGINT, p, pn)])
    assert len(gp) == len(op) > 0
    assert np.array_equal(sorted_pairs(gp.cpu().numpy(), gb.cpu().numpy()), sorted_pairs(op, ob))
    assert ht.probe_count(pcol) == len(op)
    ht.free()


@pytest.mark.parametrize("typ,npdt", [(abi.KG_INT, np.int32), (abi.KG_SMALLINT, np.int16), (abi.KG_REAL, np.float64)])
def test_join_other_key_types(ctx, typ, npdt):
    rng = np.random.default_rng(int(typ) + 11)
    if npdt == np.float64:
        b = np.round(rng.normal(size=3000), 2)
        p = np.round(rng.normal(size=9000), 2)
        b[:3] = -0.0
        p[:5] = 0.0
        p[5] = np.nan
    else:
        b = rng.integers(-300, 300, 3000).astype(npdt)
        p = rng.integers(-300, 300, 9000).astype(npdt)
    ht = HashTable(ctx, 3000)
    ht.insert(Column(typ, torch.as_tensor(b).to(DEV)))
    gp, gb = ht.probe(Column(typ, torch.as_tensor(p).to(DEV)))
    ot = {abi.KG_INT: po.INT, abi.KG_SMALLINT: po.SMALLINT, abi.KG_REAL: po.REAL}[typ]
    op, ob = po.hash_join([po.Col(ot, b)], [po.Col(ot, p)])
    assert np.array_equal(sorted_pairs(gp.cpu().numpy(), gb.cpu().numpy()), sorted_pairs(op, ob))
    ht.free()


def test_join_empty_sides(ctx):
    ht = HashTable(ctx, 10)
    assert ht.probe_count(col64(np.arange(100))) == 0
    ht.insert(col64(np.arange(10)))
    assert ht.probe_count(col64(np.zeros(0))) == 0
    ht.free()


def test_join_capacity_error(ctx):
    ht = HashTable(ctx, 10)
    with pytest.raises(abi.KgError):
        ht.insert(col64(np.arange(11)))
    ht.free()


@pytest.mark.parametrize("ngroups,n", [(1, 1000), (1000, 200_000), (150_000, 1_000_000)])
def test_group_by_i64_vs_oracle(ctx, ngroups, n):
    rng = np.random.default_rng(ngroups)
    keys = rng.integers(0, ngroups, n).astype(np.int64) * 7919 - 3
    keys[: min(5, n)] = np.iinfo(np.int64).min
    vals = rng.integers(-2**62, 2**62, n).astype(np.int64)
    k, s, c, first = ctx.group_by_i64(torch.as_tensor(keys).to(DEV), torch.as_tensor(vals).to(DEV),
                                      with_first=True)
    k, s, c, first = (x.cpu().numpy() for x in (k, s, c, first))
    order = np.argsort(first, kind="stable")
    kv, _, v, _ = po.group_by([po.Col(po.BIGINT, keys)], [(po.SUM, po.Col(po.BIGINT, vals)), (po.COUNT, None)],
                              max_groups=len(k) + 1)
    assert len(k) == len(kv)
    assert np.array_equal(k[order], kv[:, 0])
    assert np.array_equal(s[order], v[:, 0]) and np.array_equal(c[order], v[:, 1])


def test_full_size_join_properties(ctx):
    """16M build keys / 256M probes, uniform: cardinality by closed form, fused
    GROUP BY totals equal the matched rows and their value sum (exact int64)."""
    nb, npr = 1 << 24, 1 << 28
    print(f"join full nb={nb} np={npr}", flush=True)
    bk, pay, pk, pv = join_tables(nb, npr, theta=0.0, device=DEV)
    ht = HashTable(ctx, nb)
    ht.insert(Column(abi.KG_BIGINT, bk))
    pc = Column(abi.KG_BIGINT, pk)
    hit = pk < nb
    want = int(hit.sum().item())
    assert ht.probe_count(pc) == want
    k, s, c, _ = ht.probe_agg(pc, pv, max_groups=nb)
    assert int(c.sum().item()) == want
    assert int(s.sum().item()) == int(pv[hit].sum().item())
    assert len(k) == int(torch.unique(pk[hit]).numel())
    ht.free()


@pytest.mark.parametrize("base", [0, 1 << 40, -(1 << 62)])
def test_join_narrow_table_and_widening(ctx, base):
    """Build keys spanning less than 2^32 (any base) use the 32-bit-offset table;
    a second build batch far outside that range converts it to the wide form.
    Pairs, counts and the fused probe + GROUP BY match the oracle before and
    after the conversion; probe keys just outside the range miss."""
    rng = np.random.default_rng(77)
    b1 = (base + rng.permutation(20_000)).astype(np.int64)
    p = np.concatenate([base + rng.integers(-100, 20_100, 50_000), [base - 1, base + (1 << 32)]]).astype(np.int64)
    v = rng.integers(0, 1 << 31, len(p)).astype(np.int64)
    ht = HashTable(ctx, 40_000)
    ht.insert(Column(abi.KG_BIGINT, torch.as_tensor(b1).to(DEV)))
    pc = Column(abi.KG_BIGINT, torch.as_tensor(p).to(DEV))

    def check(build):
        gp, gb = ht.probe(pc)
        op, ob = po.hash_join([po.Col(po.BIGINT, build)], [po.Col(po.BIGINT, p)])
        assert ht.probe_count(pc) == len(op) == len(gp)
        assert np.array_equal(sorted_pairs(gp.cpu().numpy(), gb.cpu().numpy()), sorted_pairs(op, ob))
        k, s, c, _ = ht.probe_agg(pc, torch.as_tensor(v).to(DEV), max_groups=len(build))
        k, s, c = (x.cpu().numpy() for x in (k, s, c))
        order = np.argsort(k)
        k, s, c = k[order], s[order], c[order]
        wk, inv = np.unique(build[ob], return_inverse=True)
        wc = np.bincount(inv, minlength=len(wk))
        ws = np.zeros(len(wk), dtype=np.int64)
        np.add.at(ws, inv, v[op])
        assert np.array_equal(k, wk) and np.array_equal(c, wc) and np.array_equal(s, ws)

    check(b1)
    b2 = np.array([base + (1 << 33) + i for i in range(1000)] + [np.iinfo(np.int64).min], dtype=np.int64)
    ht.insert(Column(abi.KG_BIGINT, torch.as_tensor(b2).to(DEV)))
    p2 = np.concatenate([p, b2[:500], [np.iinfo(np.int64).min]]).astype(np.int64)
    v = rng.integers(0, 1 << 31, len(p2)).astype(np.int64)
    pc = Column(abi.KG_BIGINT, torch.as_tensor(p2).to(DEV))
    p = p2
    allb = np.concatenate([b1, b2])
    # build row ids of the second batch restart at 0 (rows of its own column)
    gp, gb = ht.probe(pc)
    pairs = set(zip(gp.cpu().numpy().tolist(), gb.cpu().numpy().tolist()))
    want = set()
    for i, key in enumerate(p2.tolist()):
        for j in np.nonzero(b1 == key)[0].tolist():
            want.add((i, j))
        for j in np.nonzero(b2 == key)[0].tolist():
            want.add((i, j))
    assert pairs == want
    assert ht.size() == len(allb)
    ht.free()


@pytest.mark.parametrize("regions", [8, 64])
@pytest.mark.parametrize("wide", [False, True])
def test_partitioned_probe_matches_direct(ctx, regions, wide):
    """Probing through a region partition of the probe keys (forced here; chosen
    by table size otherwise) gives the same pairs, counts and grouped sums as
    the direct probe, on narrow and wide tables, with NULL probe keys."""
    rng = np.random.default_rng(regions + (7 if wide else 0))
    nb, npr = 50_000, 400_000
    b = rng.permutation(nb).astype(np.int64) * 3
    if wide:
        b[:5] = np.iinfo(np.int64).min      # wide form: the free-slot marker value as a key
        b[5] = np.iinfo(np.int64).max
    p = rng.integers(-10, 3 * nb + 10, npr).astype(np.int64)
    p[:11] = np.iinfo(np.int64).min
    pn = (rng.random(npr) < 0.05).astype(np.int8)
    v = rng.integers(0, 1 << 31, npr).astype(np.int64)
    ht = HashTable(ctx, nb)
    ht.insert(Column(abi.KG_BIGINT, torch.as_tensor(b).to(DEV)))
    pc = Column(abi.KG_BIGINT, torch.as_tensor(p).to(DEV), torch.as_tensor(pn).to(DEV))
    vd = torch.as_tensor(v).to(DEV)
    res = {}
    for reg in (0, regions):
        ctx.set_probe_regions(reg)
        gp, gb = ht.probe(pc)
        k, s, c, f = ht.probe_agg(pc, vd, max_groups=nb, with_first=True)
        order = np.argsort(k.cpu().numpy())
        res[reg] = (ht.probe_count(pc), sorted_pairs(gp.cpu().numpy(), gb.cpu().numpy()),
                    *(x.cpu().numpy()[order] for x in (k, s, c, f)))
    ctx.set_probe_regions(-1)
    a, d = res[0], res[regions]
    assert a[0] == d[0] == len(a[1]) > 0
    for x, y in zip(a[1:], d[1:]):
        assert np.array_equal(x, y)
    op, ob = po.hash_join([po.Col(po.BIGINT, b)], [po.Col(po.BIGINT, p, pn)])
    assert np.array_equal(a[1], sorted_pairs(op, ob))
    ht.free()


def test_join_tune_and_cache(ctx):
    """The join tuner returns a valid configuration, applies it, reuses it for
    the same (rows, cardinality, skew) class, and probes under it stay exact."""
    bk, pay, pk, pv = join_tables(200_000, 4_000_000, theta=0.0, device=DEV, seed=5)
    bc, pc = Column(abi.KG_BIGINT, bk), Column(abi.KG_BIGINT, pk)
    t = ctx.join_tune(bc, pc)
    assert t["load_factor"] in (0.35, 0.5, 0.65)
    assert t["probe_wg_per_cu"] in (8, 16, 32) and t["probe_regions"] in (0, 64, 128)
    assert t["from_cache"] == 0 and t["trials"] == 7 and t["probe_ms"] > 0
    again = ctx.join_tune(bc, pc)
    assert again["from_cache"] == 1 and again["tuning_class"] == t["tuning_class"]
    ht = HashTable(ctx, 200_000, load_factor=t["load_factor"])
    ht.insert(bc)
    hp = pk.cpu().numpy()
    assert ht.probe_count(pc) == int((hp < 200_000).sum())
    ht.free()
    # a heavily skewed probe side is another class
    skewed = torch.zeros_like(pk)
    s = ctx.join_tune(bc, Column(abi.KG_BIGINT, skewed))
    assert s["tuning_class"] != t["tuning_class"] and s["from_cache"] == 0
    ctx.set_probe_regions(0)


@pytest.mark.parametrize("kind,regions,lf", [("narrow", -1, 0.5), ("narrow", 8192, 0.8), ("wide", -1, 0.5),
                                             ("wide", 8192, 0.8), ("narrow_sel", -1, 0.35)])
def test_region_build_equals_direct(kind, regions, lf):
    """The region-partitioned build (kg_ctx_set_build_regions) against direct
    inserts: 2.1M build rows with duplicate keys and NULLs, in two build calls
    (the second into a non-empty table), narrow (keys within 2^32) and wide
    (keys across int64, with the free-marker key INT64_MIN) forms, at the
    default region size and with regions small enough (load factor 0.8) that
    many keys run off their region's end and take the spill path. The join
    pairs and the fused GROUP BY must be equal to the direct build's and the
    pair set to the oracle's."""
    rng = np.random.default_rng(sum(kind.encode()) + regions % 97)
    nb = 2_100_000
    if kind.startswith("narrow"):
        keys = rng.integers(0, 1_500_000, nb) + 10**12
    else:
        keys = rng.integers(-2**63, 2**63 - 1, nb, dtype=np.int64)
        keys[rng.integers(0, nb, 5)] = np.iinfo(np.int64).min
        keys[: nb // 3] = keys[nb // 3: 2 * (nb // 3)]  # duplicates
    nulls = (rng.random(nb) < 0.02).astype(np.int8)
    probe = np.concatenate([keys[rng.integers(0, nb, 600_000)], rng.integers(-2**62, 2**62, 100_000)])
    pv = rng.integers(0, 1 << 20, len(probe)).astype(np.int64)
    bcol = Column(abi.KG_BIGINT, torch.as_tensor(keys).to(DEV), torch.as_tensor(nulls).to(DEV))
    pcol, pvals = col64(probe), torch.as_tensor(pv).to(DEV)
    sel = None
    if kind == "narrow_sel":
        sel = torch.as_tensor(np.nonzero(rng.random(nb) < 0.9)[0].astype(np.int32)).to(DEV)
    results = []
    for mode in (0, regions):
        c = Context(0)
        c.set_build_regions(mode)
        ht = HashTable(c, nb, load_factor=lf)
        rows = sel if sel is not None else torch.arange(nb, dtype=torch.int32, device=DEV)
        half = rows.numel() // 2
        print(f"region build {kind} mode={mode} lf={lf}", flush=True)
        ht.insert(bcol, rows=rows[:half])
        ht.insert(bcol, rows=rows[half:])
        cnt = ht.probe_count(pcol)
        gp, gb = ht.probe(pcol)
        k, s, n, _ = ht.probe_agg(pcol, pvals, max_groups=nb)
        agg = np.stack([x.cpu().numpy() for x in (k, s, n)], 1)
        results.append((ht.size(), cnt, sorted_pairs(gp.cpu().numpy(), gb.cpu().numpy()),
                        agg[np.lexsort((agg[:, 2], agg[:, 1], agg[:, 0]))]))
        ht.free()
        c.close()
    (s0, c0, p0, a0), (s1, c1, p1, a1) = results
    assert s0 == s1 == int((nulls[(sel.cpu().numpy() if sel is not None else np.arange(nb))] == 0).sum())
    assert c0 == c1 and np.array_equal(p0, p1) and np.array_equal(a0, a1)
    hs = sel.cpu().numpy() if sel is not None else np.arange(nb)
    bmask = np.zeros(nb, dtype=bool)
    bmask[hs] = True
    bmask &= nulls == 0
    op, ob = po.hash_join([po.Col(po.BIGINT, np.where(bmask, keys, 0), (~bmask).astype(np.int8))],
                          [po.Col(po.BIGINT, probe)])
    assert np.array_equal(p1, sorted_pairs(op, ob))


@pytest.mark.parametrize("kind,regions", [("uniform", 1024), ("uniform", 128), ("zipf1", 1024), ("dup_nulls", 512),
                                          ("wide", 1024), ("wide_marker", 1024)])
def test_region_agg_equals_atomic(kind, regions):
    """Probe + GROUP BY partitioned by table region (kg_ctx_set_agg_regions)
    against the per-slot atomic form on the same table: uniform and Zipf(1)
    probes (frequent keys summed during the partition) with 10% misses,
    duplicate build keys with NULL probe keys (narrow table), a wide table
    (keys across int64), and a wide table whose probes carry the free-marker key
    INT64_MIN (the call then runs the atomic form). Groups equal as sets; the
    unique-key cases also equal the oracle's."""
    rng = np.random.default_rng(sum(kind.encode()) + regions)
    pnull = None
    if kind in ("uniform", "zipf1"):
        nb, npr = 300_000, 3_000_000
        bk, _, pk, pv = join_tables(nb, npr, theta=1.0 if kind == "zipf1" else 0.0, device=DEV,
                                    seed=7 + regions)
        keys, probe, vals = bk.cpu().numpy(), pk.cpu().numpy(), pv.cpu().numpy()
    elif kind == "dup_nulls":
        nb, npr = 400_000, 2_000_000
        keys = rng.integers(0, 250_000, nb) + 5 * 10**9
        probe = np.concatenate([keys[rng.integers(0, nb, npr - 200_000)], rng.integers(0, 10**10, 200_000)])
        vals = rng.integers(-(1 << 40), 1 << 40, npr)
        pnull = (rng.random(npr) < 0.03).astype(np.int8)
    else:
        nb, npr = 300_000, 2_000_000
        keys = rng.integers(-2**63, 2**63 - 1, nb, dtype=np.int64)
        keys[: nb // 4] = keys[nb // 4: nb // 2]
        probe = np.concatenate([keys[rng.integers(0, nb, npr - 100_000)], rng.integers(-2**62, 2**62, 100_000)])
        if kind == "wide_marker":
            keys[rng.integers(0, nb, 3)] = np.iinfo(np.int64).min
            probe[rng.integers(0, npr, 50)] = np.iinfo(np.int64).min
        vals = rng.integers(0, 1 << 32, npr)
    bcol = col64(keys)
    pcol = Column(abi.KG_BIGINT, torch.as_tensor(np.ascontiguousarray(probe, dtype=np.int64)).to(DEV),
                  torch.as_tensor(pnull).to(DEV) if pnull is not None else None)
    pvals = torch.as_tensor(np.ascontiguousarray(vals, dtype=np.int64)).to(DEV)
    results, paths = [], []
    for mode in (0, regions):
        c = Context(0)
        c.set_agg_regions(mode)
        ht = HashTable(c, nb, load_factor=0.5)
        ht.insert(bcol)
        print(f"region agg {kind} mode={mode}", flush=True)
        k, s, n, _ = ht.probe_agg(pcol, pvals, max_groups=nb)
        paths.append(c.last_path())
        agg = np.stack([x.cpu().numpy() for x in (k, s, n)], 1)
        results.append(agg[np.lexsort((agg[:, 2], agg[:, 1], agg[:, 0]))])
        ht.free()
        c.close()
    assert paths == ["hashed", "hashed" if kind == "wide_marker" else "regions"]
    assert np.array_equal(results[0], results[1])
    if kind in ("uniform", "zipf1"):
        ok = np.empty(nb, dtype=np.int64)
        osum = np.empty(nb, dtype=np.int64)
        ocnt = np.empty(nb, dtype=np.int64)
        g = po.lib().kor_q_join_agg(nb, keys.ctypes.data, npr, probe.ctypes.data, vals.ctypes.data, nb,
                                    ok.ctypes.data, osum.ctypes.data, ocnt.ctypes.data)
        want = np.stack([ok[:g], osum[:g], ocnt[:g]], 1)
        assert np.array_equal(results[1], want[np.lexsort((want[:, 2], want[:, 1], want[:, 0]))])
