"""The Skinner adaptive join on the HIP path (kg_skinner_join): UCT-chosen
join orders, device episodes over slices of the leftmost table, tuple-id
dedup. Checked against the reference's skinner_join fixtures (the
*_expected.tbl outputs of end_to_end_test/skinner_join/*_test.cc) and against
the oracle's multi-way join on random join graphs.

Bar: the result tuple set is identical (tuple ids = positions in the
children), for every budget and seed — the order the agent picks must not
change the set."""
import numpy as np
import pytest
import torch

import kdb
import queries
from engines import GpuEngine
from kushdb_amd import abi
from kushdb_amd.chain import skinner_chain_query
from kushdb_amd.runtime import Column, Context, skinner_join
from kushdb_amd.workloads import skinner_tables
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def E():
    return GpuEngine()


@pytest.mark.parametrize("t", sorted(queries.SKINNER) + ["multiple_tables_join"])
@pytest.mark.parametrize("budget,seed", [(0, 1), (7, 2), (150, 3)])
def test_skinner_fixture(E, t, budget, seed):
    print(f"skinner fixture {t} budget {budget} seed {seed}", flush=True)
    got = queries.skinner(E, t, budget=budget, seed=seed)
    assert sorted(got) == sorted(kdb.expected(f"skinner_join/{t}_expected.tbl"))
    st = E.last_stats
    assert st["distinct_tuples"] == len(got) and st["result_tuples"] >= len(got) and st["episodes"] >= 1


KEY_TYPES = [(abi.KG_INT, np.int32), (abi.KG_SMALLINT, np.int16), (abi.KG_BIGINT, np.int64),
             (abi.KG_REAL, np.float64), (abi.KG_DATE, np.int32), (abi.KG_ENUM, np.int32)]

GRAPHS = {
    "chain5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
    "star4": (4, [(0, 1), (0, 2), (0, 3)]),
    "cycle4": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "clique3": (3, [(0, 1), (1, 2), (0, 2)]),
    "disconnected": (4, [(0, 1), (2, 3)]),
    "two_conds_one_pair": (3, [(0, 1), (0, 1), (1, 2)]),
}

# Key domain = largest table / FANOUT: results in the thousands for every shape.
FANOUT = {"chain5": 4, "star4": 4, "cycle4": 8, "clique3": 25, "disconnected": 4, "two_conds_one_pair": 25}


def random_case(graph, seed, sel=False):
    """Tables of 20..400 tuples, keys over small domains, NULLs, mixed key
    types and widths (mixed widths joined through a 64-bit index);
    optionally children that are selections of larger tables."""
    rng = np.random.default_rng(seed)
    nt, edges = GRAPHS[graph]
    cards = [int(rng.integers(20, 400)) for _ in range(nt)]
    if graph == "disconnected":
        cards = [int(rng.integers(5, 40)) for _ in range(nt)]
    base = [c * 2 if sel else c for c in cards]
    rows = [np.sort(rng.choice(b, c, replace=False)).astype(np.int32) if sel else None for b, c in zip(base, cards)]
    hconds, dconds = [], []
    for i, (a, b) in enumerate(edges):
        kt, dt = KEY_TYPES[(seed + i) % len(KEY_TYPES)]
        if kt in (abi.KG_INT, abi.KG_SMALLINT, abi.KG_DATE, abi.KG_ENUM) and (seed + i) % 3 == 0:
            kt2, dt2 = abi.KG_BIGINT, np.int64   # mixed widths: a 64-bit index
        else:
            kt2, dt2 = kt, dt
        dom = max(4, max(base) // FANOUT[graph])  # about FANOUT matches per key
        lk = rng.integers(0, dom, base[a]).astype(dt)
        rk = rng.integers(0, dom, base[b]).astype(dt2)
        ln = (rng.random(base[a]) < 0.1).astype(np.int8)
        rn = (rng.random(base[b]) < 0.1).astype(np.int8) if i % 2 else None
        hconds.append((a, lk, ln, b, rk, rn))
        lc = Column(kt, torch.from_numpy(lk).to(DEV), torch.from_numpy(ln).to(DEV))
        rc = Column(kt2, torch.from_numpy(rk).to(DEV), torch.from_numpy(rn).to(DEV) if rn is not None else None)
        dconds.append((a, lc, b, rc))
    tables = [(c, torch.from_numpy(r).to(DEV) if r is not None else None) for c, r in zip(cards, rows)]
    return cards, rows, hconds, tables, dconds


@pytest.mark.parametrize("graph", sorted(GRAPHS))
@pytest.mark.parametrize("seed", [1, 2])
def test_skinner_random_graphs_vs_oracle(graph, seed):
    print(f"skinner random {graph} seed {seed}", flush=True)
    ctx = Context(0)
    cards, rows, hconds, tables, dconds = random_case(graph, seed, sel=seed % 2 == 0)
    want = np.array(po.skinner_join(cards, hconds, rows), dtype=np.int32).reshape(-1, len(cards))
    print(f"  {len(want)} result tuples", flush=True)
    for budget, jseed in [(0, 1), (16, 5), (999, 9)]:
        got, st = skinner_join(ctx, tables, dconds, budget=budget, seed=jseed)
        g = got.cpu().numpy()
        assert np.array_equal(g[np.lexsort(g.T[::-1])], want), (graph, budget, jseed)
        assert st["distinct_tuples"] == len(want)
    ctx.close()


def test_skinner_bigint_1m_by_1m():
    """1M x 1M BIGINT equi-join through the device column index (keys spread
    over the int64 range, so no 32-bit narrowing applies), against the
    oracle's chained hash join."""
    ctx = Context(0)
    rng = np.random.default_rng(8)
    n = 1_000_000
    dom = rng.integers(-2**62, 2**62, 600_000)
    a = dom[rng.integers(0, len(dom), n)]
    b = dom[rng.integers(0, len(dom), n)]
    miss = rng.random(n) < 0.2
    b[miss] = rng.integers(-2**62, 2**62, int(miss.sum()))
    ca = Column(abi.KG_BIGINT, torch.from_numpy(a).to(DEV))
    cb = Column(abi.KG_BIGINT, torch.from_numpy(b).to(DEV))
    op, ob = po.hash_join([po.Col(po.BIGINT, a)], [po.Col(po.BIGINT, b)])
    want = np.stack([ob, op], 1).astype(np.int32)
    want = want[np.lexsort(want.T[::-1])]
    got, st = skinner_join(ctx, [(n, None), (n, None)], [(0, ca, 1, cb)], budget=1 << 22, seed=3)
    g = got.cpu().numpy()
    assert np.array_equal(g[np.lexsort(g.T[::-1])], want)
    assert st["distinct_tuples"] == len(want)
    ctx.close()


@pytest.mark.parametrize("kt", [abi.KG_INT, abi.KG_BIGINT, abi.KG_REAL])
def test_skinner_cycle4_intersection(kt):
    """A cyclic 4-table join (t0-t1-t2-t3-t0): the last table of every order
    is linked by two conditions, which the episode joins through the
    intersection of both index buckets. Tables of 600 tuples with NULLs,
    against the oracle's result set."""
    ctx = Context(0)
    rng = np.random.default_rng(int(kt) + 40)
    dt = {abi.KG_INT: np.int32, abi.KG_BIGINT: np.int64, abi.KG_REAL: np.float64}[kt]
    card = 600
    edges = [(0, 1), (1, 2), (2, 3), (3, 0)]
    hconds, dconds = [], []
    for i, (x, y) in enumerate(edges):
        lk = rng.integers(0, 100, card).astype(dt)
        rk = rng.integers(0, 100, card).astype(dt)
        if kt == abi.KG_REAL:
            lk[lk == 0] = -0.0                        # -0.0 joins 0.0
        ln = (rng.random(card) < 0.05).astype(np.int8)
        hconds.append((x, lk, ln, y, rk, None))
        dconds.append((x, Column(kt, torch.from_numpy(lk).to(DEV), torch.from_numpy(ln).to(DEV)), y,
                       Column(kt, torch.from_numpy(rk).to(DEV))))
    want = np.array(po.skinner_join([card] * 4, hconds), dtype=np.int32).reshape(-1, 4)
    print(f"  cycle4 {len(want)} result tuples", flush=True)
    for budget, seed in [(0, 1), (5000, 7)]:
        got, st = skinner_join(ctx, [(card, None)] * 4, dconds, budget=budget, seed=seed)
        g = got.cpu().numpy()
        assert np.array_equal(g[np.lexsort(g.T[::-1])], want), (kt, budget, seed)
    ctx.close()


def test_skinner_empty_and_single_table():
    ctx = Context(0)
    k = Column(abi.KG_INT, torch.arange(10, dtype=torch.int32, device=DEV))
    e = Column(abi.KG_INT, torch.zeros(0, dtype=torch.int32, device=DEV))
    got, st = skinner_join(ctx, [(10, None), (0, None)], [(0, k, 1, e)])
    assert got.shape[0] == 0 and st["episodes"] == 0
    got, st = skinner_join(ctx, [(10, None)], [])
    assert sorted(got.view(-1).tolist()) == list(range(10))
    with pytest.raises(abi.KgError):
        skinner_join(ctx, [(10, None), (10, None)], [(0, k, 0, k)])   # a table joined with itself
    ctx.close()


def test_skinner_chain_query_matches_fixed_order():
    """The five-table chain query (TEXT prefix + LIKE + int filters) with the
    join order left to the agent: the same tuples and MINs as the fixed-order
    device pipeline, whose result the chain tests check against the oracle."""
    from kushdb_amd.chain import chain_query
    ctx = Context(0)
    tabs, spec, htabs = skinner_tables(scale=0.01, theta=1.0, device=DEV, seed=21)
    t0, a0, b0, _ = chain_query(ctx, tabs, spec, [0, 1, 2, 3, 4])
    for budget, seed in [(0, 1), (5000, 4)]:
        t1, a1, b1, st = skinner_chain_query(ctx, tabs, spec, budget=budget, seed=seed)
        g0, g1 = t0.cpu().numpy(), t1.cpu().numpy()
        assert np.array_equal(g0[np.lexsort(g0.T[::-1])], g1[np.lexsort(g1.T[::-1])])
        assert a0 == a1 and b0 == b1
        assert st["episodes"] >= 1 and len(st["last_order"]) == 5
    ctx.close()


def test_cross_product_fixture(E):
    """CrossProductOperator (cross_product_test.cc) on the device: the join with
    no conditions, every episode a cross product of its slice."""
    assert sorted(queries.cross_product(E)) == sorted(kdb.expected("cross_product/cross_product_expected.tbl"))
