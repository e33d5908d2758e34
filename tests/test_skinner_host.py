"""Skinner join, host side (no GPU): the oracle's multi-way join result set
pinned on the reference's skinner_join fixtures and on brute force, and the
UCT join-order agent (kg_uct_*, runtime/skinner_join_executor.cc:559-838).

The agent only picks orders; the result set does not depend on them, so the
agent's own choices are checked by properties (permutations, the prefix, the
connectivity heuristic, each first table tried once before UCB, convergence
to the rewarded order, determinism under a fixed seed), not against the
reference's random sequence (parity unpinned: the reference's std::shuffle /
default_random_engine draws are not restated)."""
import itertools

import numpy as np
import pytest

import kdb
import queries
from engines import OracleEngine
from kushdb_amd import abi
from kushdb_amd.runtime import UctAgent
from oracle import pyoracle as po

E = OracleEngine()


@pytest.mark.parametrize("t", sorted(queries.SKINNER) + ["multiple_tables_join"])
def test_oracle_skinner_fixture(t):
    assert sorted(queries.skinner(E, t)) == sorted(kdb.expected(f"skinner_join/{t}_expected.tbl"))


def brute_force(cards, conds):
    out = []
    for tp in itertools.product(*[range(c) for c in cards]):
        ok = True
        for lt, lk, ln, rt, rk, rn in conds:
            if (ln is not None and ln[tp[lt]]) or (rn is not None and rn[tp[rt]]) or lk[tp[lt]] != rk[tp[rt]]:
                ok = False
                break
        if ok:
            out.append(tp)
    return out


SHAPES = {
    "pair": [(0, 1)],
    "chain4": [(0, 1), (1, 2), (2, 3)],
    "star4": [(0, 1), (0, 2), (0, 3)],
    "cycle3": [(0, 1), (1, 2), (2, 0)],
    "disconnected3": [(0, 1)],
    "double_edge": [(0, 1), (0, 1), (1, 2)],
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_multiway_join_vs_brute_force(shape, seed):
    rng = np.random.default_rng(seed * 31 + len(shape))
    edges = SHAPES[shape]
    nt = 1 + max(max(e) for e in edges)
    if shape == "disconnected3":
        nt = 3
    cards = [int(rng.integers(1, 7)) for _ in range(nt)]
    conds = []
    for a, b in edges:
        lk = rng.integers(0, 4, cards[a]).astype(np.int32)
        rk = rng.integers(0, 4, cards[b]).astype(np.int64)
        ln = (rng.random(cards[a]) < 0.2).astype(np.int8)
        rn = None if seed else (rng.random(cards[b]) < 0.2).astype(np.int8)
        conds.append((a, lk, ln, b, rk, rn))
    assert po.skinner_join(cards, conds) == brute_force(cards, conds)


def test_oracle_empty_table():
    k = np.arange(3, dtype=np.int32)
    assert po.skinner_join([3, 0], [(0, k, None, 1, k[:0], None)]) == []


# ---------------------------------------------------------------- UCT agent
CHAIN5 = [(0, 1), (1, 2), (2, 3), (3, 4)]


def connected_order(order, edges):
    adj = {(a, b) for a, b in edges} | {(b, a) for a, b in edges}
    return all(any((order[i], order[j]) in adj for j in range(i)) for i in range(1, len(order)))


def run(agent, rounds, reward):
    orders = []
    for _ in range(rounds):
        o = agent.next()
        orders.append(o)
        agent.reward(reward(o))
    return orders


def test_uct_orders_are_connected_permutations():
    a = UctAgent(5, CHAIN5, seed=3)
    for o in run(a, 200, lambda o: 0.1):
        assert sorted(o) == list(range(5))
        assert connected_order(o, CHAIN5), o


def test_uct_tries_every_first_table_once_first():
    a = UctAgent(5, CHAIN5, seed=5)
    firsts = [o[0] for o in run(a, 5, lambda o: 0.0)]
    assert sorted(firsts) == list(range(5))


def test_uct_converges_on_rewarded_order():
    a = UctAgent(5, CHAIN5, seed=7)
    orders = run(a, 300, lambda o: 1.0 if o[:2] == [2, 3] else 0.01)
    assert all(o[:2] == [2, 3] for o in orders[-100:])


def test_uct_prefix_and_seed():
    a = UctAgent(4, [(0, 1), (1, 2), (2, 3)], seed=11, prefix=(1,))
    b = UctAgent(4, [(0, 1), (1, 2), (2, 3)], seed=11, prefix=(1,))
    oa = run(a, 60, lambda o: 0.3 if o[1] == 2 else 0.2)
    ob = run(b, 60, lambda o: 0.3 if o[1] == 2 else 0.2)
    assert oa == ob
    assert all(o[0] == 1 for o in oa)


def test_uct_disconnected_graph_orders_every_table():
    a = UctAgent(4, [(0, 1)], seed=2)
    for o in run(a, 50, lambda o: 0.5):
        assert sorted(o) == [0, 1, 2, 3]


def test_uct_misuse():
    a = UctAgent(3, [(0, 1), (1, 2)], seed=1)
    with pytest.raises(abi.KgError):
        a.reward(0.5)          # no order outstanding
    a.next()
    with pytest.raises(abi.KgError):
        a.next()               # the last order has no reward yet
    with pytest.raises(abi.KgError):
        a.reward(-1.0)
    a.reward(0.0)
    with pytest.raises(abi.KgError):
        UctAgent(3, [], prefix=(0, 0))


def test_oracle_cross_product_fixture():
    assert sorted(queries.cross_product(E)) == sorted(kdb.expected("cross_product/cross_product_expected.tbl"))
