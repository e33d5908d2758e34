"""The oracle functions timed as the bench's CPU columns, pinned on closed
forms: the C2 join's match count and SUM against numpy set membership, and
the C3 chain query (filters, four chained joins, MIN over TEXT and INT)
against a numpy restatement that follows each tuple's foreign keys from t4
back to t0 (each row of t_i references exactly one row of t_{i-1})."""
import numpy as np

from kushdb_amd.workloads import SKINNER_SIZES, text_column, vocabulary
from oracle import pyoracle as po


def host_chain_tables(scale, seed=5):
    """Host-only tables of the chain query's shape (sizes SKINNER_SIZES * scale)."""
    rng = np.random.default_rng(seed)
    sizes = [max(4, int(n * scale)) for n in SKINNER_SIZES]
    vocab = vocabulary(nwords=5000, seed=seed)
    ht = []
    for i, n in enumerate(sizes):
        t = {"id": np.arange(n, dtype=np.int32), "v": rng.integers(0, 1000, n).astype(np.int32)}
        if i > 0:
            t["fk"] = rng.integers(0, sizes[i - 1], n).astype(np.int32)
        if i in (0, 1):
            off, heap, _ = text_column(n, vocab, rng)
            name = "name" if i == 0 else "s"
            t[name + "_off"], t[name + "_heap"] = off, heap
        ht.append(t)
    return ht


def test_time_join_vs_isin():
    rng = np.random.default_rng(1)
    b = rng.permutation(200_000).astype(np.int64) * 3
    p = rng.integers(0, 700_000, 500_000)
    v = rng.integers(0, 1 << 32, len(p))
    m, s, bs, ps = po.time_join(b, p, v)
    hit = np.isin(p, b)
    assert m == int(hit.sum())
    assert s == int(v[hit].sum())
    assert bs >= 0 and ps >= 0


def test_chain_query_vs_fk_walk():
    ht = host_chain_tables(0.01)
    n, name, vmin = po.chain_query(ht)
    # closed form: walk every t4 row back through its foreign keys
    names = [bytes(ht[0]["name_heap"][ht[0]["name_off"][r]:ht[0]["name_off"][r + 1]]) for r in range(len(ht[0]["id"]))]
    svals = [bytes(ht[1]["s_heap"][ht[1]["s_off"][r]:ht[1]["s_off"][r + 1]]) for r in range(len(ht[1]["id"]))]
    import re
    like = re.compile(rb"\A(?:.|\n)*e.a(?:.|\n)*\Z".replace(rb"(?:.|\n)", rb"[^\n]"))
    ok1 = np.array([like.match(s) is not None for s in svals])
    ok0 = np.array([x.startswith(b"a") for x in names])
    r4 = np.nonzero(ht[4]["v"] < 50)[0]
    r3 = ht[4]["fk"][r4]
    r2 = ht[3]["fk"][r3]
    r1 = ht[2]["fk"][r2]
    r0 = ht[1]["fk"][r1]
    keep = ok1[r1] & ok0[r0]
    assert n == int(keep.sum()) and n > 0
    if n:
        assert vmin == int(ht[4]["v"][r4[keep]].min())
        assert name == min(names[r] for r in r0[keep])
