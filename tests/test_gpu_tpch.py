"""TPC-H Q3, Q5, Q7, Q9, Q10, Q11, Q12, Q14, Q18 and Q19 end to end on the device operators. Q3 as the reference's
q03_test.cc plans it: TEXT select (kg_text_select), scan-selects (kg_scan_select),
two hash joins (kg_ht_build / kg_ht_probe over gathered keys), the derived
revenue (kg_expr_eval), GROUP BY over three hashed keys (kg_hash_group_agg)
and ORDER BY revenue DESC, o_orderdate (kg_order_by). Checked against the
reference's q03_expected.tbl (CHECK_EQ_TBL: REAL within 1e-5) and against the
oracle's plan (same rows in the same order; revenue within 1e-12 relative)."""
import numpy as np
import pytest
import torch

import kdb
import tpch_q03
from kushdb_amd import abi
from kushdb_amd.expr import ColRef, Literal, Mul, Sub, expr_eval
from kushdb_amd.runtime import Column, Context, HashTable, TextColumn, gather, order_by, text_select
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def dev(v, typ):
    return Column(typ, torch.from_numpy(np.ascontiguousarray(v)).to(DEV))


def device_q03(ctx, c):
    d = po.build_date(*tpch_q03.DATE)
    seg = TextColumn.from_strings(c["c_mktsegment"], device=DEV)
    rows_c = text_select(ctx, seg, abi.KG_TEXT_EQ, "AUTOMOBILE")
    odate = dev(c["o_orderdate"], abi.KG_DATE)
    rows_o = ctx.scan_select([odate], [(0, abi.KG_LT, d)])
    # customer JOIN orders ON c_custkey = o_custkey (build: selected customers)
    ck = gather(ctx, dev(c["c_custkey"], abi.KG_INT), rows_c)
    ok = gather(ctx, dev(c["o_custkey"], abi.KG_INT), rows_o)
    ht = HashTable(ctx, max(len(ck), 1))
    ht.insert(ck)
    pi, _ = ht.probe(ok)
    ht.free()
    co = rows_o[pi.long()].contiguous()                    # orders rows of the join
    # ... JOIN lineitem ON o_orderkey = l_orderkey (build: customer-orders)
    rows_l = ctx.scan_select([dev(c["l_shipdate"], abi.KG_DATE)], [(0, abi.KG_GT, d)])
    bk = gather(ctx, dev(c["o_orderkey"], abi.KG_INT), co)
    pk = gather(ctx, dev(c["l_orderkey"], abi.KG_INT), rows_l)
    ht = HashTable(ctx, max(len(bk), 1))
    ht.insert(bk)
    pj, bj = ht.probe(pk)
    ht.free()
    li = rows_l[pj.long()].contiguous()
    orow = co[bj.long()].contiguous()
    k1 = gather(ctx, dev(c["l_orderkey"], abi.KG_INT), li)
    k2 = gather(ctx, odate, orow)
    k3 = gather(ctx, dev(c["o_shippriority"], abi.KG_INT), orow)
    ext = gather(ctx, dev(c["l_extendedprice"], abi.KG_REAL), li)
    disc = gather(ctx, dev(c["l_discount"], abi.KG_REAL), li)
    rev = expr_eval(ctx, len(li), [ext, disc],
                    Mul(ColRef(0, abi.KG_REAL), Sub(Literal(1.0), ColRef(1, abi.KG_REAL))))
    first, vals, _ = ctx.hash_group_agg([k1, k2, k3, Column(abi.KG_REAL, rev.data)], [], [0, 1, 2],
                                        [(abi.KG_AGG_SUM, 3)], max_groups=len(li) + 1)
    frows = torch.from_numpy(first).to(DEV)
    g1, g2, g3 = gather(ctx, k1, frows), gather(ctx, k2, frows), gather(ctx, k3, frows)
    grev = Column(abi.KG_REAL, torch.from_numpy(vals[:, 0].view(np.float64).copy()).to(DEV))
    perm = order_by(ctx, [grev, g2], [False, True]).cpu().numpy()
    a1, a2, a3 = g1.data.cpu().numpy(), g2.data.cpu().numpy(), g3.data.cpu().numpy()
    r = grev.data.cpu().numpy()
    return [(int(a1[i]), float(r[i]), int(a2[i]), int(a3[i])) for i in perm]


def test_tpch_q03():
    ctx = Context(0)
    c = tpch_q03.columns()
    got = device_q03(ctx, c)
    tpch_q03.check(got, kdb.expected("tpch/q03_expected.tbl"))
    want = tpch_q03.oracle_q03(c)
    assert [(k, d, p) for k, _, d, p in got] == [(k, d, p) for k, _, d, p in want]
    assert np.allclose([r for _, r, _, _ in got], [r for _, r, _, _ in want], rtol=1e-12, atol=0)
    ctx.close()


def device_q12(ctx):
    """Q12 on the device: the lineitem filter as one expression over a TEXT and
    three DATE columns (kg_expr_select), the join (kg_ht_build on the selected
    lineitem keys, kg_ht_probe with all orders), the CASE counts over the
    orders' TEXT priority (kg_expr_eval at the joined rows), GROUP BY l_shipmode
    through its string hash (kg_text_hash codes, kg_hash_group_agg), the
    codes checked exact in one pass (kg_text_hash_verify) so that no two
    strings share a group, ORDER BY l_shipmode (kg_order_by over TEXT)."""
    import tpch_q12
    from kushdb_amd.expr import And, Case, Eq, Geq, Lt, Neq, Or, expr_select
    from kushdb_amd.runtime import text_hash
    c = tpch_q12.columns()
    sm = TextColumn.from_strings(c["l_shipmode"], device=DEV)
    pri = TextColumn.from_strings(c["o_orderpriority"], device=DEV)
    commit, receipt, ship = (dev(c[k], abi.KG_DATE) for k in ("l_commitdate", "l_receiptdate", "l_shipdate"))
    T, D = abi.KG_TEXT, abi.KG_DATE
    flt = And([Or([Eq(ColRef(0, T), Literal("RAIL")), Eq(ColRef(0, T), Literal("MAIL"))]),
               Lt(ColRef(1, D), ColRef(2, D)), Lt(ColRef(3, D), ColRef(1, D)),
               Geq(ColRef(2, D), Literal(tpch_q12.LO)), Lt(ColRef(2, D), Literal(tpch_q12.HI))])
    rows_l = expr_select(ctx, len(c["l_orderkey"]), [sm, commit, receipt, ship], flt)
    bk = gather(ctx, dev(c["l_orderkey"], abi.KG_INT), rows_l)
    ht = HashTable(ctx, max(len(bk), 1))
    ht.insert(bk)
    orows, bi = ht.probe(dev(c["o_orderkey"], abi.KG_INT))
    ht.free()
    lrows = rows_l[bi.long()].contiguous()
    codes, _ = text_hash(ctx, sm, lrows)
    from kushdb_amd.runtime import text_hash_verify
    urgent = Or([Eq(ColRef(0, T), Literal("1-URGENT")), Eq(ColRef(0, T), Literal("2-HIGH"))])
    high = expr_eval(ctx, len(c["o_orderkey"]), [pri], Case(urgent, Literal(1), Literal(0)), rows=orows.contiguous())
    low = expr_eval(ctx, len(c["o_orderkey"]), [pri],
                    Case(And([Neq(ColRef(0, T), Literal("1-URGENT")), Neq(ColRef(0, T), Literal("2-HIGH"))]),
                         Literal(1), Literal(0)), rows=orows.contiguous())
    first, vals, _ = ctx.hash_group_agg([Column(abi.KG_BIGINT, codes), high, low], [], [0],
                                        [(abi.KG_AGG_SUM, 1), (abi.KG_AGG_SUM, 2), (abi.KG_AGG_COUNT, -1)],
                                        max_groups=64)
    names = [c["l_shipmode"][int(r)] for r in lrows.cpu().numpy()[first]]
    # every row of a group has the group's string: no two strings share a code
    assert text_hash_verify(ctx, sm, codes, lrows) == 0
    gsm = TextColumn.from_strings(names, device=DEV)
    perm = order_by(ctx, [gsm], [True]).cpu().numpy()
    return [(names[i], int(vals[i, 0]), int(vals[i, 1])) for i in perm]


def test_tpch_q12():
    import tpch_q12
    ctx = Context(0)
    got = device_q12(ctx)
    assert tpch_q12.lines(got) == kdb.expected("tpch/q12_expected.tbl")
    assert got == tpch_q12.oracle_q12(tpch_q12.columns())
    ctx.close()


def _dev_join(ctx, bkeys, pkeys):
    """(probe positions, build positions) of equal keys (kg_ht_build / kg_ht_probe)."""
    ht = HashTable(ctx, max(len(bkeys), 1))
    ht.insert(bkeys)
    pr, br = ht.probe(pkeys)
    ht.free()
    return pr.long(), br.long()


def device_q05(ctx):
    """Q5 on the device: five hash joins over gathered keys, the two-key
    supplier join as a join on s_suppkey plus the checked s_nationkey =
    n_nationkey (kg_expr_select), revenue (kg_expr_eval), GROUP BY n_name
    through its string hash with the exact per-group string check, ORDER BY
    revenue DESC (kg_order_by)."""
    import tpch_q05
    from kushdb_amd.expr import Eq, Geq, expr_select
    from kushdb_amd.runtime import text_hash
    c = tpch_q05.columns()
    I = abi.KG_INT
    col = {k: dev(v, abi.KG_DATE if k == "o_orderdate" else abi.KG_REAL if k.startswith("l_e") or k == "l_discount"
                  else I) for k, v in c.items() if k not in ("r_name", "n_name")}
    rname = TextColumn.from_strings(c["r_name"], device=DEV)
    nname = TextColumn.from_strings(c["n_name"], device=DEV)
    reg = text_select(ctx, rname, abi.KG_TEXT_EQ, "MIDDLE EAST")
    pr, _ = _dev_join(ctx, gather(ctx, col["r_regionkey"], reg), col["n_regionkey"])
    nat = pr.int()
    pr, br = _dev_join(ctx, gather(ctx, col["n_nationkey"], nat), col["c_nationkey"])
    cust, cnat = pr.int(), nat[br]
    ords = ctx.scan_select([col["o_orderdate"]], [(0, abi.KG_GEQ, po.build_date(*tpch_q05.LO)),
                                                  (0, abi.KG_LT, po.build_date(*tpch_q05.HI))])
    pr, br = _dev_join(ctx, gather(ctx, col["c_custkey"], cust), gather(ctx, col["o_custkey"], ords))
    orow, onat = ords[pr], cnat[br]
    pr, br = _dev_join(ctx, gather(ctx, col["o_orderkey"], orow.contiguous()), col["l_orderkey"])
    lrow, lnat = pr.int(), onat[br]
    pr, br = _dev_join(ctx, gather(ctx, col["l_suppkey"], lrow), col["s_suppkey"])
    srow, jl, jn = pr.int(), lrow[br].contiguous(), lnat[br].contiguous()
    sn = gather(ctx, col["s_nationkey"], srow)
    nn = gather(ctx, col["n_nationkey"], jn)
    keep = expr_select(ctx, len(sn), [sn, nn], Eq(ColRef(0, I), ColRef(1, I))).long()
    jl, jn = jl[keep].contiguous(), jn[keep].contiguous()
    ext, disc = gather(ctx, col["l_extendedprice"], jl), gather(ctx, col["l_discount"], jl)
    rev = expr_eval(ctx, len(jl), [ext, disc], Mul(ColRef(0, abi.KG_REAL), Sub(Literal(1.0), ColRef(1, abi.KG_REAL))))
    codes, _ = text_hash(ctx, nname, jn)
    first, vals, _ = ctx.hash_group_agg([Column(abi.KG_BIGINT, codes), Column(abi.KG_REAL, rev.data)], [], [0],
                                        [(abi.KG_AGG_SUM, 1), (abi.KG_AGG_COUNT, -1)], max_groups=64)
    names = [c["n_name"][int(r)] for r in jn.cpu().numpy()[first]]
    for g, s in enumerate(names):
        assert len(text_select(ctx, nname, abi.KG_TEXT_EQ, s, jn)) == vals[g, 1]
    r = vals[:, 0].view(np.float64).copy()
    perm = order_by(ctx, [Column(abi.KG_REAL, torch.from_numpy(r).to(DEV))], [False]).cpu().numpy()
    return [(names[i], float(r[i])) for i in perm]


def test_tpch_q05():
    import tpch_q05
    ctx = Context(0)
    got = device_q05(ctx)
    tpch_q05.check(got, kdb.expected("tpch/q05_expected.tbl"))
    want = tpch_q05.oracle_q05(tpch_q05.columns())
    assert [n for n, _ in got] == [n for n, _ in want]
    assert np.allclose([r for _, r in got], [r for _, r in want], rtol=1e-12, atol=0)
    ctx.close()


def test_tpch_q14():
    """Q14 on the device: scan-select, join with part, the CASE over
    STARTS_WITH(p_type) and the revenue product as expressions at the joined
    rows, two ungrouped SUMs (kg_scan_group_agg), then 100 * a1 / a2."""
    import tpch_q14
    from kushdb_amd.expr import Case, Eq, StartsWith
    c = tpch_q14.columns()
    ctx = Context(0)
    rows = ctx.scan_select([dev(c["l_shipdate"], abi.KG_DATE)],
                           [(0, abi.KG_GEQ, po.build_date(*tpch_q14.LO)), (0, abi.KG_LT, po.build_date(*tpch_q14.HI))])
    pr, br = _dev_join(ctx, gather(ctx, dev(c["l_partkey"], abi.KG_INT), rows), dev(c["p_partkey"], abi.KG_INT))
    li = rows[br].contiguous()
    ext = gather(ctx, dev(c["l_extendedprice"], abi.KG_REAL), li)
    disc = gather(ctx, dev(c["l_discount"], abi.KG_REAL), li)
    R = abi.KG_REAL
    rev_e = Mul(ColRef(0, R), Sub(Literal(1.0), ColRef(1, R)))
    rev = expr_eval(ctx, len(li), [ext, disc], rev_e)
    ptype = TextColumn.from_strings(c["p_type"], device=DEV)
    promo = expr_eval(ctx, len(c["p_partkey"]), [ptype],
                      Case(StartsWith(ColRef(0, abi.KG_TEXT), Literal("PROMO")), Literal(1), Literal(0)),
                      rows=pr.int().contiguous())
    part_rev = expr_eval(ctx, len(li), [ext, disc, Column(abi.KG_INT, promo.data)],
                         Case(Eq(ColRef(2, abi.KG_INT), Literal(1)), rev_e, Literal(0.0)))
    _, vals, _ = ctx.scan_group_agg([Column(R, part_rev.data), Column(R, rev.data)], [], [],
                                    [(abi.KG_AGG_SUM, 0), (abi.KG_AGG_SUM, 1)])
    a1, a2 = vals[0, 0:1].view(np.float64)[0], vals[0, 1:2].view(np.float64)[0]
    got = 100.0 * a1 / a2
    assert abs(got - float(kdb.expected("tpch/q14_expected.tbl")[0])) < 1e-5
    assert abs(got - tpch_q14.oracle_q14(c)) <= 1e-12 * abs(got)
    ctx.close()


def test_tpch_q19():
    """Q19 on the device: the part and lineitem filters as expressions over
    TEXT / INT / REAL columns (kg_expr_select), the join, the post-join
    select pairing each part conjunction (a class code from the conjunctions'
    flag columns at the joined part rows) with its quantity range, and the ungrouped SUM whose
    empty input gives the NULL row."""
    import tpch_q19
    from kushdb_amd.expr import And, Case, Eq, Geq, Leq, Or, expr_select
    c = tpch_q19.columns()
    ctx = Context(0)
    T, I, R = abi.KG_TEXT, abi.KG_INT, abi.KG_REAL
    brand, cont = (TextColumn.from_strings(c[k], device=DEV) for k in ("p_brand", "p_container"))
    size = dev(c["p_size"], I)

    def part_conj(j):
        b, ct, hi = tpch_q19.PART[j]
        return And([Geq(ColRef(2, I), Literal(1)), Leq(ColRef(2, I), Literal(hi)), Eq(ColRef(0, T), Literal(b)),
                    Or([Eq(ColRef(1, T), Literal(x)) for x in ct])])
    # One flag column per conjunction (an expression takes at most 64 nodes),
    # then their OR.
    np_ = len(c["p_partkey"])
    flags = [Column(I, expr_eval(ctx, np_, [brand, cont, size], Case(part_conj(j), Literal(1), Literal(0))).data)
             for j in range(3)]
    prows = expr_select(ctx, np_, flags, Or([Eq(ColRef(j, I), Literal(1)) for j in range(3)]))
    instr, mode = (TextColumn.from_strings(c[k], device=DEV) for k in ("l_shipinstruct", "l_shipmode"))
    qty = dev(c["l_quantity"], R)
    lflt = And([Eq(ColRef(0, T), Literal("DELIVER IN PERSON")),
                Or([Eq(ColRef(1, T), Literal("AIR")), Eq(ColRef(1, T), Literal("AIR REG"))]),
                Or([And([Geq(ColRef(2, R), Literal(lo)), Leq(ColRef(2, R), Literal(hi))]) for lo, hi in tpch_q19.QTY])])
    lrows = expr_select(ctx, len(c["l_partkey"]), [instr, mode, qty], lflt)
    pr, br = _dev_join(ctx, gather(ctx, dev(c["p_partkey"], I), prows), gather(ctx, dev(c["l_partkey"], I), lrows))
    jp, jl = prows[br].contiguous(), lrows[pr].contiguous()
    # class of the joined part row: 1, 2 or 3 for the conjunction it meets (the
    # brands make them exclusive)
    from kushdb_amd.expr import Add
    cls = expr_eval(ctx, np_, flags, Add(Add(ColRef(0, I), Mul(ColRef(1, I), Literal(2))),
                                         Mul(ColRef(2, I), Literal(3))), rows=jp)
    q = gather(ctx, qty, jl)
    keep = expr_select(ctx, len(jl), [Column(I, cls.data), q],
                       Or([And([Eq(ColRef(0, I), Literal(j + 1)), Geq(ColRef(1, R), Literal(lo)),
                                Leq(ColRef(1, R), Literal(hi))]) for j, (lo, hi) in enumerate(tpch_q19.QTY)]))
    kl = jl[keep.long()].contiguous()
    ext, disc = gather(ctx, dev(c["l_extendedprice"], R), kl), gather(ctx, dev(c["l_discount"], R), kl)
    rev = expr_eval(ctx, len(kl), [ext, disc], Mul(ColRef(0, R), Sub(Literal(1.0), ColRef(1, R))))
    codes, vals, nulls = ctx.scan_group_agg([Column(R, rev.data)], [], [], [(abi.KG_AGG_SUM, 0)])
    got = None if (len(codes) == 0 or nulls[0, 0]) else float(vals[0, 0:1].view(np.float64)[0])
    assert tpch_q19.line(got) == kdb.expected("tpch/q19_expected.tbl")[0]
    assert got == tpch_q19.oracle_q19(c)
    assert len(pr) > 0   # the joins produce pairs; only the post-join select empties them
    ctx.close()


def test_tpch_q18():
    """Q18's subquery on the device: GROUP BY l_orderkey SUM(l_quantity) over
    all of lineitem (kg_hash_group_agg, thousands of groups) equal to the
    oracle group for group, then the HAVING select (kg_scan_select on the
    sums), whose empty result empties the joins downstream (a build on no rows,
    probed by all orders) and the output."""
    import tpch_q18
    c = tpch_q18.columns()
    ctx = Context(0)
    ok_col, qty = dev(c["l_orderkey"], abi.KG_INT), dev(c["l_quantity"], abi.KG_REAL)
    n = len(c["l_orderkey"])
    first, vals, _ = ctx.hash_group_agg([ok_col, qty], [], [0], [(abi.KG_AGG_SUM, 1)], max_groups=n + 1)
    keys, sums = tpch_q18.oracle_subquery(c)
    assert np.array_equal(c["l_orderkey"][first], keys)
    got = vals[:, 0].view(np.float64)
    assert np.all(np.abs(got - sums) <= 1e-12 * np.abs(sums))
    s = Column(abi.KG_REAL, torch.from_numpy(got.copy()).to(DEV))
    having = ctx.scan_select([s], [(0, abi.KG_GT, tpch_q18.BAR)])
    hk = gather(ctx, Column(abi.KG_INT, torch.from_numpy(keys.astype(np.int32)).to(DEV)), having)
    orders = dev(kdb.col("o_orderkey", "int")[0], abi.KG_INT)
    pr, _ = _dev_join(ctx, hk, orders)
    assert len(having) == 0 and len(pr) == 0 and kdb.expected("tpch/q18_expected.tbl") == []
    ctx.close()


def test_tpch_q10():
    """Q10 on the device: the orders date range, three hash joins, the TEXT
    l_returnflag select (kg_text_select), revenue, GROUP BY the seven customer
    and nation columns and ORDER BY revenue DESC. Every grouping column is an
    attribute of the joined customer row (c_custkey is unique, checked here)
    or of its nation, so the device groups by (c_custkey, c_acctbal,
    n_nationkey) (kg_hash_group_agg): the same groups as the seven values."""
    import tpch_q10
    c = tpch_q10.columns()
    assert len(np.unique(c["c_custkey"])) == len(c["c_custkey"])
    ctx = Context(0)
    I, R = abi.KG_INT, abi.KG_REAL
    ords = ctx.scan_select([dev(c["o_orderdate"], abi.KG_DATE)],
                           [(0, abi.KG_GEQ, po.build_date(*tpch_q10.LO)), (0, abi.KG_LT, po.build_date(*tpch_q10.HI))])
    pr, br = _dev_join(ctx, gather(ctx, dev(c["o_custkey"], I), ords), dev(c["c_custkey"], I))
    crow, orow = pr.int(), ords[br]
    pr, br = _dev_join(ctx, dev(c["n_nationkey"], I), gather(ctx, dev(c["c_nationkey"], I), crow))
    crow, orow, nrow = crow[pr].contiguous(), orow[pr].contiguous(), br.int().contiguous()
    li = text_select(ctx, TextColumn.from_strings(c["l_returnflag"], device=DEV), abi.KG_TEXT_EQ, "R")
    pr, br = _dev_join(ctx, gather(ctx, dev(c["o_orderkey"], I), orow), gather(ctx, dev(c["l_orderkey"], I), li))
    lrow, crow, nrow = li[pr].contiguous(), crow[br].contiguous(), nrow[br].contiguous()
    ext, disc = gather(ctx, dev(c["l_extendedprice"], R), lrow), gather(ctx, dev(c["l_discount"], R), lrow)
    rev = expr_eval(ctx, len(lrow), [ext, disc], Mul(ColRef(0, R), Sub(Literal(1.0), ColRef(1, R))))
    keys = [gather(ctx, dev(c["c_custkey"], I), crow), gather(ctx, dev(c["c_acctbal"], R), crow),
            gather(ctx, dev(c["n_nationkey"], I), nrow)]
    first, vals, _ = ctx.hash_group_agg(keys + [Column(R, rev.data)], [], [0, 1, 2], [(abi.KG_AGG_SUM, 3)],
                                        max_groups=len(lrow) + 1)
    r = vals[:, 0].view(np.float64).copy()
    perm = order_by(ctx, [Column(R, torch.from_numpy(r).to(DEV))], [False]).cpu().numpy()
    fc, fn = crow.cpu().numpy()[first], nrow.cpu().numpy()[first]
    got = [(int(c["c_custkey"][fc[i]]), c["c_name"][fc[i]], float(r[i]), float(c["c_acctbal"][fc[i]]),
            c["n_name"][fn[i]], c["c_address"][fc[i]], c["c_phone"][fc[i]], c["c_comment"][fc[i]]) for i in perm]
    tpch_q10.check(got, kdb.expected("tpch/q10_expected.tbl"))
    want = tpch_q10.oracle_q10(c)
    assert [g[:2] + g[3:] for g in got] == [w[:2] + w[3:] for w in want]
    assert np.allclose([g[2] for g in got], [w[2] for w in want], rtol=1e-12, atol=0)
    ctx.close()


def test_tpch_q11():
    """Q11 on the device: TEXT select on nation, two joins, the subquery's
    SUM(ps_supplycost * Float(ps_availqty)) as an ungrouped product aggregate
    (kg_scan_group_agg with a second factor), the same product grouped by
    ps_partkey (kg_hash_group_agg), the cross-product select value > value1
    (kg_scan_select against the subquery's value), ORDER BY value DESC."""
    import tpch_q11
    c = tpch_q11.columns()
    ctx = Context(0)
    I, R = abi.KG_INT, abi.KG_REAL
    nat = text_select(ctx, TextColumn.from_strings(c["n_name"], device=DEV), abi.KG_TEXT_EQ, "GERMANY")
    pr, _ = _dev_join(ctx, gather(ctx, dev(c["n_nationkey"], I), nat), dev(c["s_nationkey"], I))
    sup = pr.int().contiguous()
    pr, _ = _dev_join(ctx, gather(ctx, dev(c["s_suppkey"], I), sup), dev(c["ps_suppkey"], I))
    ps = pr.int().contiguous()
    cost, qty, part = (gather(ctx, dev(c[k], t), ps) for k, t in
                       (("ps_supplycost", R), ("ps_availqty", I), ("ps_partkey", I)))
    _, tot, _ = ctx.scan_group_agg([cost, qty], [], [], [(abi.KG_AGG_SUM, 0, 1)])
    bar = float(tot[0, 0:1].view(np.float64)[0]) * 0.0001
    first, vals, _ = ctx.hash_group_agg([part, cost, qty], [], [0], [(abi.KG_AGG_SUM, 1, 2)],
                                        max_groups=len(ps) + 1)
    value = Column(R, torch.from_numpy(vals[:, 0].view(np.float64).copy()).to(DEV))
    keep = ctx.scan_select([value], [(0, abi.KG_GT, bar)]).long()
    kv = value.data[keep].contiguous()
    perm = order_by(ctx, [Column(R, kv)], [False]).cpu().numpy()
    keys = c["ps_partkey"][ps.cpu().numpy()[first]][keep.cpu().numpy()]
    got = [(int(keys[i]), float(kv[i])) for i in perm]
    tpch_q11.check(got, kdb.expected("tpch/q11_expected.tbl"))
    want = tpch_q11.oracle_q11(c)
    assert [k for k, _ in got] == [k for k, _ in want]
    assert np.allclose([v for _, v in got], [v for _, v in want], rtol=1e-12, atol=0)
    ctx.close()


def test_tpch_q07():
    """Q7 on the device: the nation select (TEXT), five joins, the post-join
    select over the two nations through a per-nation code (CASE on n_name,
    EGYPT 1 / KENYA 2, gathered at both nation rows), EXTRACT(YEAR) and
    revenue as expressions, GROUP BY (code1, code2, year) over hashed keys,
    ORDER BY the three (the codes order as the names do)."""
    import tpch_q07
    from kushdb_amd.expr import And, Case, Eq, ExtractYear, Or, expr_select
    c = tpch_q07.columns()
    ctx = Context(0)
    I, R, D, T = abi.KG_INT, abi.KG_REAL, abi.KG_DATE, abi.KG_TEXT
    names = TextColumn.from_strings(c["n_name"], device=DEV)
    code = expr_eval(ctx, len(c["n_name"]), [names],
                     Case(Eq(ColRef(0, T), Literal("EGYPT")), Literal(1),
                          Case(Eq(ColRef(0, T), Literal("KENYA")), Literal(2), Literal(0))))
    codecol = Column(I, code.data)
    nat = ctx.scan_select([codecol], [(0, abi.KG_GT, 0)])
    pr, br = _dev_join(ctx, gather(ctx, dev(c["n_nationkey"], I), nat), dev(c["s_nationkey"], I))
    srow, snat = pr.int(), nat[br]
    pr, br = _dev_join(ctx, gather(ctx, dev(c["n_nationkey"], I), nat), dev(c["c_nationkey"], I))
    crow, cnat = pr.int(), nat[br]
    pr, br = _dev_join(ctx, gather(ctx, dev(c["c_custkey"], I), crow), dev(c["o_custkey"], I))
    orow, onat = pr.int(), cnat[br]
    ship = dev(c["l_shipdate"], D)
    li = ctx.scan_select([ship], [(0, abi.KG_GEQ, po.build_date(*tpch_q07.LO)),
                                  (0, abi.KG_LEQ, po.build_date(*tpch_q07.HI))])
    pr, br = _dev_join(ctx, gather(ctx, dev(c["o_orderkey"], I), orow), gather(ctx, dev(c["l_orderkey"], I), li))
    lrow, lnat = li[pr], onat[br]
    pr, br = _dev_join(ctx, gather(ctx, dev(c["s_suppkey"], I), srow), gather(ctx, dev(c["l_suppkey"], I),
                                                                               lrow.contiguous()))
    lrow, n2, n1 = lrow[pr].contiguous(), lnat[pr].contiguous(), snat[br].contiguous()
    c1, c2 = gather(ctx, codecol, n1), gather(ctx, codecol, n2)
    keep = expr_select(ctx, len(lrow), [c1, c2],
                       Or([And([Eq(ColRef(0, I), Literal(1)), Eq(ColRef(1, I), Literal(2))]),
                           And([Eq(ColRef(0, I), Literal(2)), Eq(ColRef(1, I), Literal(1))])])).long()
    lrow = lrow[keep].contiguous()
    c1, c2 = Column(I, c1.data[keep].contiguous()), Column(I, c2.data[keep].contiguous())
    sd, ext, disc = (gather(ctx, dev(c[k], t), lrow) for k, t in
                     (("l_shipdate", D), ("l_extendedprice", R), ("l_discount", R)))
    year = expr_eval(ctx, len(lrow), [sd], ExtractYear(ColRef(0, D)))
    rev = expr_eval(ctx, len(lrow), [ext, disc], Mul(ColRef(0, R), Sub(Literal(1.0), ColRef(1, R))))
    yc = Column(I, year.data)
    first, vals, _ = ctx.hash_group_agg([c1, c2, yc, Column(R, rev.data)], [], [0, 1, 2], [(abi.KG_AGG_SUM, 3)],
                                        max_groups=64)
    fr = torch.from_numpy(first).to(DEV)
    g1, g2, g3 = gather(ctx, c1, fr), gather(ctx, c2, fr), gather(ctx, yc, fr)
    perm = order_by(ctx, [g1, g2, g3], [True, True, True]).cpu().numpy()
    a1, a2, a3 = g1.data.cpu().numpy(), g2.data.cpu().numpy(), g3.data.cpu().numpy()
    r = vals[:, 0].view(np.float64)
    nm = {1: "EGYPT", 2: "KENYA"}
    got = [(nm[int(a1[i])], nm[int(a2[i])], int(a3[i]), float(r[i])) for i in perm]
    tpch_q07.check(got, kdb.expected("tpch/q07_expected.tbl"))
    want = tpch_q07.oracle_q07(c)
    assert [g[:3] for g in got] == [w[:3] for w in want]
    assert np.allclose([g[3] for g in got], [w[3] for w in want], rtol=1e-12, atol=0)
    ctx.close()


def test_tpch_q09():
    """Q9 on the device: CONTAINS on p_name (kg_text_select), five joins (the
    two-key lineitem join as l_partkey plus a checked l_suppkey), EXTRACT(YEAR)
    and the profit expression, GROUP BY (nation row, year) over hashed keys
    (nation names are unique per nation row), ORDER BY nation (TEXT), year DESC."""
    import tpch_q09
    from kushdb_amd.expr import Eq, ExtractYear, expr_select
    c = tpch_q09.columns()
    ctx = Context(0)
    I, R, D = abi.KG_INT, abi.KG_REAL, abi.KG_DATE
    assert len(set(c["n_name"])) == len(c["n_name"])
    pr, br = _dev_join(ctx, dev(c["n_nationkey"], I), dev(c["s_nationkey"], I))
    srow, snat = pr.int(), br.int()
    part = text_select(ctx, TextColumn.from_strings(c["p_name"], device=DEV), abi.KG_TEXT_CONTAINS, "moccasin")
    pr, br = _dev_join(ctx, gather(ctx, dev(c["p_partkey"], I), part), dev(c["ps_partkey"], I))
    psrow, prow = pr.int(), part[br]
    pr, br = _dev_join(ctx, gather(ctx, dev(c["s_suppkey"], I), srow), gather(ctx, dev(c["ps_suppkey"], I), psrow))
    psrow, prow, srow2, nrow = psrow[pr], prow[pr].contiguous(), srow[br], snat[br]
    pr, br = _dev_join(ctx, gather(ctx, dev(c["p_partkey"], I), prow), dev(c["l_partkey"], I))
    ls = gather(ctx, dev(c["l_suppkey"], I), pr.int().contiguous())
    ss = gather(ctx, dev(c["s_suppkey"], I), srow2[br].contiguous())
    keep = expr_select(ctx, len(ls), [ls, ss], Eq(ColRef(0, I), ColRef(1, I))).long()
    lrow, psrow, nrow = pr.int()[keep].contiguous(), psrow[br][keep].contiguous(), nrow[br][keep].contiguous()
    pr, br = _dev_join(ctx, gather(ctx, dev(c["l_orderkey"], I), lrow), dev(c["o_orderkey"], I))
    orow, lrow, psrow, nrow = pr.int(), lrow[br].contiguous(), psrow[br].contiguous(), nrow[br].contiguous()
    ext, disc, qty = (gather(ctx, dev(c[k], R), lrow) for k in ("l_extendedprice", "l_discount", "l_quantity"))
    cost = gather(ctx, dev(c["ps_supplycost"], R), psrow)
    amount = expr_eval(ctx, len(lrow), [ext, disc, cost, qty],
                       Sub(Mul(ColRef(0, R), Sub(Literal(1.0), ColRef(1, R))), Mul(ColRef(2, R), ColRef(3, R))))
    year = expr_eval(ctx, len(orow), [gather(ctx, dev(c["o_orderdate"], D), orow)], ExtractYear(ColRef(0, D)))
    nk, yk = Column(I, nrow), Column(I, year.data)
    first, vals, _ = ctx.hash_group_agg([nk, yk, Column(R, amount.data)], [], [0, 1], [(abi.KG_AGG_SUM, 2)],
                                        max_groups=1024)
    fr = torch.from_numpy(first).to(DEV)
    gn, gy = gather(ctx, nk, fr).data.cpu().numpy(), gather(ctx, yk, fr).data.cpu().numpy()
    gnames = [c["n_name"][int(n)] for n in gn]
    perm = order_by(ctx, [TextColumn.from_strings(gnames, device=DEV), Column(I, torch.from_numpy(gy).to(DEV))],
                    [True, False]).cpu().numpy()
    r = vals[:, 0].view(np.float64)
    got = [(gnames[i], int(gy[i]), float(r[i])) for i in perm]
    tpch_q09.check(got, kdb.expected("tpch/q09_expected.tbl"))
    want = tpch_q09.oracle_q09(c)
    assert [g[:2] for g in got] == [w[:2] for w in want]
    assert np.allclose([g[2] for g in got], [w[2] for w in want], rtol=1e-12, atol=1e-9)
    ctx.close()


def test_tpch_q18_lower_bar():
    """Q18's whole plan with a HAVING bar low enough that about 60 orders
    pass: the subquery GROUP BY and HAVING select, the joins with orders,
    customer and lineitem again, the GROUP BY over (c_name, c_custkey,
    o_orderkey, o_orderdate, o_totalprice) (c_name is an attribute of the
    uniquely keyed customer row, so the device groups by the other four) with
    SUM(l_quantity), and ORDER BY o_totalprice DESC, o_orderdate, against the
    oracle's restatement of the plan."""
    import tpch_q18
    c = tpch_q18.full_columns()
    assert len(np.unique(c["c_custkey"])) == len(c["c_custkey"])
    assert len(np.unique(c["o_orderkey"])) == len(c["o_orderkey"])
    keys, sums = tpch_q18.oracle_subquery(c)
    bar = float(np.sort(sums)[-60])
    want = tpch_q18.oracle_rows(c, bar)
    assert 40 <= len(want) <= 60
    ctx = Context(0)
    I, R = abi.KG_INT, abi.KG_REAL
    n = len(c["l_orderkey"])
    first, vals, _ = ctx.hash_group_agg([dev(c["l_orderkey"], I), dev(c["l_quantity"], R)], [], [0],
                                        [(abi.KG_AGG_SUM, 1)], max_groups=n + 1)
    got_sums = vals[:, 0].view(np.float64).copy()
    s = Column(R, torch.from_numpy(got_sums).to(DEV))
    having = ctx.scan_select([s], [(0, abi.KG_GT, bar)])
    hk = gather(ctx, Column(I, torch.from_numpy(c["l_orderkey"][first].astype(np.int32)).to(DEV)), having)
    print(f"q18 bar {bar}: {len(having)} orders above it", flush=True)
    # subquery JOIN orders ON l_orderkey = o_orderkey
    orow, _ = _dev_join(ctx, hk, dev(c["o_orderkey"], I))
    # JOIN customer ON o_custkey = c_custkey
    pr, crow = _dev_join(ctx, dev(c["c_custkey"], I), gather(ctx, dev(c["o_custkey"], I), orow.int()))
    orow = orow[pr].contiguous()
    # JOIN lineitem ON o_orderkey = l_orderkey
    lrow, br = _dev_join(ctx, gather(ctx, dev(c["o_orderkey"], I), orow.int()), dev(c["l_orderkey"], I))
    orow, crow = orow[br].int().contiguous(), crow[br].int().contiguous()
    gcols = [gather(ctx, dev(c["c_custkey"], I), crow), gather(ctx, dev(c["o_orderkey"], I), orow),
             gather(ctx, dev(c["o_orderdate"], abi.KG_DATE), orow), gather(ctx, dev(c["o_totalprice"], R), orow),
             gather(ctx, dev(c["l_quantity"], R), lrow.int())]
    gfirst, gvals, _ = ctx.hash_group_agg(gcols, [], [0, 1, 2, 3], [(abi.KG_AGG_SUM, 4)], max_groups=len(lrow) + 1)
    fo, fc = orow.cpu().numpy()[gfirst], crow.cpu().numpy()[gfirst]
    tp = Column(R, torch.from_numpy(c["o_totalprice"][fo].copy()).to(DEV))
    od = Column(abi.KG_DATE, torch.from_numpy(c["o_orderdate"][fo].copy()).to(DEV))
    perm = order_by(ctx, [tp, od], [False, True]).cpu().numpy()
    rs = gvals[:, 0].view(np.float64)
    got = [(c["c_name"][fc[i]], int(c["c_custkey"][fc[i]]), int(c["o_orderkey"][fo[i]]), int(c["o_orderdate"][fo[i]]),
            float(c["o_totalprice"][fo[i]]), float(rs[i])) for i in perm]
    ctx.close()
    assert [g[:5] for g in got] == [w[:5] for w in want]
    assert all(abs(g[5] - w[5]) <= 1e-12 * abs(w[5]) for g, w in zip(got, want))
