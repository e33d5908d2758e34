"""Multi-way chain join the way Skinner's per-episode work runs it, on the
device: each child filtered to a selection (ScanSelect / TEXT predicates), a
column index on the join column of the next table over its selection
(MemoryColumnIndex), index-nested-loop expansion one table at a time
(permutable_skinner_join_translator.cc:359-433), result tuples deduplicated in
a tuple-id set (the valid-tuple handler, :842-912), then MIN aggregates over the
joined rows. Tuple ids are positions in the filtered children, as in the
reference. The join order is given (the UCT search that picks it stays on the
host). Tables: five chained tables t0..t4 where t_i.fk references
t_{i-1}.id (kushdb_amd.workloads.skinner_tables)."""
import torch

from . import abi
from .runtime import ColumnIndex, TupleSet, gather, index_join_expand, text_minmax, text_select, tuple_rows


def chain_filters(ctx, tabs, spec):
    """Selection (ascending row ids, int32) of every table."""
    sels = []
    for t, f in zip(tabs, spec):
        rows = None
        for kind, col, op, lit in f:
            if kind == "text":
                rows = text_select(ctx, t[col], op, lit, rows)
            else:
                c = t[col]
                if rows is None:
                    rows = ctx.scan_select([c], [(0, op, lit)])
                else:
                    g = gather(ctx, c, rows)
                    keep = ctx.scan_select([g], [(0, op, lit)])
                    rows = rows[keep.long()]
        if rows is None:
            n = len(t["id"])
            rows = torch.arange(n, dtype=torch.int32, device=t["id"].data.device)
        sels.append(rows.contiguous())
    return sels


def chain_join(ctx, tabs, sels, order):
    """Join along the chain in `order` (each next table adjacent to a joined
    one). Returns [n, ntables] tuple ids (positions in each selection), columns
    in table order, and the number of index probes and inserts made."""
    ntab = len(tabs)
    first = order[0]
    dev = sels[first].device
    cur = torch.arange(sels[first].numel(), dtype=torch.int32, device=dev).view(-1, 1)
    cols = [first]
    probes = inserts = 0
    for nxt in order[1:]:
        if nxt - 1 in cols:      # nxt.fk = (nxt-1).id: probe with (nxt-1).id, index on nxt.fk
            src, src_key, idx_col = nxt - 1, tabs[nxt - 1]["id"], tabs[nxt]["fk"]
        else:                    # (nxt+1).fk = nxt.id: probe with (nxt+1).fk, index on nxt.id
            src, src_key, idx_col = nxt + 1, tabs[nxt + 1]["fk"], tabs[nxt]["id"]
        idx = ColumnIndex(ctx, idx_col, sels[nxt])
        inserts += sels[nxt].numel()
        probes += cur.shape[0]
        cur = index_join_expand(ctx, cur.contiguous(), cols.index(src), src_key, sels[src], idx)
        idx.free()
        cols.append(nxt)
    perm = [cols.index(t) for t in range(ntab)]
    return cur[:, perm].contiguous(), probes, inserts


def chain_query(ctx, tabs, spec, order, text_min=(0, "name"), int_min=(4, "v")):
    """Filters, chain join, dedup and MIN(t0.name), MIN(t4.v). Returns
    (distinct tuples [n, 5], MIN text bytes or None, MIN int or None, stats)."""
    sels = chain_filters(ctx, tabs, spec)
    res, probes, inserts = chain_join(ctx, tabs, sels, order)
    ts = TupleSet(ctx, res.shape[1], max(res.shape[0], 1))
    ts.insert(res)
    tuples = ts.extract()
    min_text = min_int = None
    if tuples.shape[0]:
        ti, col = text_min
        _, min_text = text_minmax(ctx, tabs[ti][col], tuple_rows(ctx, tuples, ti, sels[ti]))
        ii, icol = int_min
        v = gather(ctx, tabs[ii][icol], tuple_rows(ctx, tuples, ii, sels[ii]))
        _, vals, _ = ctx.scan_group_agg([v], [], [], [(abi.KG_AGG_MIN, 0)])
        min_int = int(vals[0, 0])
    ts.free()
    stats = {"selected": [int(s.numel()) for s in sels], "index_probes": probes, "index_inserts": inserts,
             "tuples": int(res.shape[0])}
    return tuples, min_text, min_int, stats


def skinner_chain_query(ctx, tabs, spec, budget=0, seed=1, text_min=(0, "name"), int_min=(4, "v")):
    """The same query with the join order left to the Skinner driver
    (kg_skinner_join: UCT-chosen orders per episode, device episodes, tuple-id
    dedup), as ExecutePermutableSkinnerJoin runs it
    (runtime/skinner_join_executor.cc:840-885). Returns (distinct tuples [n, 5],
    MIN text bytes or None, MIN int or None, stats)."""
    from .runtime import skinner_join
    sels = chain_filters(ctx, tabs, spec)
    tables = [(s.numel(), s) for s in sels]
    conds = [(t - 1, tabs[t - 1]["id"], t, tabs[t]["fk"]) for t in range(1, len(tabs))]
    tuples, stats = skinner_join(ctx, tables, conds, budget=budget, seed=seed)
    min_text = min_int = None
    if tuples.shape[0]:
        ti, col = text_min
        _, min_text = text_minmax(ctx, tabs[ti][col], tuple_rows(ctx, tuples, ti, sels[ti]))
        ii, icol = int_min
        v = gather(ctx, tabs[ii][icol], tuple_rows(ctx, tuples, ii, sels[ii]))
        _, vals, _ = ctx.scan_group_agg([v], [], [], [(abi.KG_AGG_MIN, 0)])
        min_int = int(vals[0, 0])
    stats["selected"] = [int(s.numel()) for s in sels]
    return tuples, min_text, min_int, stats
