"""Multi-way chain join the way Skinner's per-episode work runs it, on the
device: each child filtered to a selection (ScanSelect / TEXT predicates), a
column index on the join column of the next table over its selection
(MemoryColumnIndex), index-nested-loop expansion one table at a time
(permutable_skinner_join_translator.cc:359-433), result tuples deduplicated in
a tuple-id set (the valid-tuple handler, :842-912), then MIN aggregates over the
joined rows. Tuple ids are positions in the filtered children, as in the
reference. The join order is given (the UCT search that picks it stays on the
host). Tables: five chained tables t0..t4 where t_i.fk references
t_{i-1}.id (kushdb_amd.workloads.skinner_tables). Every step is a library call;
nothing here runs a torch kernel (tensors only hold the device buffers)."""
from . import abi
from .runtime import Column, ColumnIndex, TupleSet, gather, index_join_expand, text_minmax, text_select, tuple_rows


def selected(tab, sel):
    """Number of tuples of a table under its selection (None = every row)."""
    return len(tab["id"]) if sel is None else int(sel.numel())


def chain_filters(ctx, tabs, spec):
    """Selection of every table: ascending row ids (int32), or None when the
    table has no filter (every row, tuple id = row id)."""
    sels = []
    for t, f in zip(tabs, spec):
        rows = None
        for kind, col, op, lit in f:
            if kind == "text":
                rows = text_select(ctx, t[col], op, lit, rows)
            elif rows is None:
                rows = ctx.scan_select([t[col]], [(0, op, lit)])
            else:
                # Filter the selected rows, then map the survivors' positions
                # back to row ids (a gather of the selection itself).
                keep = ctx.scan_select([gather(ctx, t[col], rows)], [(0, op, lit)])
                rows = gather(ctx, Column(abi.KG_INT, rows), keep).data
        sels.append(rows)
    return sels


def chain_join(ctx, tabs, sels, order):
    """Join along the chain in `order` (each next table adjacent to a joined
    one). Returns ([n, ntables] tuple ids (positions in each selection) with
    columns in join order, the table of each column (= order), and the number
    of index probes and inserts made)."""
    first = order[0]
    cur = None                     # the identity tuples (0), (1), ... of the first table
    ncur = selected(tabs[first], sels[first])
    cols = [first]
    probes = inserts = 0
    for nxt in order[1:]:
        if nxt - 1 in cols:      # nxt.fk = (nxt-1).id: probe with (nxt-1).id, index on nxt.fk
            src, src_key, idx_col = nxt - 1, tabs[nxt - 1]["id"], tabs[nxt]["fk"]
        else:                    # (nxt+1).fk = nxt.id: probe with (nxt+1).fk, index on nxt.id
            src, src_key, idx_col = nxt + 1, tabs[nxt + 1]["fk"], tabs[nxt]["id"]
        idx = ColumnIndex(ctx, idx_col, sels[nxt])
        inserts += selected(tabs[nxt], sels[nxt])
        probes += ncur
        cur = index_join_expand(ctx, cur, cols.index(src), src_key, sels[src], idx, n=ncur)
        ncur = cur.shape[0]
        idx.free()
        cols.append(nxt)
    return cur, cols, probes, inserts


def chain_query(ctx, tabs, spec, order, text_min=(0, "name"), int_min=(4, "v")):
    """Filters, chain join, dedup and MIN(t0.name), MIN(t4.v). Returns
    (distinct tuples [n, 5] with columns in join order — table order for the
    order [0, 1, 2, 3, 4]; stats["columns"] names the table of each —, MIN text
    bytes or None, MIN int or None, stats)."""
    sels = chain_filters(ctx, tabs, spec)
    res, cols, probes, inserts = chain_join(ctx, tabs, sels, order)
    ts = TupleSet(ctx, res.shape[1], max(res.shape[0], 1))
    ts.insert(res)
    tuples = ts.extract()
    min_text = min_int = None
    if tuples.shape[0]:
        ti, col = text_min
        _, min_text = text_minmax(ctx, tabs[ti][col], tuple_rows(ctx, tuples, cols.index(ti), sels[ti]))
        ii, icol = int_min
        v = gather(ctx, tabs[ii][icol], tuple_rows(ctx, tuples, cols.index(ii), sels[ii]))
        _, vals, _ = ctx.scan_group_agg([v], [], [], [(abi.KG_AGG_MIN, 0)])
        min_int = int(vals[0, 0])
    ts.free()
    stats = {"selected": [selected(t, s) for t, s in zip(tabs, sels)], "index_probes": probes,
             "index_inserts": inserts, "tuples": int(res.shape[0]), "columns": cols}
    return tuples, min_text, min_int, stats


def skinner_chain_query(ctx, tabs, spec, budget=0, seed=1, text_min=(0, "name"), int_min=(4, "v")):
    """The same query with the join order left to the Skinner driver
    (kg_skinner_join: UCT-chosen orders per episode, device episodes, tuple-id
    dedup), as ExecutePermutableSkinnerJoin runs it
    (runtime/skinner_join_executor.cc:840-885). Returns (distinct tuples [n, 5]
    in table order, MIN text bytes or None, MIN int or None, stats)."""
    from .runtime import skinner_join
    sels = chain_filters(ctx, tabs, spec)
    tables = [(selected(t, s), s) for t, s in zip(tabs, sels)]
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
    stats["selected"] = [selected(t, s) for t, s in zip(tabs, sels)]
    return tuples, min_text, min_int, stats
