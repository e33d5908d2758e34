"""Seeded synthetic tables for the star, scan-aggregate and join workloads.

Generated on the device with torch's Philox generator (deterministic for a
seed); the CPU checker receives copies of the same tensors.
"""
import datetime

import torch

SEED = 0x4B555348


def julian_day(y, m, d):
    """Julian day number of a civil date (the integer DATE encoding of
    runtime/date.cc:10-25; equal to the proleptic Gregorian ordinal + 1721425)."""
    return datetime.date(y, m, d).toordinal() + 1721425


DAY0 = julian_day(1992, 1, 1)
NDAYS = 2557  # 1992-01-01 .. 1998-12-31


def _gen(device, seed):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def scan_agg_table(n, device="cuda", seed=SEED + 1):
    """60M-row shape: two enum group columns (3 and 2 codes), four float64
    measures U[0,1e5), three int32 measures U{1..50}, day uniform over 2557
    days. The filter day <= cutoff keeps about 98% of rows."""
    g = _gen(device, seed)
    t = {
        "day": torch.randint(DAY0, DAY0 + NDAYS, (n,), generator=g, device=device, dtype=torch.int32),
        "g1": torch.randint(0, 3, (n,), generator=g, device=device, dtype=torch.int32),
        "g2": torch.randint(0, 2, (n,), generator=g, device=device, dtype=torch.int32),
    }
    for i in range(4):
        t[f"m{i + 1}"] = torch.rand((n,), generator=g, device=device, dtype=torch.float64) * 1e5
    for i in range(3):
        t[f"i{i + 1}"] = torch.randint(1, 51, (n,), generator=g, device=device, dtype=torch.int32)
    t["cutoff"] = DAY0 + int(0.98 * NDAYS) - 1
    return t


def star_tables(nfact=1_000_000, ndim=10_000, device="cuda", seed=SEED):
    """Fact: fk U{0..ndim-1}, qty U{1..50}, price U[1,1e5), day over 2557 days.
    Dim: pk 0..ndim-1, cat U over 25 codes. Filter: a 30% day range."""
    g = _gen(device, seed)
    f = {
        "fk": torch.randint(0, ndim, (nfact,), generator=g, device=device, dtype=torch.int32),
        "qty": torch.randint(1, 51, (nfact,), generator=g, device=device, dtype=torch.int32),
        "price": 1.0 + torch.rand((nfact,), generator=g, device=device, dtype=torch.float64) * (1e5 - 1.0),
        "day": torch.randint(DAY0, DAY0 + NDAYS, (nfact,), generator=g, device=device, dtype=torch.int32),
    }
    d = {
        "pk": torch.arange(ndim, device=device, dtype=torch.int32),
        "cat": torch.randint(0, 25, (ndim,), generator=g, device=device, dtype=torch.int32),
    }
    lo = DAY0 + 600
    return f, d, (lo, lo + int(0.3 * NDAYS))


def zipf_ranks(n, nkeys, theta, g, device):
    """n ranks in [0, nkeys) with P(rank = i) proportional to 1 / (i+1)^theta
    (inverse-CDF sampling; theta = 0 is uniform)."""
    if theta == 0:
        return torch.randint(0, nkeys, (n,), generator=g, device=device, dtype=torch.int64)
    w = torch.arange(1, nkeys + 1, device=device, dtype=torch.float64).pow_(-theta)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    out = torch.empty(n, dtype=torch.int64, device=device)
    chunk = 1 << 26
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        u = torch.rand((e - s,), generator=g, device=device, dtype=torch.float64)
        out[s:e] = torch.searchsorted(cdf, u).clamp_(max=nkeys - 1)
    return out


def join_tables(nbuild=1 << 24, nprobe=1 << 28, theta=0.0, miss=0.1, device="cuda", seed=SEED + 2):
    """Build: nbuild unique int64 keys (a random permutation of 0..nbuild-1)
    with an int32 payload. Probe: nprobe int64 keys drawn Zipf(theta) over the
    build keys, a `miss` fraction replaced by keys >= nbuild, and an int64
    value U[0, 2^32)."""
    g = _gen(device, seed)
    bkeys = torch.randperm(nbuild, generator=g, device=device, dtype=torch.int64)
    payload = torch.randint(-2**31, 2**31 - 1, (nbuild,), generator=g, device=device, dtype=torch.int32)
    ranks = zipf_ranks(nprobe, nbuild, theta, g, device)
    pkeys = bkeys[ranks]
    del ranks
    missmask = torch.rand((nprobe,), generator=g, device=device) < miss
    pkeys[missmask] = nbuild + torch.randint(0, 1 << 40, (int(missmask.sum().item()),), generator=g, device=device,
                                             dtype=torch.int64)
    del missmask
    pvals = torch.randint(0, 1 << 32, (nprobe,), generator=g, device=device, dtype=torch.int64)
    return bkeys, payload, pkeys, pvals


SKINNER_SIZES = (200_000, 1_000_000, 4_000_000, 25_000_000, 50_000_000)


def vocabulary(nwords=50_000, seed=SEED + 3):
    """nwords distinct lowercase ASCII words of 3..10 letters."""
    import numpy as np
    rng = np.random.default_rng(seed)
    words = set()
    while len(words) < nwords:
        ln = int(rng.integers(3, 11))
        words.add("".join(chr(97 + int(c)) for c in rng.integers(0, 26, ln)))
    return sorted(words)


def text_column(n, vocab, rng):
    """n strings of 8..40 chars: vocabulary words joined by spaces. Returns
    (offsets int64 [n+1], heap uint8, python list)."""
    import numpy as np
    out = []
    widx = rng.integers(0, len(vocab), n * 8)
    target = rng.integers(8, 41, n)
    k = 0
    for i in range(n):
        s = vocab[widx[k]]
        k += 1
        while len(s) < target[i]:
            s = s + " " + vocab[widx[k % len(widx)]]
            k += 1
        out.append(s[: target[i]] if len(s) > 40 else s)
    enc = [x.encode() for x in out]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(b) for b in enc])
    heap = np.frombuffer(b"".join(enc), dtype=np.uint8).copy()
    return off, heap, out


def skinner_tables(scale=1.0, theta=0.0, device="cuda", seed=SEED + 4):
    """Five chained tables t0..t4 of SKINNER_SIZES * scale rows. Every table has
    a dense int32 `id` and an int32 `v` U{0..999}; t_i (i >= 1) has `fk`
    referencing t_{i-1}.id (uniform or Zipf(theta)); t0 has a TEXT `name`, t1 a
    TEXT `s`. Filters (1-10%): t0.name STARTS_WITH "a", t1.s LIKE "%e_a%"
    (the planner's general LIKE form), t4.v < 50.

    Returns (device tables, filter spec, host copies)."""
    import numpy as np
    from . import abi
    from .runtime import Column, TextColumn
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    sizes = [max(4, int(n * scale)) for n in SKINNER_SIZES]
    v_is_dev = torch.device(device).type == "cuda"
    vocab = vocabulary(seed=seed)
    tabs, htabs = [], []
    for i, n in enumerate(sizes):
        t = {"id": torch.arange(n, dtype=torch.int32, device=device),
             "v": torch.randint(0, 1000, (n,), generator=g, device=device, dtype=torch.int32)}
        if i > 0:
            t["fk"] = zipf_ranks(n, sizes[i - 1], theta, g, device).to(torch.int32)
            if theta > 0:  # spread the hot keys over the referenced table
                perm = torch.randperm(sizes[i - 1], generator=g, device=device).to(torch.int32)
                t["fk"] = perm[t["fk"].long()]
        ht = {k: v.cpu().numpy() for k, v in t.items()}
        # Column objects need device tensors; on the CPU the tensors stay raw
        # (host copies are what CPU callers use).
        dt = {k: Column(abi.KG_INT, v) for k, v in t.items()} if v_is_dev else dict(t)
        if i in (0, 1):
            name = "name" if i == 0 else "s"
            off, heap, lst = text_column(n, vocab, rng)
            dt[name] = TextColumn(torch.as_tensor(off).to(device), torch.as_tensor(heap).to(device))
            ht[name + "_off"], ht[name + "_heap"], ht[name + "_list"] = off, heap, lst
        tabs.append(dt)
        htabs.append(ht)
    spec = [[("text", "name", abi.KG_TEXT_STARTS_WITH, "a")],
            [("text", "s", abi.KG_TEXT_LIKE, "%e_a%")],
            [], [],
            [("int", "v", abi.KG_LT, 50)]]
    return tabs, spec, htabs


# -- batched / served mix -----------------------------------------------------------
# Resident tables: fact tables of uneven size with the columns below, one 10k-row
# dimension (pk dense, cat over 25 codes) and one build table of int64 keys (a
# permutation of 0..nbuild-1). Fact k64 keys are uniform over 0..1.1*nbuild-1
# (10% misses).
SERVE_FACT_SIZES = (10_000, 100_000, 1_000_000, 4_000_000, 16_000_000, 50_000_000)
SERVE_BUILD_ROWS = 1 << 20
FACT_COLS = ("day", "g1", "g2", "fk", "m1", "m2", "m3", "m4", "i1", "i2", "i3", "k64", "k64b")
# Templates the batch engine runs, and the served mix: those plus the five-table
# chain join with TEXT filters and MIN aggregates (c3_chain), which serving runs
# on the plain path.
BATCH_TEMPLATES = ("c1_scan_agg", "c1_small", "two_filter", "two_filter_ungrouped", "c0_star", "star_dim_filter",
                   "c2_join", "c2_join_16m", "selective_join")
SERVE_TEMPLATES = BATCH_TEMPLATES + ("c3_chain",)
# Chain tables of the served mix: the five chained tables at this fraction of
# their full sizes (10k / 50k / 200k / 1.25M / 2.5M rows).
SERVE_CHAIN_SCALE = 0.05
# LIKE patterns of served chain queries (t1.s LIKE p).
CHAIN_LIKES = ("%e_a%", "%a_e%", "%o_i%", "%i_e%", "%a_a%", "%e_o%")


def serve_tables(fact_sizes=SERVE_FACT_SIZES, ndim=10_000, nbuild=SERVE_BUILD_ROWS, device="cuda", seed=SEED + 5,
                 nbig=None, chain_scale=SERVE_CHAIN_SCALE):
    """Fact tables, the 10k-row dimension, the build table of nbuild int64
    keys and the big build table of nbig keys (16 * nbuild by default: 16M for
    the serving mix), each a permutation of 0..rows-1. Fact k64 / k64b keys are
    uniform over 1.1 times the build / big table's keys (10% misses). "chain"
    holds the five chained tables of skinner_tables at chain_scale (device
    columns and host copies)."""
    nbig = 16 * nbuild if nbig is None else nbig
    facts = []
    for j, n in enumerate(fact_sizes):
        g = _gen(device, seed + 16 * j)
        t = {
            "day": torch.randint(DAY0, DAY0 + NDAYS, (n,), generator=g, device=device, dtype=torch.int32),
            "g1": torch.randint(0, 3, (n,), generator=g, device=device, dtype=torch.int32),
            "g2": torch.randint(0, 2, (n,), generator=g, device=device, dtype=torch.int32),
            "fk": torch.randint(0, ndim, (n,), generator=g, device=device, dtype=torch.int32),
            "m1": 1.0 + torch.rand((n,), generator=g, device=device, dtype=torch.float64) * (1e5 - 1.0),
        }
        for i in range(2, 5):
            t[f"m{i}"] = torch.rand((n,), generator=g, device=device, dtype=torch.float64) * 1e5
        for i in range(1, 4):
            t[f"i{i}"] = torch.randint(1, 51, (n,), generator=g, device=device, dtype=torch.int32)
        t["k64"] = torch.randint(0, int(nbuild * 1.1), (n,), generator=g, device=device, dtype=torch.int64)
        t["k64b"] = torch.randint(0, int(nbig * 1.1), (n,), generator=g, device=device, dtype=torch.int64)
        facts.append(t)
    g = _gen(device, seed + 7)
    dim = {"pk": torch.arange(ndim, device=device, dtype=torch.int32),
           "cat": torch.randint(0, 25, (ndim,), generator=g, device=device, dtype=torch.int32)}
    build = {"key": torch.randperm(nbuild, generator=g, device=device).to(torch.int64),
             "payload": torch.randint(0, 1 << 30, (nbuild,), generator=g, device=device, dtype=torch.int32)}
    g = _gen(device, seed + 9)
    big = {"key": torch.randperm(nbig, generator=g, device=device).to(torch.int64),
           "payload": torch.randint(0, 1 << 30, (nbig,), generator=g, device=device, dtype=torch.int32)}
    ctabs, _, chost = skinner_tables(scale=chain_scale, device=device)
    return {"facts": facts, "dim": dim, "build": build, "big": big, "chain": {"tabs": ctabs, "host": chost}}


def serve_queries(nq, fact_sizes=SERVE_FACT_SIZES, nbuild=SERVE_BUILD_ROWS, seed=SEED + 6, templates=None,
                  nbig=None):
    """nq query specs: template, fact table index and parameters. Column
    indices refer to FACT_COLS, the dimension's (pk, cat) and the build / big
    build table's (key, payload). Templates and tables are drawn uniformly.
    c3_chain specs (kind "chain") join the five chain tables with
    t0.name STARTS_WITH p0, t1.s LIKE p1 and t4.v < vmax, for MIN(t0.name),
    MIN(t4.v) and the count of result tuples."""
    nbig = 16 * nbuild if nbig is None else nbig
    import numpy as np
    from . import abi
    rng = np.random.RandomState(seed & 0x7FFFFFFF)
    names = list(templates or SERVE_TEMPLATES)
    S, A, C = abi.KG_AGG_SUM, abi.KG_AGG_AVG, abi.KG_AGG_COUNT
    D = DAY0
    specs = []
    for _ in range(nq):
        tpl = names[rng.randint(len(names))]
        table = int(rng.randint(len(fact_sizes)))
        s = {"template": tpl, "table": table, "kind": "scan", "keys": [], "preds": []}
        if tpl == "c1_scan_agg":
            s["preds"] = [(0, abi.KG_LEQ, D + int(NDAYS * rng.uniform(0.8, 1.0)))]
            s["keys"] = [(1, 3), (2, 2)]
            s["aggs"] = [(S, c) for c in range(4, 11)] + [(A, c) for c in range(4, 11)] + [(C, -1)]
        elif tpl == "c1_small":
            s["preds"] = [(0, abi.KG_LEQ, D + int(NDAYS * rng.uniform(0.3, 1.0)))]
            s["keys"] = [(1, 3)]
            s["aggs"] = [(S, 4), (A, 9), (C, -1)]
        elif tpl == "two_filter":
            d0 = D + int(rng.randint(0, int(NDAYS * 0.7)))
            s["preds"] = [(0, abi.KG_GEQ, d0), (0, abi.KG_LT, d0 + int(NDAYS * rng.uniform(0.1, 0.3))),
                          (8, abi.KG_LEQ, int(rng.randint(5, 51)))]
            s["keys"] = [(2, 2)]
            s["aggs"] = [(S, 5), (S, 10), (C, -1)]
        elif tpl == "two_filter_ungrouped":
            s["preds"] = [(0, abi.KG_GEQ, D + int(rng.randint(0, int(NDAYS * 0.9)))),
                          (1, abi.KG_EQ, int(rng.randint(0, 3)))]
            s["aggs"] = [(S, 6), (A, 7), (C, -1)]
        elif tpl in ("c0_star", "star_dim_filter", "selective_join"):
            s.update(kind="join", build="dim", probe_key=3, build_key=0, build_codes=1, group_card=25,
                     build_preds=[])
            if tpl == "c0_star":
                d0 = D + int(rng.randint(0, NDAYS - 767))
                s["preds"] = [(0, abi.KG_GEQ, d0), (0, abi.KG_LT, d0 + 767)]
                s["aggs"] = [(S, 4, 8), (C, -1)]          # SUM(price * qty), COUNT(*)
            elif tpl == "star_dim_filter":
                d0 = D + int(rng.randint(0, NDAYS - 767))
                s["preds"] = [(0, abi.KG_GEQ, d0), (0, abi.KG_LT, d0 + 767)]
                s["build_preds"] = [(1, abi.KG_LT, int(rng.randint(5, 26)))]
                s["aggs"] = [(S, 5), (C, -1)]
            else:
                d0 = D + int(rng.randint(0, int(NDAYS * 0.9)))
                w = max(1, int(NDAYS * rng.uniform(0.01, 0.1)))
                s["preds"] = [(0, abi.KG_GEQ, d0), (0, abi.KG_LT, d0 + w), (9, abi.KG_LT, int(rng.randint(10, 51)))]
                s["aggs"] = [(C, -1), (S, 8)]
        elif tpl in ("c2_join", "c2_join_16m"):
            # A key range of 10-50% of the build (or the 16M-key big build)
            # table is built; the fact's k64 (k64b) keys probe it.
            nb = nbuild if tpl == "c2_join" else nbig
            lo = int(rng.randint(0, nb // 2))
            hi = lo + int(nb * rng.uniform(0.1, 0.5))
            s.update(kind="join", build="build" if tpl == "c2_join" else "big",
                     probe_key=11 if tpl == "c2_join" else 12, build_key=0, build_codes=-1, group_card=1,
                     build_preds=[(0, abi.KG_GEQ, lo), (0, abi.KG_LT, hi)])
            s["preds"] = [(0, abi.KG_GEQ, D + int(rng.randint(0, int(NDAYS * 0.5))))]
            s["aggs"] = [(C, -1), (S, 10)]
        elif tpl == "c3_chain":
            s.update(kind="chain", p0=chr(97 + int(rng.randint(26))), p1=CHAIN_LIKES[rng.randint(len(CHAIN_LIKES))],
                     vmax=int(rng.randint(10, 101)), aggs=[])
        else:
            raise ValueError(tpl)
        specs.append(s)
    return specs


def serve_columns(tables):
    """Column objects of serve_tables: (facts [FACT_COLS order], dim, builds),
    builds = {"build": [key, payload], "big": [key, payload]}."""
    from . import abi
    from .runtime import Column
    types = {"day": abi.KG_DATE, "g1": abi.KG_ENUM, "g2": abi.KG_ENUM, "fk": abi.KG_INT, "m1": abi.KG_REAL,
             "m2": abi.KG_REAL, "m3": abi.KG_REAL, "m4": abi.KG_REAL, "i1": abi.KG_INT, "i2": abi.KG_INT,
             "i3": abi.KG_INT, "k64": abi.KG_BIGINT, "k64b": abi.KG_BIGINT}
    facts = [[Column(types[c], t[c]) for c in FACT_COLS] for t in tables["facts"]]
    dim = [Column(abi.KG_INT, tables["dim"]["pk"]), Column(abi.KG_ENUM, tables["dim"]["cat"])]
    builds = {b: [Column(abi.KG_BIGINT, tables[b]["key"]), Column(abi.KG_INT, tables[b]["payload"])]
              for b in ("build", "big")}
    if "chain" in tables:
        builds["chain"] = tables["chain"]["tabs"]
    return facts, dim, builds


def make_query(spec, facts, dim, builds, load_factor=0.0):
    """runtime.Query for a spec of serve_queries over serve_columns."""
    from .runtime import Query
    if spec["kind"] == "chain":
        return chain_query_for(builds["chain"], spec["p0"], spec["p1"], spec["vmax"],
                               indexes=builds.get("chain_idx"))
    cols = facts[spec["table"]]
    if spec["kind"] == "scan":
        return Query.scan_agg(cols, spec["preds"], spec["keys"], spec["aggs"], table_id=spec["table"])
    bcols = dim if spec["build"] == "dim" else builds[spec["build"]]
    return Query.join_agg(cols, spec["preds"], spec["probe_key"], bcols, spec["build_key"], spec["aggs"],
                          build_preds=spec["build_preds"], build_codes=spec["build_codes"],
                          group_card=spec["group_card"], table_id=spec["table"], load_factor=load_factor)


def chain_indexes(ctx, tabs):
    """Stored column indexes on the fk columns of chain tables t1.. over all
    rows (None for t0): what an unfiltered child probes instead of an index
    built per query."""
    from .runtime import ColumnIndex
    return [None] + [ColumnIndex(ctx, t["fk"]) for t in tabs[1:]]


def chain_query_for(tabs, p0, p1, vmax, skinner_budget=0, seed=1, indexes=None):
    """runtime.Query (KG_Q_CHAIN) over the five chain tables of
    skinner_tables: t0.name STARTS_WITH p0, t1.s LIKE p1, t4.v < vmax;
    MIN(t0.name), MIN(t4.v). indexes: chain_indexes (optional)."""
    from . import abi
    from .runtime import Query
    tables = []
    for i, t in enumerate(tabs):
        cols = [t["id"], t["v"]] + ([t["fk"]] if i > 0 else [])
        d = {"cols": cols, "id_col": 0, "fk_col": 2 if i > 0 else -1,
             "fk_index": indexes[i] if indexes is not None else None}
        if i == 0:
            d.update(text=t["name"], text_op=abi.KG_TEXT_STARTS_WITH, pattern=p0.encode())
        elif i == 1:
            d.update(text=t["s"], text_op=abi.KG_TEXT_LIKE, pattern=p1.encode())
        if i == len(tabs) - 1:
            d["preds"] = [(1, abi.KG_LT, vmax)]
        tables.append(d)
    return Query.chain(tables, min_text=(0, tabs[0]["name"]), min_int=(len(tabs) - 1, 1),
                       skinner_budget=skinner_budget, seed=seed)
