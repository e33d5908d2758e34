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
        dt = {k: Column(abi.KG_INT, v) for k, v in t.items()}
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
