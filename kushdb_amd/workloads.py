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
