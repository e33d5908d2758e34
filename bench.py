"""Benchmark of the MI355X pipeline-operator path (one GPU).

Headline workload: ScanSelect(day <= cutoff, ~98%) -> GroupByAggregate over
two enum keys (3 x 2 = 6 groups) with SUM and AVG of four float64 and three
int32 measures plus COUNT(*), on 60M resident rows. A step is one
kg_scan_group_agg call over all rows (the result, 6 groups, is read back).

Also reported ("extra"): join build (16M unique int64 keys), join probe
(256M int64 keys, 10% misses, uniform), and the probe fused with GROUP BY the
matched build key (SUM(int64), COUNT(*)).

Prints ONE JSON line.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "probe/agg rows/s + HBM GB/s fraction; batched vs plain queries/s at p99"
HBM_PEAK_GBS = 8000.0
SCAN_AGG_BYTES_PER_ROW = 4 + 2 * 4 + 4 * 8 + 3 * 4  # day, 2 enum codes, 4 f64, 3 i32 = 56 B
PROBE_BYTES_PER_ROW = 8 + 64                       # streamed key + one 64 B bucket
PROBE_AGG_BYTES_PER_ROW = 8 + 8 + 64 + 64          # key + value + bucket + accumulator update
BUILD_BYTES_PER_ROW = 8 + 4 + 64                   # key + payload + one 64 B insert


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=60_000_000)
    ap.add_argument("--cpu-rows", type=int, default=60_000_000, help="rows of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-join", action="store_true")
    ap.add_argument("--build-rows", type=int, default=1 << 24)
    ap.add_argument("--probe-rows", type=int, default=1 << 28)
    ap.add_argument("--thetas", type=float, nargs="+", default=[0.0, 0.5, 1.0])
    ap.add_argument("--cpu-probe-rows", type=int, default=1 << 25, help="probe rows of the C2 CPU baseline sample")
    ap.add_argument("--lf", type=float, default=0.0, help="join table load factor (0: autotuned)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-skinner", action="store_true", help="C3 without the Skinner-driven runs")
    ap.add_argument("--only", choices=["c0_star", "c2_join", "c3_chain", "serving", "c2_build", "c2_probe_agg"],
                    help="run one workload alone (for profiling) and print only its section")
    ap.add_argument("--serve-queries", type=int, default=512)
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-saturation-only", action="store_true",
                    help="serving: one batched pass of the mix at once (for kernel traces)")
    ap.add_argument("--serve-rate-only", type=float, default=0.0,
                    help="serving: one batched pass of the mix at this offered rate, q/s (for kernel traces)")
    ap.add_argument("--serve-caps", type=int, nargs="+", default=[100_000_000, 0],
                    help="batch work caps (rows) tried for the batched mode; 0 = no cap")
    ap.add_argument("--serve-max-batch", type=int, default=128)
    ap.add_argument("--serve-lanes", type=int, default=1, help="plain lanes of batched serving")
    ap.add_argument("--serve-hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for this process (0: leave the environment's value)")
    ap.add_argument("--serve-lane-scan-rows", type=int, default=0,
                    help="batched serving: scans of at least this many rows arriving with nothing pending run "
                         "on an idle plain lane (0: never)")
    ap.add_argument("--serve-arena", type=int, default=1024, help="MiB of device arena per batch slot")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def roofline(bytes_per_launch, ms, traffic=None):
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic}


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")
# Kernels whose HBM traffic the bench reports, by a substring of their name.
TRAFFIC_KERNELS = {"k_group_agg_v4": "k_group_agg_v4<", "k_ht_build": "k_ht_build(",
                   "k_join_group_agg": "k_join_group_agg<",
                   "k_ht_probe_count": "k_ht_probe<0,", "k_ht_probe_agg": "k_ht_probe<2,",
                   "k_batch_shared": "k_batch_shared(", "k_batch_join": "k_batch_join("}
COPY_BYTES = 1 << 30


def measured_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/traffic.json), or None when not measured."""
    try:
        with open(TRAFFIC_JSON) as f:
            v = json.load(f)["kernels"].get(kernel, {}).get("bytes_per_launch")
            return int(v) if v is not None else None
    except (OSError, ValueError, KeyError):
        return None


def copy_calibration():
    """A plain device copy of COPY_BYTES (three launches): the known-size
    reference for FETCH_SIZE / WRITE_SIZE."""
    import torch
    src = torch.ones(COPY_BYTES, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "launches": 3}))
    return 0


def pmc_query_summary(fetch_csv, write_csv, name, first_kernel):
    """HBM bytes of one query run alone under --pmc (bench.py --only <name>
    --warmup 0 --steps 1): FETCH_SIZE + WRITE_SIZE summed over every dispatch
    from the first one whose kernel name contains first_kernel (the query's
    first kernel; everything before it is table generation) to the end,
    scaled by the calibration already in profiles/traffic.json, stored under
    "queries"/name with the kernels dispatched."""
    import csv

    def dispatches(path, counter):
        acc = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            d = acc.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0])
            d[1] += float(r["Counter_Value"])
        seq = [acc[k] for k in sorted(acc)]
        start = next(i for i, (k, _) in enumerate(seq) if first_kernel in k)
        return seq[start:]

    with open(TRAFFIC_JSON) as fh:
        out = json.load(fh)
    ff, wf = out["calibration"]["fetch_factor"], out["calibration"]["write_factor"]
    fseq, wseq = dispatches(fetch_csv, "FETCH_SIZE"), dispatches(write_csv, "WRITE_SIZE")
    assert [k for k, _ in fseq] == [k for k, _ in wseq], "the two passes dispatched different kernels"
    f, w = sum(v for _, v in fseq), sum(v for _, v in wseq)
    kernels = {}
    for k, _ in fseq:
        short = k.split("(")[0].replace("void ", "")[:60]
        kernels[short] = kernels.get(short, 0) + 1
    out.setdefault("queries", {})[name] = {
        "first_kernel": first_kernel, "dispatches": len(fseq), "fetch_kib": f, "write_kib": w,
        "bytes_per_query": round((f * ff + w * wf) * 1024), "kernels": kernels}
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out["queries"][name], indent=1))
    return 0


def pmc_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the reported kernels,
    scaled by the copy kernel's reading of its known size, into
    profiles/traffic.json."""
    import csv

    def per_kernel(path, counter):
        """Per kernel, the largest dispatch's reading: the bench's full-size
        launch (smaller launches of the same kernel are tuning samples or
        other workloads)."""
        acc = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            acc.setdefault(r["Kernel_Name"], {}).setdefault(r["Dispatch_Id"], 0.0)
            acc[r["Kernel_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
        return {k: max(v.values()) for k, v in acc.items()}

    def largest(d):
        k = max(d, key=d.get)
        return k, d[k]

    cname, cf = largest(per_kernel(copy_fetch_csv, "FETCH_SIZE"))
    _, cw = largest(per_kernel(copy_write_csv, "WRITE_SIZE"))
    ff, wf = COPY_BYTES / (cf * 1024), COPY_BYTES / (cw * 1024)
    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    out = {"calibration": {"kernel": cname[:80], "bytes": COPY_BYTES, "fetch_kib": cf, "write_kib": cw,
                           "fetch_factor": ff, "write_factor": wf},
           "kernels": {}}
    for short, sub in TRAFFIC_KERNELS.items():
        fk = [v for k, v in fetch.items() if sub in k]
        wk = [v for k, v in write.items() if sub in k]
        if not fk or not wk:
            continue
        f, w = max(fk), max(wk)
        out["kernels"][short] = {"fetch_kib": f, "write_kib": w,
                                 "bytes_per_launch": round((f * ff + w * wf) * 1024)}
    try:
        with open(TRAFFIC_JSON) as fh:
            prev = json.load(fh).get("queries")
        if prev:
            out["queries"] = prev
    except (OSError, ValueError):
        pass
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))
    return 0


def bench_scan_agg(ctx, a):
    import torch
    from kushdb_amd import abi
    from kushdb_amd.runtime import Column
    from kushdb_amd.workloads import scan_agg_table
    n = a.rows
    log(f"scan-agg: generating {n} rows")
    t = scan_agg_table(n, device="cuda:0")
    cols = [Column(abi.KG_DATE, t["day"]), Column(abi.KG_ENUM, t["g1"]), Column(abi.KG_ENUM, t["g2"])]
    cols += [Column(abi.KG_REAL, t[f"m{i}"]) for i in range(1, 5)]
    cols += [Column(abi.KG_INT, t[f"i{i}"]) for i in range(1, 4)]
    aggs = [(abi.KG_AGG_SUM, 3 + i) for i in range(7)] + [(abi.KG_AGG_AVG, 3 + i) for i in range(7)]
    aggs.append((abi.KG_AGG_COUNT, -1))
    preds = [(0, abi.KG_LEQ, t["cutoff"])]
    keys = [(1, 3), (2, 2)]
    for _ in range(a.warmup):
        ctx.scan_group_agg(cols, preds, keys, aggs)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.kernel_stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = ctx.scan_group_agg(cols, preds, keys, aggs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, launches = ctx.kernel_stats(reset=True)
    ctx.set_timing(False)
    assert launches == a.steps, (launches, a.steps)
    groups = len(res[0])
    return t, wall, kms / launches, groups


def cpu_baseline_scan_agg(t, nrows):
    """The oracle's single-threaded scan + group-by (same operator chain and
    evaluation order as the reference's generated code) on the first nrows."""
    import numpy as np
    from oracle import pyoracle as po
    h = {k: (v[:nrows].cpu().numpy() if hasattr(v, "cpu") else v) for k, v in t.items()}
    fm = [np.ascontiguousarray(h[f"m{i}"]) for i in range(1, 5)]
    im = [np.ascontiguousarray(h[f"i{i}"]) for i in range(1, 4)]
    fp = (ctypes.c_void_p * 4)(*[x.ctypes.data for x in fm])
    ip = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in im])
    out = np.zeros((16, 17), dtype=np.int64)
    n = len(h["day"])
    t0 = time.perf_counter()
    po.lib().kor_q_scan_agg(n, h["day"].ctypes.data, h["cutoff"], h["g1"].ctypes.data, h["g2"].ctypes.data, 4, fp,
                            3, ip, 16, out.ctypes.data)
    dt = time.perf_counter() - t0
    return n / dt, dt, n


def bench_join_theta(ctx, a, theta, cpu):
    """C2 at one Zipf skew: build (16M unique int64 keys), probe counting the
    join cardinality, probe emitting every (probe row, build row) pair, and
    the probe fused with GROUP BY the matched build key (SUM(int64), COUNT(*)),
    each at full size. With cpu, the oracle's chained join table timed on the
    host over the full build and a probe sample."""
    import torch
    from kushdb_amd import abi
    from kushdb_amd.runtime import Column, HashTable
    from kushdb_amd.workloads import join_tables
    nb, npr = a.build_rows, a.probe_rows
    log(f"join: generating build {nb} / probe {npr} (theta={theta})")
    bk, pay, pk, pv = join_tables(nb, npr, theta=theta, device="cuda:0")
    bcol, pcol = Column(abi.KG_BIGINT, bk), Column(abi.KG_BIGINT, pk)
    out = {}
    lf = a.lf
    if lf == 0:
        tune = ctx.join_tune(bcol, pcol)
        lf = tune["load_factor"]
        out["join_tuning"] = tune
        log(f"join tuning (theta={theta}): {tune}")
    steps = max(1, min(a.steps, 5))
    # build
    build_ms = []
    for i in range(steps + 1):
        ht = HashTable(ctx, nb, load_factor=lf)
        torch.cuda.synchronize()
        ctx.set_timing(True)
        ctx.kernel_stats(reset=True)
        ht.insert(bcol)
        kms, _ = ctx.kernel_stats(reset=True)
        ctx.set_timing(False)
        if i > 0:
            build_ms.append(kms)
        if i < steps:
            ht.free()
    bms = sum(build_ms) / len(build_ms)
    out["join_build"] = {"rows": nb, "theta": theta, "load_factor": lf, "kernel_ms": round(bms, 3),
                         "rows_per_s": nb / (bms * 1e-3),
                         "roofline": roofline(nb * BUILD_BYTES_PER_ROW, bms,
                                              query_traffic("c2_build") if theta == 0 else None)}
    # probe (join cardinality)
    ht.probe_count(pcol)
    ctx.set_timing(True)
    ctx.kernel_stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        card = ht.probe_count(pcol)
    wall = time.perf_counter() - t0
    kms, launches = ctx.kernel_stats(reset=True)
    ctx.set_timing(False)
    pms = kms / steps   # every timed span of the steps' calls
    out["join_probe"] = {"rows": npr, "theta": theta, "matches": card, "kernel_ms": round(pms, 3),
                         "ms_per_step": round(wall / steps * 1e3, 3), "rows_per_s": npr / (pms * 1e-3),
                         "roofline": roofline(npr * PROBE_BYTES_PER_ROW, pms,
                                              measured_traffic("k_ht_probe_count") if theta == 0 else None)}
    # probe emitting the pairs (two int32 row ids per match)
    op = torch.empty(card, dtype=torch.int32, device="cuda:0")
    ob = torch.empty(card, dtype=torch.int32, device="cuda:0")
    ht.probe(pcol, out_cap=card, out=(op, ob))
    ctx.set_timing(True)
    ctx.kernel_stats(reset=True)
    for _ in range(steps):
        gp, _ = ht.probe(pcol, out_cap=card, out=(op, ob))
    kms, launches = ctx.kernel_stats(reset=True)
    ctx.set_timing(False)
    ems = kms / steps
    assert gp.numel() == card
    out["join_probe_emit"] = {"rows": npr, "theta": theta, "pairs": card, "kernel_ms": round(ems, 3),
                              "rows_per_s": npr / (ems * 1e-3),
                              "roofline": roofline(npr * PROBE_BYTES_PER_ROW + card * 8, ems)}
    del op, ob, gp
    # probe fused with GROUP BY matched key
    ht.probe_agg(pcol, pv, max_groups=nb)
    ctx.set_timing(True)
    ctx.kernel_stats(reset=True)
    for _ in range(steps):
        k, s, c, _ = ht.probe_agg(pcol, pv, max_groups=nb)
    kms, launches = ctx.kernel_stats(reset=True)
    ctx.set_timing(False)
    ams = kms / steps
    groups = int(k.numel())
    # SURVEY cost model: matched rows (key 8 + value 8 + 64 B update), every
    # probe row's key and bucket (8 + 64), 24 B per group out.
    agg_bytes = npr * PROBE_BYTES_PER_ROW + card * (8 + 64) + groups * 24
    out["join_probe_agg"] = {"rows": npr, "theta": theta, "groups": groups, "kernel_ms": round(ams, 3),
                             "rows_per_s": npr / (ams * 1e-3),
                             "roofline": roofline(agg_bytes, ams,
                                                  query_traffic("c2_probe_agg") if theta == 0 else None)}
    ht.free()
    if cpu:
        from oracle import pyoracle as po
        ns = min(a.cpu_probe_rows, npr)
        hb, hp, hv = bk.cpu().numpy(), pk[:ns].cpu().numpy(), pv[:ns].cpu().numpy()
        log(f"cpu baseline C2 (oracle, 1 thread): build {nb}, probe sample {ns}")
        m, _, bs, ps = po.time_join(hb, hp, hv)
        out["cpu_baseline"] = {"value": ns / ps, "unit": "probe rows/s", "cores": 1, "kind": "port",
                               "build_rows_per_s": nb / bs,
                               "sample": f"the full {nb}-key build ({bs:.2f} s) and the first {ns} probe rows "
                                         f"({ps:.2f} s, {m} matches) of the same tables, chained join table as in "
                                         f"runtime/hash_table.cc, {platform.processor() or platform.machine()}"}
    del bk, pay, pk, pv
    torch.cuda.empty_cache()
    return out


def bench_c2_op(ctx, a, op):
    """One C2 operation at full size and theta = --thetas[0], once, for
    counter runs: 'c2_build' generates the tables and builds the 16M-key table
    (load factor --lf, 0.35 when 0: the tuner's choice at this size);
    'c2_probe_agg' also runs one probe + GROUP BY matched key after the build.
    The operation's dispatches are the last ones of the process."""
    import torch
    from kushdb_amd import abi
    from kushdb_amd.runtime import Column, HashTable
    from kushdb_amd.workloads import join_tables
    bk, _, pk, pv = join_tables(a.build_rows, a.probe_rows, theta=a.thetas[0], device="cuda:0")
    torch.cuda.synchronize()
    ht = HashTable(ctx, a.build_rows, load_factor=a.lf or 0.35)
    ht.insert(Column(abi.KG_BIGINT, bk))
    out = {"op": op, "build_rows": a.build_rows, "load_factor": a.lf or 0.35}
    if op == "c2_probe_agg":
        k, _, _, _ = ht.probe_agg(Column(abi.KG_BIGINT, pk), pv, max_groups=a.build_rows)
        out.update(probe_rows=a.probe_rows, groups=int(k.numel()))
    torch.cuda.synchronize()
    ht.free()
    return out


def pmc_counters(name, *csvs):
    """Per kernel (largest dispatch of each), every counter of the given
    rocprofv3 --pmc counter_collection CSVs, summed over the dispatch's
    instances, as JSON under profiles/counters.json[name]."""
    import csv
    per = {}
    for path in csvs:
        acc = {}
        for r in csv.DictReader(open(path)):
            k = (r["Kernel_Name"], r["Dispatch_Id"])
            acc.setdefault(k, {}).setdefault(r["Counter_Name"], 0.0)
            acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
        for (kern, did), cs in acc.items():
            short = kern.split("(")[0].replace("void ", "").replace("kg::", "")[:70]
            for cn, v in cs.items():
                d = per.setdefault(short, {}).setdefault(cn, [])
                d.append(v)
    out = {k: {cn: {"dispatches": len(v), "max": max(v), "sum": sum(v)} for cn, v in cs.items()}
           for k, cs in per.items()}
    path = os.path.join(ROOT, "profiles", "counters.json")
    try:
        with open(path) as fh:
            allc = json.load(fh)
    except (OSError, ValueError):
        allc = {}
    allc[name] = out
    with open(path, "w") as fh:
        json.dump(allc, fh, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))
    return 0


def bench_join(ctx, a):
    """C2 at theta 0 (the keys the earlier records use) and every other skew
    in --thetas."""
    out = {}
    for theta in a.thetas:
        r = bench_join_theta(ctx, a, theta, cpu=not a.no_cpu and theta == a.thetas[0])
        if theta == a.thetas[0]:
            out.update(r)
        else:
            out[f"c2_theta_{theta}"] = r
    return out


C0_BYTES = 1_000_000 * (4 + 4 + 4 + 8) + 10_000 * (4 + 4)   # fk, day, qty, price; dim pk, cat


def bench_star_c0(ctx, a):
    """C0: 1M-row fact filtered to a 30% day range, joined with a 10k-row
    dimension on fk = pk, grouped by the dimension's 25 categories with
    SUM(price * Float(qty)) and COUNT(*): dim build + one fused
    scan/probe/aggregate call per query. Latency-bound at this size."""
    import torch
    from kushdb_amd import abi
    from kushdb_amd.runtime import Column, HashTable
    from kushdb_amd.workloads import star_tables
    f, d, (lo, hi) = star_tables(device="cuda:0")
    cols = [Column(abi.KG_INT, f["fk"]), Column(abi.KG_INT, f["qty"]), Column(abi.KG_REAL, f["price"]),
            Column(abi.KG_DATE, f["day"])]
    preds = [(3, abi.KG_GEQ, lo), (3, abi.KG_LT, hi)]
    aggs = [(abi.KG_AGG_SUM, 2, 1), (abi.KG_AGG_COUNT, -1)]
    pk, cat = Column(abi.KG_INT, d["pk"]), Column(abi.KG_ENUM, d["cat"])

    def query():
        ht = HashTable(ctx, 10_000)
        ht.insert(pk)
        r = ht.scan_probe_group_agg(cols, preds, 0, cat, 25, aggs, max_groups=64)
        ht.free()
        return r
    # Alone (--only, e.g. under a profiler) the given warmup/steps; else 3 / 20.
    warm, steps = (a.warmup, max(1, a.steps)) if a.only else (3, 20)
    for _ in range(warm):
        query()
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.kernel_stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        codes, vals, _ = query()
    wall = (time.perf_counter() - t0) / steps
    kms, launches = ctx.kernel_stats(reset=True)
    ctx.set_timing(False)
    # timed kernels per query: the dim build and the fused probe kernel
    per_query_kernel_ms = kms / steps
    out = {"workload": "1M-row fact (30% day range) join 10k-row dim, GROUP BY 25 cats: SUM(price*qty), COUNT(*)",
           "groups": int(len(codes)), "query_ms": round(wall * 1e3, 4), "kernel_ms": round(per_query_kernel_ms, 4),
           "rows_per_s": 1_000_000 / wall,
           "roofline": roofline(C0_BYTES, per_query_kernel_ms, query_traffic("c0_star"))}
    if not a.no_cpu:
        import numpy as np
        from oracle import pyoracle as po
        h = {k: v.cpu().numpy() for k, v in f.items()}
        hd = {k: v.cpu().numpy() for k, v in d.items()}
        oc, osum, ocnt = np.zeros(64, np.int32), np.zeros(64), np.zeros(64, np.int64)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            g = po.lib().kor_q_star(len(h["fk"]), h["fk"].ctypes.data, h["qty"].ctypes.data, h["price"].ctypes.data,
                                    h["day"].ctypes.data, lo, hi, len(hd["pk"]), hd["pk"].ctypes.data,
                                    hd["cat"].ctypes.data, 64, oc.ctypes.data, osum.ctypes.data, ocnt.ctypes.data)
            times.append(time.perf_counter() - t0)
        assert g == len(codes) and np.array_equal(np.sort(ocnt[:g]), np.sort(vals[:, 1]))
        ct = sorted(times)[2]
        out["cpu_baseline"] = {"value": 1_000_000 / ct, "unit": "fact rows/s", "cores": 1, "kind": "port",
                               "sample": f"the whole query (1M fact rows, 10k dim rows), median of 5 runs, "
                                         f"{ct * 1e3:.1f} ms, scan-select + chained hash join + aggregate "
                                         f"table in the reference's orders"}
    return out


def bench_chain_c3(ctx, a):
    """C3: five chained tables (200k, 1M, 4M, 25M, 50M rows), filters incl.
    STARTS_WITH / LIKE on TEXT and v < 50, chain join through device column
    indexes in the order t0..t4, tuple-id dedup, MIN(t0.name), MIN(t4.v)."""
    import torch
    from kushdb_amd.chain import chain_query
    from kushdb_amd.workloads import SKINNER_SIZES, skinner_tables
    log("C3: generating chain tables")
    tabs, spec, htabs = skinner_tables(scale=1.0, theta=0.0, device="cuda:0")
    torch.cuda.synchronize()
    order = [0, 1, 2, 3, 4]
    # Alone (--only, e.g. under a profiler) the given warmup/steps; else 1 / 3.
    warm, steps = (a.warmup, max(1, a.steps)) if a.only else (1, 3)
    for _ in range(warm):
        chain_query(ctx, tabs, spec, order)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tuples, mn_text, mn_int, stats = chain_query(ctx, tabs, spec, order)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    n_in = sum(SKINNER_SIZES)
    # referenced bytes: filter columns of t0 (name: 8-B offset + bytes), t1 (s), t4 (v: 4 B); join
    # keys of the selected rows (4 B each); 64 B per index insert, index probe and tuple insert.
    text_bytes = sum(int(htabs[i][c + "_heap"].nbytes) + 8 * len(htabs[i]["id"]) for i, c in ((0, "name"), (1, "s")))
    alg = (text_bytes + 4 * SKINNER_SIZES[4] + 4 * sum(stats["selected"]) +
           64 * (stats["index_inserts"] + stats["index_probes"] + stats["tuples"]))
    out = {"workload": "5-table chain join (200k..50M rows), TEXT prefix + LIKE + int filters, device indexes, "
                       "tuple dedup, MIN(text), MIN(int)", "order": order, "stats": stats,
           "query_ms": round(wall * 1e3, 3), "rows_per_s": n_in / wall,
           "min_int": mn_int, "roofline": roofline(alg, wall * 1e3, c3_traffic())}
    if not a.no_cpu:
        # The oracle's chain query on the same full-size tables: its result is
        # also the full-size check of the device result.
        from oracle import pyoracle as po
        log("cpu baseline C3 (oracle, 1 thread)")
        t0 = time.perf_counter()
        cn, cname, cvmin = po.chain_query(htabs)
        ct = time.perf_counter() - t0
        as_bytes = lambda x: x.encode() if isinstance(x, str) else (bytes(x) if x is not None else None)
        same = cn == tuples.shape[0] and cvmin == mn_int and as_bytes(cname) == as_bytes(mn_text)
        out["cpu_baseline"] = {"value": n_in / ct, "unit": "input rows/s", "cores": 1, "kind": "port",
                               "sample": f"the whole query over the same {n_in} rows, {ct:.2f} s: filters, four "
                                         f"left-deep chained hash joins, MIN; result equal to the device's: {same}"}
        out["oracle_equal"] = bool(same)
    # The same query as one native plan (kg_chain_run: the same entry points
    # called from C++, intermediate memory from the context's block cache).
    from kushdb_amd.runtime import chain_run
    from kushdb_amd.workloads import chain_query_for
    q = chain_query_for(tabs, "a", "%e_a%", 50)
    chain_run(ctx, q)
    t0 = time.perf_counter()
    for _ in range(steps):
        nn, nvmin, nrow = chain_run(ctx, q)
    nwall = (time.perf_counter() - t0) / steps
    ntext = bytes(htabs[0]["name_heap"][htabs[0]["name_off"][nrow]:htabs[0]["name_off"][nrow + 1]]) \
        if nrow is not None else None
    as_b = lambda x: x.encode() if isinstance(x, str) else (bytes(x) if x is not None else None)
    out["native_plan"] = {"query_ms": round(nwall * 1e3, 3), "rows_per_s": n_in / nwall,
                          "equal_to_stepwise": bool(nn == tuples.shape[0] and nvmin == mn_int and
                                                    ntext == as_b(mn_text))}
    log(f"C3 native plan: {nwall * 1e3:.2f} ms ({out['native_plan']})")
    # With stored column indexes on the fk columns (built once, as the
    # reference's .kdbindex files are): unfiltered children (t2, t3) are
    # probed through them instead of being indexed per query.
    from kushdb_amd.workloads import chain_indexes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    idx = chain_indexes(ctx, tabs)
    torch.cuda.synchronize()
    ibuild = time.perf_counter() - t0
    qs = chain_query_for(tabs, "a", "%e_a%", 50, indexes=idx)
    chain_run(ctx, qs)
    t0 = time.perf_counter()
    for _ in range(steps):
        sn, svmin, srow = chain_run(ctx, qs)
    swall_i = (time.perf_counter() - t0) / steps
    stext = bytes(htabs[0]["name_heap"][htabs[0]["name_off"][srow]:htabs[0]["name_off"][srow + 1]]) \
        if srow is not None else None
    out["native_plan_stored_indexes"] = {
        "query_ms": round(swall_i * 1e3, 3), "rows_per_s": n_in / swall_i,
        "index_build_once_ms": round(ibuild * 1e3, 2),
        "equal_to_stepwise": bool(sn == tuples.shape[0] and svmin == mn_int and stext == as_b(mn_text))}
    log(f"C3 native plan, stored fk indexes: {swall_i * 1e3:.2f} ms ({out['native_plan_stored_indexes']})")
    # The reference's choice per child: the stored index of an unfiltered child
    # (t2, t3), an index built per query over a filtered child's selection (t1, t4).
    qr = chain_query_for(tabs, "a", "%e_a%", 50, indexes=[None, None, idx[2], idx[3], None])
    chain_run(ctx, qr)
    t0 = time.perf_counter()
    for _ in range(steps):
        rn, rvmin, rrow = chain_run(ctx, qr)
    rwall = (time.perf_counter() - t0) / steps
    rtext = bytes(htabs[0]["name_heap"][htabs[0]["name_off"][rrow]:htabs[0]["name_off"][rrow + 1]]) \
        if rrow is not None else None
    out["native_plan_stored_unfiltered"] = {
        "query_ms": round(rwall * 1e3, 3), "rows_per_s": n_in / rwall,
        "equal_to_stepwise": bool(rn == tuples.shape[0] and rvmin == mn_int and rtext == as_b(mn_text))}
    log(f"C3 native plan, stored indexes of the unfiltered children: {rwall * 1e3:.2f} ms")
    # The same query with the join order chosen per episode by the UCT agent (kg_skinner_join).
    from kushdb_amd.chain import skinner_chain_query
    for budget in () if a.no_skinner else (1 << 20, 1 << 22):
        skinner_chain_query(ctx, tabs, spec, budget=budget, seed=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st_tuples, st_text, st_int, sst = skinner_chain_query(ctx, tabs, spec, budget=budget, seed=1)
        torch.cuda.synchronize()
        swall = (time.perf_counter() - t0) / steps
        assert st_tuples.shape[0] == tuples.shape[0] and st_int == mn_int and st_text == mn_text
        out[f"skinner_budget_{budget}"] = {
            "query_ms": round(swall * 1e3, 3), "rows_per_s": n_in / swall,
            "episodes": sst["episodes"], "orders_tried": sst["distinct_orders"], "last_order": sst["last_order"],
            "partial_tuples": sst["partial_tuples"], "result_tuples": sst["result_tuples"]}
        log(f"C3 skinner budget {budget}: {swall * 1e3:.2f} ms, {sst['episodes']} episodes, "
            f"{sst['distinct_orders']} orders, last {sst['last_order']}")
    return out


def query_traffic(name):
    """HBM bytes of one whole query (every dispatch of it, FETCH_SIZE +
    WRITE_SIZE scaled by the copy calibration; profiles/traffic.json
    "queries"), or None when not measured."""
    try:
        with open(TRAFFIC_JSON) as f:
            v = json.load(f).get("queries", {}).get(name, {}).get("bytes_per_query")
            return int(v) if v is not None else None
    except (OSError, ValueError):
        return None


def c3_traffic():
    return query_traffic("c3_chain")


def bench_serving(ctx, a):
    """Batched/served mix: Poisson arrivals of queries from 9 templates over
    resident fact tables of 10k..50M rows. Plain one-at-a-time GPU path vs
    continuous batching. The p99 bar is the plain path's p99 at 70% of its
    saturation rate; for each mode the highest offered rate whose p99 stays
    within the bar is found, and the achieved queries/s there are compared."""
    import numpy as np
    import torch
    from kushdb_amd import abi
    from kushdb_amd.runtime import serve
    from kushdb_amd.workloads import (SERVE_FACT_SIZES, SERVE_TEMPLATES, chain_indexes, make_query, serve_columns,
                                      serve_queries, serve_tables)
    log(f"serving: generating tables {SERVE_FACT_SIZES}")
    tables = serve_tables(device="cuda:0")
    torch.cuda.synchronize()
    facts, dim, build = serve_columns(tables)
    # Stored indexes on the chain tables' fk columns (unfiltered children probe them).
    build["chain_idx"] = chain_indexes(ctx, build["chain"])
    specs = serve_queries(a.serve_queries)
    queries = [make_query(s, facts, dim, build) for s in specs]
    # The join tuner's load factors for the two build tables (cached on the
    # context: join queries of both serving modes take them).
    for bname, pcol in (("build", 11), ("big", 12)):
        log(f"serving: join tuning for the {bname} table: {ctx.join_tune(build[bname][0], facts[-1][pcol])}")
    nq = len(queries)
    tpl_of = np.array([sp["template"] for sp in specs])
    rng = np.random.RandomState(7)
    expo = rng.exponential(1.0, size=nq)

    def run(mode, rate, cap=0, keep=False):
        arr = np.zeros(nq) if rate is None else np.cumsum(expo / rate)
        res, lat, st = serve(ctx, queries, arr, mode, slots=a.serve_slots, max_batch=a.serve_max_batch,
                             arena_bytes=a.serve_arena << 20, max_groups=64, max_batch_rows=cap,
                             lanes=a.serve_lanes, lane_scan_rows=a.serve_lane_scan_rows)
        out = {"offered": rate, "cap_rows": cap, "qps": st["qps"], "p50_ms": st["p50_s"] * 1e3,
               "p99_ms": st["p99_s"] * 1e3, "mean_batch": st["mean_batch"], "batches": st["batches"],
               "fallbacks": st["fallbacks"], "lane_scans": st["lane_scans"]}
        # Latency by template: [p50, p99] in ms.
        lat_ms = np.asarray(lat[:nq]) * 1e3
        out["by_template"] = {t: [round(float(np.percentile(lat_ms[tpl_of == t], q)), 3) for q in (50, 99)]
                              for t in SERVE_TEMPLATES if (tpl_of == t).any()}
        return (out, res) if keep else out

    t0 = tables["chain"]["host"][0]

    def chain_text(row):
        return bytes(t0["name_heap"][t0["name_off"][row]:t0["name_off"][row + 1]])

    def same_results(ra, rb):
        """Per query: equal status, group keys and order, counts and integer
        sums exactly, FP64 sums and averages within 1e-12 relative."""
        from kushdb_amd.workloads import FACT_COLS
        bad = 0
        for i, (sp, x, y) in enumerate(zip(specs, ra, rb)):
            ok = x[0] == y[0] and x[1].shape == y[1].shape and np.array_equal(x[1], y[1]) and \
                np.array_equal(x[3], y[3])
            if ok and sp["kind"] == "chain" and x[0] == abi.KG_OK:
                # tuple count and MIN int exact; the MIN strings' bytes equal
                ok = x[2][0, 0] == y[2][0, 0] and (x[3][0, 1] or x[2][0, 1] == y[2][0, 1]) and \
                    (x[3][0, 2] or chain_text(x[2][0, 2]) == chain_text(y[2][0, 2]))
            if not ok and len(examples) < 8:
                examples.append({"query": i, "template": sp["template"], "table": sp["table"],
                                 "status": [int(x[0]), int(y[0])], "groups": [len(x[1]), len(y[1])]})
            for j, ag in enumerate(sp["aggs"]):
                if not ok:
                    break
                real = ag[0] == abi.KG_AGG_AVG or (ag[0] == abi.KG_AGG_SUM and (
                    (len(ag) > 2 and ag[2] >= 0) or FACT_COLS[ag[1]].startswith("m")))
                if real:
                    u, v = x[2][:, j].view(np.float64), y[2][:, j].view(np.float64)
                    ok = bool(np.all(np.abs(u - v) <= 1e-12 * np.maximum(np.abs(u), np.abs(v)) + 0.0))
                else:
                    ok = np.array_equal(x[2][:, j], y[2][:, j])
                if not ok and len(examples) < 8:
                    examples.append({"query": i, "template": sp["template"], "table": sp["table"], "agg": j,
                                     "plain": x[2][:3, j].tolist(), "batched": y[2][:3, j].tolist()})
            bad += 0 if ok else 1
        return bad

    def best_rate(mode, start, bar, cap=0, factor=1.25, limit=64, refine=0):
        """Highest offered rate (start * factor^k, then `refine` geometric
        bisection steps below the first failing rate) whose p99 stays within
        bar."""
        best, runs, rate = None, [], start
        failed = None
        while rate <= start * limit:
            r = run(mode, rate, cap)
            runs.append(r)
            if r["p99_ms"] > bar or r["qps"] < 0.9 * rate:
                failed = rate
                break
            best = r
            rate *= factor
        lo = best["offered"] if best else None
        for _ in range(refine if (best and failed) else 0):
            mid = (lo * failed) ** 0.5
            r = run(mode, mid, cap)
            runs.append(r)
            if r["p99_ms"] > bar or r["qps"] < 0.9 * mid:
                failed = mid
            else:
                best, lo = r, mid
        return best, runs

    run(abi.KG_SERVE_BATCHED, None)
    if a.serve_rate_only > 0:  # one batched pass at a given rate (for kernel traces)
        r = run(abi.KG_SERVE_BATCHED, a.serve_rate_only)
        log(f"batched at {a.serve_rate_only:.0f} q/s offered: {r}")
        return {"batched_at_rate": r}
    if a.serve_saturation_only:  # one batched pass of the mix (for kernel traces)
        batched_sat = run(abi.KG_SERVE_BATCHED, None)
        log(f"saturation: batched {batched_sat['qps']:.0f} q/s")
        return {"batched_saturation": batched_sat}
    run(abi.KG_SERVE_PLAIN, None)
    plain_sat, plain_res = run(abi.KG_SERVE_PLAIN, None, keep=True)
    batched_sat, batched_res = run(abi.KG_SERVE_BATCHED, None, keep=True)
    examples = []
    mismatched = same_results(plain_res, batched_res)
    log(f"saturation: plain {plain_sat['qps']:.0f} q/s, batched {batched_sat['qps']:.0f} q/s; "
        f"{mismatched} of {nq} results differ between the two")
    base = 0.7 * plain_sat["qps"]
    plain_at = run(abi.KG_SERVE_PLAIN, base)
    bar = plain_at["p99_ms"]
    # The run that sets the bar meets it; higher plain rates are tried from there.
    more, plain_runs = best_rate(abi.KG_SERVE_PLAIN, base * 1.1, bar, factor=1.1, limit=2, refine=2)
    plain_best = more if (more and more["qps"] > plain_at["qps"]) else plain_at
    plain_runs = [plain_at] + plain_runs
    batched = []
    for cap in a.serve_caps:
        bb, runs = best_rate(abi.KG_SERVE_BATCHED, base, bar, cap=cap, refine=3)
        batched.append({"cap_rows": cap, "best": bb, "runs": runs})
        log(f"batched cap {cap}: best {bb['qps'] if bb else None} q/s within p99 {bar:.2f} ms")
    cands = [x["best"] for x in batched if x["best"]]
    bbest = max(cands, key=lambda r: r["qps"]) if cands else None
    ratio = (bbest["qps"] / plain_best["qps"]) if (bbest and plain_best) else None
    return {
        "workload": f"{nq} queries from {len(SERVE_TEMPLATES)} templates {list(SERVE_TEMPLATES)} over fact tables "
                    f"of {list(SERVE_FACT_SIZES)} rows (uniform), Poisson arrivals",
        "p99_bar_ms": bar,
        "plain_saturation": plain_sat, "batched_saturation": batched_sat,
        "plain_best_within_bar": plain_best, "plain_runs": plain_runs,
        "batched_best_within_bar": bbest, "batched_by_cap": batched,
        "qps_ratio_at_equal_p99": ratio,
        "serving_config": {"slots": a.serve_slots, "lanes": a.serve_lanes, "max_batch": a.serve_max_batch,
                           "arena_mib": a.serve_arena, "lane_scan_rows": a.serve_lane_scan_rows,
                           "GPU_MAX_HW_QUEUES": os.environ.get("GPU_MAX_HW_QUEUES")},
        "results_plain_vs_batched": {"queries": nq, "mismatched": mismatched, "examples": examples,
                                     "bar": "status, group keys/order, counts and integer sums exact; FP64 "
                                            "within 1e-12 relative"},
        "saturation_qps_ratio": batched_sat["qps"] / plain_sat["qps"],
    }


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--copy-calibration":
        return copy_calibration()
    if len(sys.argv) > 1 and sys.argv[1] == "--pmc-summary":
        return pmc_summary(*sys.argv[2:6])
    if len(sys.argv) > 1 and sys.argv[1] == "--pmc-query-summary":
        return pmc_query_summary(*sys.argv[2:6])
    if len(sys.argv) > 1 and sys.argv[1] == "--pmc-counters":
        return pmc_counters(*sys.argv[2:])
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU sharding is a design note, "
              "not a launched configuration.", file=sys.stderr)
        return 2
    # Hardware queues of this process, read when the HIP runtime starts: batched serving keeps the
    # context stream, its batch slots and a plain lane busy at once, and with HIP's default of 4
    # queues some of those streams share a queue and wait for each other's work. Single-stream
    # workloads (every other measurement, and the plain serving path) do not depend on it.
    if a.serve_hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(a.serve_hw_queues)
    import torch
    from kushdb_amd.runtime import Context
    torch.cuda.init()
    ctx = Context(0)
    if a.only:
        run_one = {"c0_star": bench_star_c0, "c2_join": bench_join, "c3_chain": bench_chain_c3,
                   "serving": bench_serving, "c2_build": lambda c, x: bench_c2_op(c, x, "c2_build"),
                   "c2_probe_agg": lambda c, x: bench_c2_op(c, x, "c2_probe_agg")}[a.only]
        print(json.dumps({a.only: run_one(ctx, a)}), flush=True)
        return 0
    t, wall, kms, groups = bench_scan_agg(ctx, a)
    n = a.rows
    rows_per_s = n * a.steps / wall
    result = {
        "metric": METRIC,
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "scan-select (day <= cutoff, ~98%) + group-by 2 enum keys (6 groups): "
                               "SUM/AVG of 4 f64 + 3 int32 measures, COUNT(*)",
                   "rows": n, "groups": groups, "bytes_per_row": SCAN_AGG_BYTES_PER_ROW},
        "roofline": dict(roofline(n * SCAN_AGG_BYTES_PER_ROW, kms, measured_traffic("k_group_agg_v4")),
                         kernel="k_group_agg_v4", kernel_ms=round(kms, 4)),
    }
    if not a.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        cps, cdt, cn = cpu_baseline_scan_agg(t, min(a.cpu_rows, n))
        result["cpu_baseline"] = {"value": cps, "unit": "rows/s", "cores": 1, "kind": "port",
                                  "sample": f"first {cn} rows of the same table, {cdt:.2f} s, "
                                            f"{platform.processor() or platform.machine()}"}
    del t
    torch.cuda.empty_cache()
    result["extra"] = {}
    if not a.no_join:
        result["extra"].update(bench_join(ctx, a))
    if not a.no_join:
        result["extra"]["c0_star"] = bench_star_c0(ctx, a)
        result["extra"]["c3_chain"] = bench_chain_c3(ctx, a)
        torch.cuda.empty_cache()
    if not a.no_serve:
        result["extra"]["serving"] = bench_serving(ctx, a)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
