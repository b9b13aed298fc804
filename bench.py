"""Benchmark: embedding-extension mining on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2] [--scale 1.0]

A step mines every query of the workload once on the plain path (one query
at a time, phase by phase through libgeneva_hip).  The default workload is the
single large graph: Chung-Lu (exponent 2.1, 1M vertices, ~10M edges, 16
labels Zipf(1.2)) with five 5-vertex shapes.  Prints one JSON line:
value = embeddings per second over the timed steps (graph resident in HBM
before timing), the per-kernel roofline of the dominant kernel (HIP events on
the context's stream), the oracle's CPU rate on a bounded sample, and the
parity of every count against the oracle's golden counts.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ds, queries, budget_s, threads):
    """Oracle count mode on a strided sample of start vertices per query."""
    import oracle as O
    og = O.OracleGraph.of(ds)
    per_q = budget_s / max(1, len(queries))
    tot_count, tot_secs, samples = 0, 0.0, []
    for q in queries:
        # a sparse probe, then the stride that fills this query's share of the budget
        stride = 8192
        r = og.mine(q, threads=threads, start_begin=stride // 2, stride=stride)
        if r.seconds < per_q / 2:
            est = max(r.seconds, 1e-3) * stride / max(per_q, 1e-3)
            stride = max(1, min(stride, int(est) + 1))
            r = og.mine(q, threads=threads, start_begin=stride // 2, stride=stride)
        tot_count += r.count
        tot_secs += r.seconds
        samples.append(f"1/{stride} of {r.nstart * stride} starts")
    return {"value": tot_count / tot_secs if tot_secs > 0 else 0.0, "unit": "embeddings/s",
            "cores": threads, "kind": "port",
            "sample": "oracle count mode (cpumine semantics), every n-th start vertex per query: "
                      + "; ".join(samples) + f"; {tot_secs:.1f}s of CPU work"}


def copy_calibration(mib):
    """Run the library's overlap-safe copy kernel on a known size: 4n bytes read
    and 4n bytes written per element count n, in launches of n/2 elements, so
    PMC readings of FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    import ctypes as C
    import geneva_amd as G
    hip = C.CDLL(G.HIP_LIB)
    ctx = C.c_void_p()
    assert hip.gv_ctx_create(0, C.byref(ctx)) == 0
    n = (mib << 20) // 4
    buf = C.c_void_p()
    assert hip.gv_malloc(ctx, C.c_uint64(4 * (n + n // 2)), C.byref(buf)) == 0
    src = buf.value
    dst = src + 4 * (n // 2)
    hip.gv_copy_partial_emb.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    for _ in range(3):
        assert hip.gv_copy_partial_emb(ctx, src, dst, n) == 0
    assert hip.gv_ctx_synchronize(ctx) == 0
    print(json.dumps({"copy_calibration": {"elements": n, "launches_per_copy": 2,
                                           "bytes_read_per_launch": 4 * (n // 2),
                                           "bytes_written_per_launch": 4 * (n // 2), "copies": 3}}))
    hip.gv_free(ctx, buf)
    hip.gv_ctx_destroy(ctx)


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=float), q)) if len(xs) else 0.0


def batched_and_served(args):
    """Workload 3 (database of 100,000 graphs, 256 queries of 4-8 vertices):
    plain path vs batched engine (autotuned, captured iteration graph); and
    workload 5 (the same database, Poisson arrivals, queries of 3-10 vertices):
    continuous batching vs the plain one-at-a-time server."""
    import numpy as np
    import geneva_amd as G
    from geneva_amd import configs
    t0 = time.perf_counter()
    ds = configs.workload3(scale=args.db_scale)
    eng = G.Engine(ds, device=0, level_bytes=1 << 30)
    # 256 slots, two 256 MiB tables each (128 GiB of the 288 GiB HBM)
    eng.batch_config(slots=256, slot_bytes=args.slot_mib << 20, row_cap=1 << 28)
    qs = ds.queries
    log(f"workload 3 scale {args.db_scale}: nv={ds.nv} graphs={ds.ngraphs} queries={len(qs)} "
        f"ready in {time.perf_counter() - t0:.1f}s")
    tune = eng.autotune(qs[:96])
    log(f"autotune: {tune}")
    # plain path: one query at a time
    for q in qs[:8]:
        eng.mine(q)
    tp = time.perf_counter()
    plain = [eng.mine(q) for q in qs]
    tp = time.perf_counter() - tp
    # batched: all queries in one engine run (one untimed run first touches
    # every slot table)
    eng.mine_batch(qs)
    tb = time.perf_counter()
    reps = max(1, args.batch_reps)
    for _ in range(reps):
        batched, st = eng.mine_batch(qs)
    tb = (time.perf_counter() - tb) / reps
    same = all(a.count == b.count and a.graph_ids == b.graph_ids for a, b in zip(plain, batched))
    emb = sum(r.count for r in batched)
    out = {"workload": "workload3: 100,000 graphs n~U[20,200] ER/BA(m=2), 32 labels Zipf(1.1), "
                       "256 walk queries of 4-8 vertices" if args.db_scale == 1.0 else
                       f"workload3 scale {args.db_scale}",
           "queries": len(qs), "embeddings": emb,
           "plain_queries_per_s": round(len(qs) / tp, 2), "plain_embeddings_per_s": round(emb / tp, 1),
           "batched_queries_per_s": round(len(qs) / tb, 2), "batched_embeddings_per_s": round(emb / tb, 1),
           "speedup_vs_plain": round(tp / tb, 2), "counts_and_graph_ids_equal_plain": same,
           "overflowed_to_plain": int(st["overflowed"]), "graph_launches": int(st["graph_launches"]),
           "tune": tune}
    # serving sweep over workload 3's database
    serv = []
    rng = np.random.default_rng(configs.seed(5))
    for i, qps in enumerate(args.qps):
        n = int(min(args.serve_requests, max(200, qps * 1.0)))
        reqs = configs.workload5_queries(ds, n, seed_offset=i)
        arr = np.cumsum(rng.exponential(1.0 / qps, size=n))
        arr = (arr - arr[0]).tolist()
        lb, cb, sb = eng.serve(reqs, arr, batched=True, mem_budget=args.mem_budget << 30)
        lp, cp, sp = eng.serve(reqs, arr, batched=False)
        serv.append({"offered_qps": qps, "requests": n,
                     "batched": {"queries_per_s": round(n / sb["wall_s"], 1), "p50_ms": round(pct(lb, 50), 3),
                                 "p99_ms": round(pct(lb, 99), 3), "max_in_flight": int(sb["max_concurrent"]),
                                 "iterations": int(sb["iterations"]), "to_plain": int(sb["overflowed"]),
                                 "direct_plain": int(sb["direct"])},
                     "plain": {"queries_per_s": round(n / sp["wall_s"], 1), "p50_ms": round(pct(lp, 50), 3),
                               "p99_ms": round(pct(lp, 99), 3)},
                     "counts_equal": cb == cp})
        log(f"serve qps={qps}: {serv[-1]}")
    out_s = {"workload": "workload5: workload3 database resident, Poisson arrivals, walk queries of 3-10 vertices",
             "points": serv}
    eng.close()
    return out, out_s


# kernel (lean variant, as rocprofv3 names it) -> the kind bench.py reports
PMC_KINDS = [("k_walk<true, false, 2>", "k_walk<count>"), ("k_walk<true, false, 4>", "k_walk<count>"), ("k_walk<false, false, 4>", "k_walk<extend>"),
             ("k_walk<false, false, 2>", "k_walk<extend>"),
             ("k_size<false>", "k_size"), ("k_count_rows<false>", "k_count_rows"),
             ("k_count_leaves<false>", "k_count_rows"), ("k_count_chain<false>", "k_count_rows"),
             ("k_chain_mark", "k_count_rows"), ("k_chain_weights", "k_count_rows"),
             ("k_scan_tiles", "k_scan"), ("k_scan_final", "k_scan"),
             ("k_child_size", "k_isect"), ("k_mark_parents", "k_isect"), ("k_isect", "k_isect"),
             ("k_isect_big", "k_isect")]


def pmc_summary(paths, workload, cal_mib):
    """HBM bytes per launch per kernel kind from two rocprofv3 --pmc runs of the
    bench (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch) and the same two
    counters over k_copy_back moving a known number of bytes each way
    (--copy-calibration cal_mib): each counter is scaled by bytes moved / KiB
    read for the copy.  Writes profiles/traffic_workload<N>.json and prints the
    table."""
    import csv
    import re
    fetch_csv, write_csv, cal_fetch_csv, cal_write_csv = paths

    def per_kernel(path, counter):
        acc = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            m = re.search(r"(k_\w+(?:<[^>]*>)?)\(", r["Kernel_Name"].replace("(anonymous namespace)::", ""))
            if m:
                d = acc.setdefault(m.group(1), {})
                d[r["Dispatch_Id"]] = d.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
        return acc

    def mean(d):
        return sum(d.values()) / len(d)

    moved = 4 * (((cal_mib << 20) // 4) // 2)   # bytes each way per k_copy_back launch
    fetch_scale = moved / (mean(per_kernel(cal_fetch_csv, "FETCH_SIZE")["k_copy_back"]) * 1024)
    write_scale = moved / (mean(per_kernel(cal_write_csv, "WRITE_SIZE")["k_copy_back"]) * 1024)
    f, w = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    out = {"fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
           "calibration": f"k_copy_back: {moved} B read + {moved} B written per launch"}
    kinds = {}
    for kern, kind in PMC_KINDS:
        if kern not in f or kern not in w:
            continue
        k = kinds.setdefault(kind, {"kernels": [], "fetch_kib": 0.0, "write_kib": 0.0, "launches": 0})
        k["kernels"].append(f"{kern} x{len(f[kern])}")
        k["fetch_kib"] += sum(f[kern].values())
        k["write_kib"] += sum(w[kern].values())
        # a launch of the scan kind is its two kernels; the row counters are
        # counted by their counting kernel (the chain's marking and weights
        # belong to the chain count)
        if kern not in ("k_scan_tiles", "k_mark_parents", "k_isect", "k_isect_big") \
                and not kern.startswith("k_chain_"):
            k["launches"] += len(f[kern])
    print("| Kind | Kernels (dispatches) | FETCH_SIZE KiB / launch | WRITE_SIZE KiB / launch | HBM bytes / launch |")
    print("|---|---|---|---|---|")
    for kind, k in kinds.items():
        n = max(k["launches"], 1)
        fk, wk = k["fetch_kib"] / n, k["write_kib"] / n
        hb = int(fk * 1024 * fetch_scale + wk * 1024 * write_scale)
        out[kind] = {"kernels": k["kernels"], "launches": n, "fetch_kib_per_launch": round(fk, 1),
                     "write_kib_per_launch": round(wk, 1), "hbm_bytes_per_launch": hb}
        print(f"| `{kind}` | {', '.join(k['kernels'])} | {fk:.1f} | {wk:.1f} | {hb:,} |")
    path = os.path.join(ROOT, "profiles", f"traffic_workload{workload}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"scales: fetch {fetch_scale:.4f}, write {write_scale:.4f}; wrote {path}")


def pmc_table(path):
    """Per-kernel means of every counter in one rocprofv3 --pmc CSV (markdown)."""
    import csv
    import re
    acc, dur = {}, {}
    for r in csv.DictReader(open(path)):
        name = r["Kernel_Name"].replace("(anonymous namespace)::", "")
        m = re.search(r"(k_\w+(?:<[^>]*>)?)\(", name)
        name = m.group(1) if m else name[:40]
        d = acc.setdefault(name, {}).setdefault(r["Counter_Name"], {})
        d[r["Dispatch_Id"]] = d.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
        dur.setdefault(name, {})[r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    cols = sorted({c for k in acc.values() for c in k})
    print("| Kernel | Dispatches | us / dispatch | " + " | ".join(cols) + " |")
    print("|---|---|---|" + "---|" * len(cols))
    for name in sorted(acc, key=lambda k: -sum(dur[k].values())):
        n = len(dur[name])
        vals = [sum(acc[name].get(c, {}).values()) / n for c in cols]
        print(f"| `{name}` | {n} | {sum(dur[name].values()) / n / 1e3:.1f} | " +
              " | ".join(f"{v:,.0f}" for v in vals) + " |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--level-bytes", type=int, default=1 << 31)
    ap.add_argument("--no-batch", action="store_true", help="skip the workload 3/5 sections")
    ap.add_argument("--db-scale", type=float, default=1.0)
    ap.add_argument("--batch-reps", type=int, default=2)
    ap.add_argument("--qps", type=float, nargs="*", default=[200, 1000, 5000])
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--mem-budget", type=int, default=32, help="GiB of estimated tables in flight")
    ap.add_argument("--slot-mib", type=int, default=256, help="MiB per batch slot table")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel on MIB mebibytes (PMC scaling)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV", "CAL_WRITE_CSV"),
                    help="only summarise rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the bench and of "
                         "--copy-calibration into profiles/traffic_workload<N>.json (no GPU)")
    ap.add_argument("--cal-mib", type=int, default=512, help="MIB of the calibration run (--pmc-summary)")
    ap.add_argument("--pmc-table", metavar="CSV", help="only print per-kernel means of a rocprofv3 --pmc CSV")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.pmc_summary:
        return pmc_summary(args.pmc_summary, args.workload, args.cal_mib)
    if args.pmc_table:
        return pmc_table(args.pmc_table)
    if args.gpus != 1:
        log("bench.py measures one GPU: the embedding-extension path here is built and measured "
            "on a single MI355X (multi-GPU sharding is a design note, not part of this build).")
        sys.exit(2)

    import geneva_amd as G
    from geneva_amd import configs

    t0 = time.perf_counter()
    ds = configs.WORKLOADS[args.workload](scale=args.scale)
    queries = ds.queries
    log(f"workload {args.workload} scale {args.scale}: nv={ds.nv} edges={len(ds.edges)} "
        f"queries={len(queries)} generated in {time.perf_counter() - t0:.1f}s")
    t1 = time.perf_counter()
    eng = G.Engine(ds, device=0, level_bytes=args.level_bytes)
    log(f"graph resident: {eng.graph_bytes / 2**20:.1f} MiB in {time.perf_counter() - t1:.1f}s")

    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", f"workload{args.workload}_full.json")
    if args.scale == 1.0 and os.path.exists(gpath):
        golden = {c["query"]: c for c in json.load(open(gpath))["cases"]}

    for w in range(args.warmup):
        for i, q in enumerate(queries):
            r = eng.mine(q)
            log(f"warmup {w} query {i}: count={r.count} device_ms={r.ms:.2f}")

    # timed steps: nothing but the mining (lean kernels, no events)
    results = []
    t2 = time.perf_counter()
    for s in range(args.steps):
        step = []
        for q in queries:
            step.append(eng.mine(q))
        results.append(step)
    t3 = time.perf_counter()
    for s, step in enumerate(results):
        log(f"step {s}: {sum(r.count for r in step)} embeddings")
    # the same steps again with every launch bracketed by HIP events on the
    # context's stream: per-kernel device time (the events add their own gaps
    # between launches, not to the kernels)
    eng.profile(True, count_bytes=False)
    for s in range(args.steps):
        for q in queries:
            eng.mine(q)
    prof = eng.profile_read()
    # one more step with the counting kernel variants: requested bytes and work
    # units per step of every kernel kind (the work is the same every step)
    eng.profile(True, count_bytes=True)
    for q in queries:
        eng.mine(q)
    acct = eng.profile_read()
    eng.profile(False)

    ms_per_step = (t3 - t2) * 1000.0 / args.steps
    emb_per_step = sum(r.count for r in results[0])
    value = emb_per_step / (ms_per_step / 1000.0)

    # parity: every step identical, and equal to the oracle's golden counts
    stable = all([r.count for r in st] == [r.count for r in results[0]] for st in results)
    parity = {"steps_identical": stable}
    if golden is not None:
        ok = [golden[q]["count"] == r.count and golden[q]["alg_bytes"] == r.alg_bytes
              for q, r in zip(queries, results[0])]
        parity.update({"checked": len(ok), "bit_exact": sum(ok), "against": "oracle golden counts"})
    else:
        parity.update({"checked": 0, "against": "none at this scale"})

    # roofline of the dominant kernel: its launches and device time over the
    # timed region (HIP events), its requested bytes per step from the
    # accounting step
    name = max(prof.items(), key=lambda kv: kv[1][1])[0]
    launches, total_ms = prof[name][0], prof[name][1]
    step_launches, _, req_step, units_step, b_step = acct[name]
    avg_ms = total_ms / max(launches, 1)
    ms_step = total_ms / args.steps
    achieved = req_step / (ms_step / 1e3) / 1e9 if ms_step > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_workload{args.workload}.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        traffic = t.get(name, {}).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "launches": launches, "avg_launch_ms": round(avg_ms, 4),
                "alg_bytes_per_launch": req_step // max(step_launches, 1),
                "units_per_launch": units_step // max(step_launches, 1),
                # the reference figure B (every required list charged in full,
                # including rows counted without enumeration) for comparison
                "b_bytes_per_launch": b_step // max(step_launches, 1),
                "b_rate_gbps": round(b_step / (ms_step / 1e3) / 1e9, 1) if ms_step > 0 else 0.0}
    kernels = {k: {"launches": v[0], "ms": round(v[1], 3), "req_bytes_per_step": acct[k][2],
                   "units_per_step": acct[k][3], "b_bytes_per_step": acct[k][4]}
               for k, v in prof.items()}

    eng.close()
    batched = served = None
    if not args.no_batch:
        batched, served = batched_and_served(args)

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(ds, queries, args.cpu_seconds, threads)

    line = {
        "metric": "embeddings/s (plain GPU path, one query per launch sequence)",
        "value": round(value, 1),
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"workload{args.workload}: Chung-Lu gamma=2.1, 1M vertices, ~10M edges, "
                               "16 labels Zipf(1.2), 5 shapes x 5 vertices" if args.workload == 2 else
                               f"workload{args.workload}",
                   "scale": args.scale, "queries": len(queries), "path": "plain",
                   "embeddings_per_step": emb_per_step},
        "queries_per_s": round(len(queries) / (ms_per_step / 1000.0), 3),
        "per_query": [{"count": r.count, "device_ms": round(r.ms, 3), "launches": r.launches}
                      for r in results[0]],
        "alg_bytes_ref": sum(r.alg_bytes for r in results[0]),
        "parity": parity,
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
