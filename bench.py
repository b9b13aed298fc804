"""Benchmark: embedding-extension mining on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline (metric "embeddings/s and queries/s (p50/p99 latency) vs plain
GPU"): workload 3, a database of 100,000 labelled graphs (n ~ U[20,200],
Erdos-Renyi / Barabasi-Albert, 32 labels Zipf(1.1)) resident in HBM, with 256
random-walk queries of 4-8 vertices.  A step is one run of the batched engine
over all 256 queries (captured iteration graph, device work queue, tuning
picked by the autotuner); value = embeddings per second over the K timed
steps.  Beside it: the plain path (one query per launch sequence) on the same
queries, their ratio, parity of every count / byte total / matched-graph set
against the oracle's full-size fixture, the roofline of the batch engine's
dominant kernel (its algorithmic bytes B over its HIP-event time), the
workload-5 serving sweep (Poisson arrivals, p50/p99), and the workload-2 and
workload-4 sections (single large graphs, plain path), and the oracle's CPU
rates on bounded samples.  Prints one JSON line.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ids_sha256(ids):
    return hashlib.sha256("\n".join(map(str, sorted(ids))).encode()).hexdigest()


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=float), q)) if len(xs) else 0.0


def host_threads():
    return min(os.cpu_count() or 1, 16)


def oracle_sample(og, queries, budget_s, threads, max_stride=8192):
    """Oracle count mode (cpumine semantics) on every n-th start vertex of
    each query: n starts at max_stride and halves (the sample doubles) while
    the last sample of the query took under a quarter of its share of
    budget_s; the rate is that of the last sample of every query."""
    per_q = budget_s / max(1, len(queries))
    tot_count, tot_secs, samples = 0, 0.0, []
    for q in queries:
        stride = max_stride
        r = og.mine(q, threads=threads, start_begin=stride // 2, stride=stride)
        while stride > 1 and r.seconds < per_q / 4:
            stride //= 2
            r = og.mine(q, threads=threads, start_begin=stride // 2, stride=stride)
        tot_count += r.count
        tot_secs += r.seconds
        samples.append(stride)
    return tot_count, tot_secs, samples


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


def gather_ceiling(n=100_000_000, sizes_mib=(16, 64, 256, 800, 1450, 2900)):
    """Rate of independent random 4-byte reads the GPU sustains from tables
    of several sizes (torch's index_select: n random positions, each read
    once), in reads per second: the ceiling for kernels whose time goes to
    scattered single-word reads (label offsets, first list entries)."""
    import torch
    out = []
    for mib in sizes_mib:
        words = (mib << 20) // 4
        x = torch.ones(words, dtype=torch.int32, device="cuda")
        idx = torch.randint(0, words, (n,), device="cuda", dtype=torch.int64)
        for _ in range(2):
            torch.index_select(x, 0, idx)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(5):
            torch.index_select(x, 0, idx)
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / 5
        out.append({"table_mib": mib, "reads": n, "ms": round(ms, 3), "reads_per_s": round(n / ms * 1e3)})
        log(f"gather from {mib} MiB: {n / ms / 1e6:.1f} G reads/s")
        del x, idx
        torch.cuda.empty_cache()
    print(json.dumps({"gather_ceiling": out}), flush=True)


# kernel (lean variant, as rocprofv3 names it) -> the kind bench.py reports
PMC_KINDS = [(f"k_b_walk<{i}, {c}>", "k_b_walk") for i in (1, 2, 4, 8) for c in ("false", "true")] + [
             ("k_b_size", "k_b_size"), ("k_b_tscan", "k_b_scan"),
             ("k_walk<true, false, 2>", "k_walk<count>"), ("k_walk<true, false, 4>", "k_walk<count>"), ("k_walk<false, false, 4>", "k_walk<extend>"),
             ("k_walk<false, false, 2>", "k_walk<extend>"),
             ("k_size<false>", "k_size"), ("k_count_rows<false>", "k_count_rows"),
             ("k_count_leaves<false>", "k_count_rows"), ("k_count_chain<false>", "k_count_rows"),
             ("k_chain_mark", "k_count_rows"), ("k_chain_weights", "k_count_rows"),
             ("k_scan_tiles", "k_scan"), ("k_scan_final", "k_scan"),
             ("k_child_size", "k_isect"), ("k_mark_parents", "k_isect"), ("k_isect", "k_isect"),
             ("k_isect_big", "k_isect")]


def pmc_summary(paths, workload, cal_mib, last=0):
    """HBM bytes per launch per kernel kind from two rocprofv3 --pmc runs of the
    bench (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch) and the same two
    counters over k_copy_back moving a known number of bytes each way
    (--copy-calibration cal_mib): each counter is scaled by bytes moved / KiB
    read for the copy.  last > 0: only the last `last` dispatches of each
    kernel (the bench's accounting run).  Writes
    profiles/traffic_workload<N>.json and prints the table."""
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
    if last:
        # the accounting run: every dispatch from its first iteration's k_b_prep on
        def keep_last(acc, path):
            preps = []
            for r in csv.DictReader(open(path)):
                if "k_b_prep" in r["Kernel_Name"]:
                    preps.append(int(r["Dispatch_Id"]))
            preps = sorted(set(preps))
            start = preps[-last] if len(preps) >= last else 0
            out = {k: {i: v for i, v in d.items() if int(i) >= start} for k, d in acc.items()}
            return {k: d for k, d in out.items() if d}
        f, w = keep_last(f, fetch_csv), keep_last(w, write_csv)
    out = {"fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
           "calibration": f"k_copy_back: {moved} B read + {moved} B written per launch",
           "dispatches": f"from the last {last} batch iterations on" if last else "all"}
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
        # belong to the chain count); a batch walk launch is its extension and
        # its counting kernel
        if kern not in ("k_scan_tiles", "k_mark_parents", "k_isect", "k_isect_big") \
                and not kern.startswith("k_chain_") \
                and not (kern.startswith("k_b_walk<") and kern.endswith("true>")):
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




# ---------------------------------------------------------------------------
# workload 3 (headline): batched engine vs plain path


def section_w3(args):
    import numpy as np  # noqa: F401
    import geneva_amd as G
    from geneva_amd import configs
    t0 = time.perf_counter()
    ds = configs.workload3(scale=args.db_scale)
    eng = G.Engine(ds, device=0, level_bytes=1 << 30)
    # 264 slots (the 256 queries and the pieces of any too large for one
    # slot), two tables of slot_mib each
    eng.batch_config(slots=264, slot_bytes=args.slot_mib << 20, row_cap=1 << 28)
    qs = ds.queries
    log(f"workload 3 scale {args.db_scale}: nv={ds.nv} graphs={ds.ngraphs} queries={len(qs)} "
        f"graph {eng.graph_bytes / 2**20:.0f} MiB ready in {time.perf_counter() - t0:.1f}s")
    tune = eng.autotune(qs[:96])
    log(f"autotune: {tune}")
    for w in range(args.warmup):
        eng.mine_batch(qs)
        for q in qs[:16]:
            eng.mine(q)
    # timed steps: one batched run of all queries each
    runs = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        runs.append(eng.mine_batch(qs))
    t2 = time.perf_counter()
    ms_step = (t2 - t1) * 1e3 / args.steps
    batched = runs[0][0]
    st = runs[0][1]
    emb = sum(r.count for r in batched)
    # plain path: one query per launch sequence, the same queries
    tp = time.perf_counter()
    plain = [eng.mine(q) for q in qs]
    tp = (time.perf_counter() - tp) * 1e3
    stable = all([r.count for r in rs] == [r.count for r in batched] for rs, _ in runs)
    same = all(a.count == b.count and a.graph_ids == b.graph_ids and a.alg_bytes == b.alg_bytes
               for a, b in zip(plain, batched))
    parity = {"steps_identical": stable, "batched_equals_plain": same}
    gpath = os.path.join(ROOT, "tests", "golden", "workload3_full.json")
    if args.db_scale == 1.0 and os.path.exists(gpath):
        gold = {c["query"]: c for c in json.load(open(gpath))["cases"]}
        ok = [gold[q]["count"] == r.count and gold[q]["alg_bytes"] == r.alg_bytes and
              gold[q]["ngraphs"] == len(r.graph_ids) and gold[q]["graph_ids_sha256"] == ids_sha256(r.graph_ids)
              for q, r in zip(qs, batched)]
        parity.update({"checked": len(ok), "bit_exact": sum(ok),
                       "against": "oracle fixture tests/golden/workload3_full.json (counts, bytes, graph-id sets)"})
    # one accounting run: every kernel of the iteration launched on its own
    # under HIP events, and the bytes B charged to k_b_size / k_b_walk
    shared = dict(zip(("shared", "restarted"), eng.batch_share_stats()))
    eng.batch_profile(True)
    eng.profile(True, count_bytes=False)
    acct_res, acct_st = eng.mine_batch(qs)
    prof = eng.profile_read()
    b_size, b_walk = eng.batch_profile_bytes()
    eng.profile(False)
    eng.batch_profile(False)
    kinds = {"k_b_size": prof["k_size"], "k_b_scan": prof["k_scan"], "k_b_walk": prof["k_walk<extend>"],
             "k_b_prep/tiles/post": prof["batch_iteration"]}
    b_total = sum(r.alg_bytes for r in acct_res)
    dev_ms = sum(v[1] for v in kinds.values())
    tpath = os.path.join(ROOT, "profiles", "traffic_workload3.json")
    tjs = json.load(open(tpath)) if os.path.exists(tpath) else {}

    def kernel_roofline(name, b_bytes, charged):
        launches, ms = kinds[name][0], kinds[name][1]
        avg_ms = ms / max(launches, 1)
        b_launch = b_bytes // max(launches, 1)
        achieved = b_launch / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
        return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": tjs.get(name, {}).get("hbm_bytes_per_launch"),
                "launches": launches, "avg_launch_ms": round(avg_ms, 4), "total_ms": round(ms, 3),
                "b_bytes_per_launch": b_launch, "b_charged": charged}

    rs = [kernel_roofline("k_b_size", b_size, "each row's read (4L), each required list (8 + 4|N|) and the "
                                              "B of the rows it counts itself (count modes, leaf groups)"),
          kernel_roofline("k_b_walk", b_walk, "the survivors' writes (4(L+1) each) and, in a counted step, "
                                              "the lists its fused last vertex searches")]
    # the dominant kernel (most device time per step) is the line's roofline;
    # the other is reported beside it
    rs.sort(key=lambda r: -r["total_ms"])
    roofline = dict(rs[0])
    roofline["other"] = rs[1]
    # B the kernels executed (k_b_size + k_b_walk) over the whole iteration's
    # device time; the queries' own B also counts the steps shared between
    # queries once per query and the closed-form first steps of pair starts
    roofline["iteration"] = {"b_bytes": b_size + b_walk, "device_ms": round(dev_ms, 3),
                             "frac": round((b_size + b_walk) / (max(dev_ms, 1e-9) / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                             "b_bytes_of_queries": b_total}
    kern = {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in kinds.items()}
    out = {"embeddings_per_step": emb, "queries": len(qs), "ms_per_step": ms_step,
           "batched_queries_per_s": round(len(qs) / (ms_step / 1e3), 1),
           "plain_ms": round(tp, 3), "plain_queries_per_s": round(len(qs) / (tp / 1e3), 1),
           "plain_embeddings_per_s": round(emb / (tp / 1e3), 1),
           "speedup_vs_plain": round(tp / ms_step, 3),
           "overflowed_to_plain": int(st["overflowed"]), "graph_launches": int(st["graph_launches"]),
           "batch_device_ms": round(st["device_ms"], 3), "batch_wall_ms": round(st["wall_ms"], 3),
           "planning_ms": round(st["compile_ms"], 3), "plain_rerun_ms": round(st["fallback_ms"], 3),
           "pieces": int(st["pieces"]),
           "tune": tune, "parity": parity,
           "shared_tables": shared,
           "roofline": roofline, "kernels": kern,
           "plain_device_ms_sum": round(sum(r.ms for r in plain), 3)}
    return ds, eng, out


def section_served(args, ds, eng):
    """Workload 5: workload 3's database resident, Poisson arrivals of walk
    queries of 3-10 vertices at each offered rate; continuous batching vs the
    plain one-at-a-time server on the same stream; results compared."""
    import numpy as np
    from geneva_amd import configs
    rng = np.random.default_rng(configs.seed(5))
    points, streams = [], {}
    for i, qps in enumerate(args.qps):
        n = int(min(args.serve_max, max(args.serve_min, qps * args.serve_seconds)))
        reqs = configs.workload5_queries(ds, n, seed_offset=i)
        arr = np.cumsum(rng.exponential(1.0 / qps, size=n))
        arr = (arr - arr[0]).tolist()
        lb, cb, sb, gb = eng.serve(reqs, arr, batched=True, mem_budget=args.mem_budget << 30, with_graphs=True)
        lp, cp, sp, gp = eng.serve(reqs, arr, batched=False, with_graphs=True)
        streams[qps] = (reqs, arr, cp, gp)
        points.append({"offered_qps": qps, "requests": n,
                       "batched": {"queries_per_s": round(n / sb["wall_s"], 1), "p50_ms": round(pct(lb, 50), 3),
                                   "p99_ms": round(pct(lb, 99), 3), "max_in_flight": int(sb["max_concurrent"]),
                                   "max_on_batch_engine": int(sb["max_batched"]),
                                   "iterations": int(sb["iterations"]), "overflowed_to_plain": int(sb["overflowed"]),
                                   "plain_side_path": int(sb["direct"])},
                       "plain": {"queries_per_s": round(n / sp["wall_s"], 1), "p50_ms": round(pct(lp, 50), 3),
                                 "p99_ms": round(pct(lp, 99), 3)},
                       "counts_and_graph_ids_equal": cb == cp and gb == gp})
        log(f"serve qps={qps}: {points[-1]}")
    return {"workload": "workload5: the workload-3 database resident in HBM, Poisson arrivals, "
                        "random-walk queries of 3-10 vertices",
            "points": points}, streams


# ---------------------------------------------------------------------------
# workload 2 and workload 4: single large graphs, plain path


def section_w2(args):
    import geneva_amd as G
    from geneva_amd import configs
    ds = configs.workload2()
    eng = G.Engine(ds, device=0, level_bytes=1 << 31)
    qs = ds.queries
    for q in qs:
        eng.mine(q)
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        res = [eng.mine(q) for q in qs]
    ms = (time.perf_counter() - t) * 1e3 / reps
    eng.profile(True, count_bytes=False)
    for q in qs:
        eng.mine(q)
    prof = eng.profile_read()
    eng.profile(True, count_bytes=True)
    for q in qs:
        eng.mine(q)
    acct = eng.profile_read()
    eng.profile(False)
    gold = {}
    gpath = os.path.join(ROOT, "tests", "golden", "workload2_full.json")
    if os.path.exists(gpath):
        gold = {c["query"]: c for c in json.load(open(gpath))["cases"]}
    ok = [q in gold and gold[q]["count"] == r.count and gold[q]["alg_bytes"] == r.alg_bytes
          for q, r in zip(qs, res)]
    # the extension step (sizing, scan and extension walk of the materialised
    # steps): its algorithmic bytes over its device time
    ext_kinds = ("k_size", "k_scan", "k_walk<extend>")
    b_ext = sum(acct[k][4] for k in ext_kinds)
    ms_ext = sum(prof[k][1] for k in ext_kinds)
    emb = sum(r.count for r in res)
    out = {"workload": "workload2: Chung-Lu gamma=2.1, 1M vertices, ~10M edges, 16 labels Zipf(1.2), "
                       "5 shapes x 5 vertices",
           "ms_per_pass": round(ms, 3), "embeddings_per_s": round(emb / (ms / 1e3), 1),
           "parity": {"checked": len(ok), "bit_exact": sum(ok), "against": "oracle fixture workload2_full.json"},
           "per_query": [{"count": r.count, "device_ms": round(r.ms, 4), "launches": r.launches} for r in res],
           "kernels": {k: {"launches": v[0], "ms": round(v[1], 4), "b_bytes": acct[k][4]} for k, v in prof.items()
                       if v[0]},
           "extension_step_roofline": {"b_bytes": b_ext, "device_ms": round(ms_ext, 4),
                                       "frac": round(b_ext / (max(ms_ext, 1e-9) / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}
    eng.close()
    return ds, out


def section_w4(args):
    """Full-size workload 4: every query mined whole on the plain path where
    it finishes within the per-query limit (args.w4_query_cap seconds); a
    query stopped at the limit is reported with the exact count of the part
    mined and that part's share of its first extension's candidates (the
    chunked path, bounded launches).  Queries that finish quickly are also
    mined by the whole-query path to compare."""
    import geneva_amd as G
    from geneva_amd import configs
    t0 = time.perf_counter()
    ds = configs.workload4()
    eng = G.Engine(ds, device=0, level_bytes=1 << 30)
    log(f"workload 4: nv={ds.nv} edges={len(ds.edges)} ready in {time.perf_counter() - t0:.1f}s")
    per = []
    deadline = time.perf_counter() + args.w4_budget
    for i, q in enumerate(ds.queries):
        cap = min(args.w4_query_cap, max(0.2, deadline - time.perf_counter()))
        eng.time_limit(cap)
        tq = time.perf_counter()
        r = eng.mine(q)
        fin, frac = eng.last_progress()
        rec = {"query": i, "k": sum(1 for l in q.splitlines() if l.startswith("v ")), "finished": fin,
               "fraction_done": round(frac, 6), "count": r.count, "alg_bytes": r.alg_bytes,
               "device_ms": round(r.ms, 3), "wall_s": round(time.perf_counter() - tq, 3)}
        eng.time_limit(0)
        if fin and r.ms < 1000:
            whole = eng.mine(q)
            rec["whole_path_equal"] = whole.count == r.count and whole.alg_bytes == r.alg_bytes
            rec["whole_path_device_ms"] = round(whole.ms, 3)
        per.append(rec)
        log(f"workload4 query {i}: {rec}")
    # the configuration's walk queries as they come (pendant vertices on hubs
    # included): the same generator without the two-neighbour condition
    uq = ds.add_walk_queries(64, 6, 8, induced=True, min_extra=2, seed=configs.seed(4) + 100,
                             min_deg=0).queries[-64:]
    uper = []
    udeadline = time.perf_counter() + args.w4u_budget
    for i, q in enumerate(uq):
        eng.time_limit(min(args.w4u_cap, max(0.2, udeadline - time.perf_counter())))
        r = eng.mine(q)
        fin, frac = eng.last_progress()
        uper.append({"query": i, "k": sum(1 for l in q.splitlines() if l.startswith("v ")), "finished": fin,
                     "fraction_done": round(frac, 6), "count": r.count, "device_ms": round(r.ms, 3)})
    eng.time_limit(0)
    ufin = [p for p in uper if p["finished"]]
    log(f"workload 4 unconstrained: {len(ufin)} of {len(uper)} finished")
    # the oracle's full-size shards (8 start vertices each)
    spath = os.path.join(ROOT, "tests", "golden", "workload4_shards.json")
    shard_ok = shard_n = 0
    if os.path.exists(spath):
        for c in json.load(open(spath))["cases"][:16]:
            for sh in c["shards"]:
                r = eng.mine(c["query"], start_begin=sh["begin"], start_end=sh["end"])
                shard_n += 1
                shard_ok += r.count == sh["count"] and r.alg_bytes == sh["alg_bytes"]
    fin = [p for p in per if p["finished"]]
    dev = sum(p.get("whole_path_device_ms", p["device_ms"]) for p in fin)
    out = {"workload": "workload4: Chung-Lu gamma=1.9, 4M vertices, mean degree 30, weights capped at 1e5, "
                       "64 labels Zipf(1.0), 64 induced walk queries of 6-8 vertices, every query vertex on a cycle",
           "query_cap_s": args.w4_query_cap, "budget_s": args.w4_budget,
           "finished": len(fin), "queries": len(per),
           "finished_embeddings": sum(p["count"] for p in fin),
           "finished_device_ms": round(dev, 3),
           "embeddings_per_s_finished": round(sum(p["count"] for p in fin) / max(dev, 1e-9) * 1e3, 1),
           "oracle_shards": {"checked": shard_n, "bit_exact": shard_ok},
           "per_query": per,
           "unconstrained": {"queries": "the same generator without the two-neighbour condition "
                                        "(pendant vertices on hubs included)",
                             "query_cap_s": args.w4u_cap, "budget_s": args.w4u_budget,
                             "finished": len(ufin), "queries_run": len(uper),
                             "finished_embeddings": sum(p["count"] for p in ufin),
                             "finished_device_ms": round(sum(p["device_ms"] for p in ufin), 3),
                             "parity": "scaled: tests/test_gpu_workload4.py against "
                                       "workload4_unconstrained_scaled.json",
                             "per_query": uper}}
    eng.close()
    return ds, out


# ---------------------------------------------------------------------------
# CPU baselines (the oracle: cpumine semantics, on this machine's cores)


def cpu_baselines(args, w3ds, w2ds, w4ds, streams):
    import numpy as np
    import oracle as O
    from geneva_amd import configs
    th = host_threads()
    out = {"nproc": os.cpu_count(), "threads": th}
    # workload 1, the reference algorithm's own shape: phase by phase,
    # single-threaded (faithful), and the depth-first count on all threads
    w1 = configs.workload1()
    og1 = O.OracleGraph.of(w1)
    t = time.perf_counter()
    c = sum(og1.mine(q, mode="faithful").count for q in w1.queries)
    tf = time.perf_counter() - t
    t = time.perf_counter()
    c2 = sum(og1.mine(q, threads=th).count for q in w1.queries)
    tc = time.perf_counter() - t
    out["workload1_faithful_1thread"] = {"embeddings": c, "seconds": round(tf, 4),
                                         "embeddings_per_s": round(c / tf, 1)}
    out["workload1_count"] = {"embeddings": c2, "seconds": round(tc, 4), "embeddings_per_s": round(c2 / tc, 1),
                              "cores": th}
    log(f"cpu workload 1: {out['workload1_faithful_1thread']} {out['workload1_count']}")
    # workload 3 (the headline's workload): bounded sample of its 256 queries
    og3 = O.OracleGraph.of(w3ds)
    w3q = w3ds.queries[:256]
    cnt, secs, strides = oracle_sample(og3, w3q, args.cpu_seconds, th)
    log(f"cpu workload 3 sample: {cnt} embeddings in {secs:.1f}s")
    head = {"value": round(cnt / secs, 1) if secs else 0.0, "unit": "embeddings/s", "cores": th,
            "nproc": os.cpu_count(), "kind": "port",
            "sample": f"oracle count mode (cpumine semantics) on workload 3, every n-th start vertex of each of "
                      f"its {len(w3q)} queries (n = {min(strides)}..{max(strides)}), {secs:.1f}s of CPU work"}
    # workload 5: the 1,000 q/s stream replayed serially by the oracle (all
    # threads per request), first requests only
    if streams:
        qps = 1000 if 1000 in streams else sorted(streams)[0]
        reqs, arr, cp, gp = streams[qps]
        n = min(len(reqs), args.cpu_replay)
        clock, lat, agree = 0.0, [], 0
        for i in range(n):
            r = og3.mine(reqs[i], threads=th)
            clock = max(clock, arr[i]) + r.seconds
            lat.append((clock - arr[i]) * 1e3)
            agree += r.count == cp[i] and r.graph_ids == gp[i]
        out["workload5_serial_replay"] = {"offered_qps": qps, "requests": n, "p50_ms": round(pct(lat, 50), 3),
                                          "p99_ms": round(pct(lat, 99), 3), "agree_with_gpu": agree}
        log(f"cpu workload 5 replay: {out['workload5_serial_replay']}")
    if w2ds is not None:
        og2 = O.OracleGraph.of(w2ds)
        cnt, secs, strides = oracle_sample(og2, w2ds.queries, args.cpu_seconds, th)
        out["workload2_count_sample"] = {"embeddings_per_s": round(cnt / secs, 1) if secs else 0.0,
                                         "seconds": round(secs, 2), "strides": strides}
        log(f"cpu workload 2 sample: {out['workload2_count_sample']}")
    if w4ds is not None:
        # full-size workload 4 on the CPU: the oracle's shards of 8 start
        # vertices (tests/golden/workload4_shards.json) timed single-threaded
        # when they were made; the hub start vertices of most queries take
        # the oracle longer than this run can wait
        spath = os.path.join(ROOT, "tests", "golden", "workload4_shards.json")
        if os.path.exists(spath):
            sh = [s for c in json.load(open(spath))["cases"] for s in c["shards"]]
            secs = sum(s["oracle_seconds"] for s in sh)
            cnt = sum(s["count"] for s in sh)
            out["workload4_oracle_shards"] = {"shards": len(sh), "embeddings": cnt, "seconds_1thread": round(secs, 2),
                                              "embeddings_per_s_1thread": round(cnt / secs, 1) if secs else 0.0}
    return head, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--db-scale", type=float, default=1.0, help="workload 3 scale")
    ap.add_argument("--slot-mib", type=int, default=256, help="MiB per batch slot table")
    ap.add_argument("--qps", type=float, nargs="*", default=[100, 200, 500, 1000, 2000, 5000, 10000])
    ap.add_argument("--serve-seconds", type=float, default=1.5, help="seconds of arrivals per rate")
    ap.add_argument("--serve-min", type=int, default=300)
    ap.add_argument("--serve-max", type=int, default=15000)
    ap.add_argument("--mem-budget", type=int, default=32, help="GiB of estimated tables in flight")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-w2", action="store_true")
    ap.add_argument("--no-w4", action="store_true")
    ap.add_argument("--w4-query-cap", type=float, default=2.0, help="seconds per workload-4 query")
    ap.add_argument("--w4-budget", type=float, default=100.0, help="seconds for all workload-4 queries")
    ap.add_argument("--w4u-cap", type=float, default=1.0, help="seconds per unconstrained workload-4 query")
    ap.add_argument("--w4u-budget", type=float, default=40.0,
                    help="seconds for all unconstrained workload-4 queries")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-replay", type=int, default=150, help="workload-5 requests replayed on the CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel on MIB mebibytes (PMC scaling)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV", "CAL_WRITE_CSV"),
                    help="only summarise rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the bench and of "
                         "--copy-calibration into profiles/traffic_workload<N>.json (no GPU)")
    ap.add_argument("--workload", type=int, default=3, help="workload of --pmc-summary")
    ap.add_argument("--cal-mib", type=int, default=512, help="MIB of the calibration run (--pmc-summary)")
    ap.add_argument("--pmc-last", type=int, default=0,
                    help="--pmc-summary over the dispatches of the last N batch iterations (the accounting run)")
    ap.add_argument("--pmc-table", metavar="CSV", help="only print per-kernel means of a rocprofv3 --pmc CSV")
    ap.add_argument("--gather-ceiling", action="store_true",
                    help="only measure the GPU's rate of random 4-byte reads from tables of 16 MiB - 2.9 GiB")
    ap.add_argument("--only", choices=("w2", "w4"),
                    help="only run the workload-2 or workload-4 section and print its JSON (profiling runs)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.pmc_summary:
        return pmc_summary(args.pmc_summary, args.workload, args.cal_mib, args.pmc_last)
    if args.pmc_table:
        return pmc_table(args.pmc_table)
    if args.gather_ceiling:
        return gather_ceiling()
    if args.gpus != 1:
        log("bench.py measures one GPU: this build runs and measures the embedding-extension path on a "
            "single MI355X; sharding over several GPUs is a design note, not part of the build.")
        sys.exit(2)

    if args.only:
        _, sec = (section_w2 if args.only == "w2" else section_w4)(args)
        print(json.dumps({args.only: sec}), flush=True)
        return
    w3ds, eng, w3 = section_w3(args)
    log(f"workload 3: {w3['batched_queries_per_s']} q/s batched vs {w3['plain_queries_per_s']} plain")
    served, streams = (None, {}) if args.no_served else section_served(args, w3ds, eng)
    eng.close()
    w2ds = w2 = w4ds = w4 = None
    if not args.no_w2:
        w2ds, w2 = section_w2(args)
        log(f"workload 2: {w2['embeddings_per_s']:.3e} embeddings/s, parity {w2['parity']}")
    if not args.no_w4:
        w4ds, w4 = section_w4(args)
    cpu = cpus = None
    if not args.no_cpu:
        cpu, cpus = cpu_baselines(args, w3ds, w2ds, w4ds, streams)
    value = w3["embeddings_per_step"] / (w3["ms_per_step"] / 1e3)
    line = {
        "metric": "embeddings/s and queries/s (p50/p99 latency) vs plain GPU; counts bit-exact",
        "value": round(value, 1),
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(w3["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": "workload3: 100,000 graphs n~U[20,200] (ER / BA m=2), avg degree 4, 32 labels "
                               "Zipf(1.1); 256 random-walk queries of 4-8 vertices; batched engine, one run of all "
                               "queries per step" if args.db_scale == 1.0 else f"workload3 scale {args.db_scale}",
                   "path": "batched", "queries": w3["queries"], "embeddings_per_step": w3["embeddings_per_step"]},
        "queries_per_s": w3["batched_queries_per_s"],
        "batched": {k: v for k, v in w3.items() if k not in ("roofline", "ms_per_step")},
        "roofline": w3["roofline"],
        "served": served,
        "workload2": w2,
        "workload4": w4,
        "cpu_baseline": cpu,
        "cpu_baselines": cpus,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
