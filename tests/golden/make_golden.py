"""Write the golden fixtures under tests/golden/ (run in the build container).

    python tests/golden/make_golden.py            # plans, workload 1, Enron, scaled workloads
    python tests/golden/make_golden.py --full2    # + full-size workload 2 counts (long)
    python tests/golden/make_golden.py --only-w4  # workload 4: scaled counts, full-size shards
                                                  # (--shards-only: the shards alone)
    python tests/golden/make_golden.py --only-w5  # workload 5: the serving query mix (3-10
                                                  # vertices) over the scaled workload-3 database
    python tests/golden/make_golden.py --only-w3full  # workload 3 at full size (100,000 graphs):
                                                      # counts, bytes and matched graph ids
    python tests/golden/make_golden.py --only-w4u  # workload 4's walk queries without the
                                                   # two-neighbour condition, scaled

Reference values come from the reference's own cpumine, compiled from its
sources into oracle/_ref/cpumine (`make -C oracle ref`), run on files written
by this project's format writers (.reorder files without skip-index entries); oracle
values come from oracle/liboracle.so.  Only data is written: query texts,
plan dumps, counts, matched graph ids and algorithmic byte totals.
"""
import gzip
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import geneva_amd as G  # noqa: E402
import oracle as O  # noqa: E402
from geneva_amd import configs  # noqa: E402

TOT = re.compile(r"write to final result cpu, tot=(\d+)")


def ref_run(prefix, query_text, timeout=600):
    """Run the reference cpumine; return (plan dump, count or None)."""
    wd = tempfile.mkdtemp()
    qf = os.path.join(wd, "q.query")
    with open(qf, "w") as f:
        f.write(query_text)
    try:
        p = subprocess.run([O.REF_CPUMINE, prefix + ".reorder", qf, prefix + ".mygraph",
                            prefix + ".metadata"], cwd=wd, capture_output=True, text=True,
                           timeout=timeout)
        out = p.stdout
    except subprocess.TimeoutExpired as e:
        out = e.stdout.decode() if isinstance(e.stdout, bytes) else (e.stdout or "")
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    lines = out.splitlines()
    plan = [l for l in lines if l.startswith(("restrict: ", "phase ", "svid:"))]
    m = TOT.search(out)
    return "\n".join(plan), (int(m.group(1)) if m else None)


def write(ds, d, name):
    prefix = os.path.join(d, name)
    ds.write(prefix)
    return prefix


def enron():
    d = tempfile.mkdtemp()
    for f in ("Enron.mygraph", "Enron.metadata"):
        with gzip.open(os.path.join(HERE, "enron", f + ".gz"), "rb") as s, open(os.path.join(d, f), "wb") as t:
            shutil.copyfileobj(s, t)
    return G.Dataset.load(os.path.join(d, "Enron.mygraph"), os.path.join(d, "Enron.metadata"))


def plan_corpus():
    qs = []
    c1 = configs.workload1()
    qs += c1.queries
    c1.add_walk_queries(40, 3, 8, seed=101)
    qs += c1.queries[10:]
    c3 = configs.workload3(scale=0.01, nqueries=30)
    c3.add_walk_queries(30, 4, 9, induced=True, seed=102)
    qs += c3.queries
    c2 = configs.workload2(scale=0.01)
    qs += c2.queries
    en = enron()
    en.add_walk_queries(40, 3, 8, seed=103)
    en.add_walk_queries(20, 4, 7, induced=True, seed=104)
    qs += en.queries
    with open(os.path.join(HERE, "enron", "Enron_4_6_0.query")) as f:
        qs.append(f.read())
    # symmetric and multi-edge-label patterns for the automorphism and
    # triangle-score branches of the planner
    def q(labels, edges):
        return "".join(f"v {i} {l}\n" for i, l in enumerate(labels)) + \
               "".join(f"e {a} {b} {l}\n" for a, b, l in edges)
    for k in range(3, 8):
        qs.append(q([1] * k, [(i, i + 1, 1) for i in range(k - 1)]))                  # path
        qs.append(q([1] * k, [(0, i, 1) for i in range(1, k)]))                       # star
        qs.append(q([1] * k, [(i, (i + 1) % k, 1) for i in range(k)]))                # cycle
        qs.append(q([1] * k, [(a, b, 1) for a in range(k) for b in range(a + 1, k)]))  # clique
    for la in ([1, 1, 1], [1, 1, 2], [1, 2, 1], [2, 1, 1], [1, 2, 3]):
        for el in ([1, 1, 1], [1, 1, 2], [1, 2, 1], [2, 1, 1], [1, 2, 3]):
            qs.append(q(la + [la[0]], [(0, 1, el[0]), (1, 2, el[1]), (0, 2, el[2]), (2, 3, 1)]))
    qs.append(q([1, 2, 1, 2, 1, 2], [(0, 1, 1), (1, 2, 1), (2, 3, 1), (3, 4, 1), (4, 5, 1), (5, 0, 1)]))
    qs.append(q([1, 1, 2, 2], [(0, 2, 1), (1, 2, 1), (0, 3, 1), (1, 3, 1)]))
    qs.append(q([3, 1, 1, 1, 2], [(0, 1, 2), (0, 2, 2), (0, 3, 2), (3, 4, 3), (1, 2, 1)]))
    return qs


def full_workload2():
    ds = configs.workload2()
    ogw = O.OracleGraph.of(ds)
    cases = []
    for qt in ds.queries:
        r = ogw.mine(qt, threads=os.cpu_count())
        cases.append({"query": qt, "count": r.count, "alg_bytes": r.alg_bytes,
                      "oracle_seconds": r.seconds, "threads": os.cpu_count()})
        print("workload2 full:", r.count, r.seconds, flush=True)
    json.dump({"workload": 2, "scale": 1.0, "cases": cases},
              open(os.path.join(HERE, "workload2_full.json"), "w"), indent=1)


def _shard_worker(og, qt, b, e, conn):
    r = og.mine(qt, threads=1, start_begin=b, start_end=e)
    conn.send((r.count, r.alg_bytes, r.seconds))
    conn.close()


def workload4_shards(nq=64, cap_s=30.0, procs=8):
    """Full-size workload 4: every query's count is far beyond any CPU run, so
    the oracle mines shards of 8 consecutive start vertices (indices into the
    id-sorted start list): per query, at the middle of the list and at every
    halving of that index towards the heavy low ids (Chung-Lu weights fall with
    the id).  Shards run single-threaded in forked workers, `procs` at a time;
    a shard that takes more than cap_s seconds is dropped."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    ds = configs.workload4()
    ogw = O.OracleGraph.of(ds)
    vl = np.asarray(ds.vlabel)
    tasks = []
    for qi, qt in enumerate(ds.queries[:nq]):
        v0 = int(re.search(r"svid:(\d+)", G.plan_dump(qt)).group(1))
        l0 = [int(l.split()[2]) for l in qt.splitlines()
              if l.startswith("v ") and int(l.split()[1]) == v0][0]
        nstart = int((vl[1:] == l0).sum())
        p = nstart // 2
        while p >= 8:
            tasks.append((qi, p))
            p //= 2
    results = {}
    running = []
    pending = list(tasks)
    while pending or running:
        while pending and len(running) < procs:
            qi, p = pending.pop(0)
            a, b = ctx.Pipe(duplex=False)
            pr = ctx.Process(target=_shard_worker, args=(ogw, ds.queries[qi], p, p + 8, b))
            pr.start()
            b.close()
            running.append((qi, p, pr, a, time.time()))
        still = []
        for qi, p, pr, a, t0 in running:
            if a.poll():
                cnt, by, secs = a.recv()
                pr.join()
                results[(qi, p)] = {"begin": p, "end": p + 8, "count": cnt, "alg_bytes": by,
                                    "oracle_seconds": round(secs, 3)}
            elif time.time() - t0 > cap_s:
                pr.kill()
                pr.join()
            else:
                still.append((qi, p, pr, a, t0))
        running = still
        time.sleep(0.01)
    cases = []
    for qi, qt in enumerate(ds.queries[:nq]):
        shards = [results[k] for k in sorted(results, key=lambda k: -k[1]) if k[0] == qi]
        cases.append({"query": qt, "shards": shards})
        print("workload4 query", qi, [(s["begin"], s["count"]) for s in shards], flush=True)
    json.dump({"workload": 4, "scale": 1.0, "threads": 1, "cap_seconds": cap_s, "cases": cases},
              open(os.path.join(HERE, "workload4_shards.json"), "w"), indent=1)


def workload4_scaled(scale=0.01, nq=24):
    ds = configs.workload4(scale=scale, nqueries=nq)
    ogw = O.OracleGraph.of(ds)
    cases = []
    for qt in ds.queries:
        r = ogw.mine(qt, threads=os.cpu_count())
        cases.append({"query": qt, "count": r.count, "alg_bytes": r.alg_bytes})
        print("workload4 scaled:", r.count, round(r.seconds, 2), flush=True)
    json.dump({"workload": 4, "scale": scale, "kw": {"nqueries": nq}, "cases": cases},
              open(os.path.join(HERE, "workload4_scaled.json"), "w"), indent=1)


def workload4_unconstrained(scale=0.01, nq=24):
    """Workload 4's induced walk queries as they come (pendant vertices
    included, min_deg=0) on the scaled graph: oracle counts and algorithmic
    bytes."""
    ds = configs.workload4(scale=scale, nqueries=nq, min_deg=0)
    ogw = O.OracleGraph.of(ds)
    cases = []
    for qt in ds.queries:
        r = ogw.mine(qt, threads=os.cpu_count())
        cases.append({"query": qt, "count": r.count, "alg_bytes": r.alg_bytes})
        print("workload4 unconstrained:", r.count, round(r.seconds, 2), flush=True)
    json.dump({"workload": 4, "scale": scale, "kw": {"nqueries": nq, "min_deg": 0}, "cases": cases},
              open(os.path.join(HERE, "workload4_unconstrained_scaled.json"), "w"), indent=1)


def workload5(scale=0.01, nq=64):
    """The serving stream's queries (random walks of 3-10 vertices, the
    workload-5 generator) over workload 3's database at `scale`: oracle counts,
    matched graph ids and algorithmic bytes, and the reference cpumine count
    where it finishes quickly."""
    ds = configs.workload3(scale=scale, nqueries=0)
    qs = configs.workload5_queries(ds, nq)
    ogw = O.OracleGraph.of(ds)
    d = tempfile.mkdtemp()
    pw = write(ds, d, "w5")
    cases = []
    for qt in qs:
        r = ogw.mine(qt, threads=os.cpu_count())
        c = {"query": qt, "count": r.count, "alg_bytes": r.alg_bytes,
             "graph_ids": sorted(r.graph_ids)}
        if r.count < 2_000_000 and os.path.exists(O.REF_CPUMINE):
            _, ref = ref_run(pw, qt, timeout=120)
            if ref is not None:
                c["ref_count"] = ref
                assert ref == r.count, (ref, r.count, qt)
        cases.append(c)
        print("workload5:", len(qt.splitlines()), r.count, len(r.graph_ids), flush=True)
    shutil.rmtree(d, ignore_errors=True)
    json.dump({"workload": 5, "scale": scale, "nqueries": nq, "cases": cases},
              open(os.path.join(HERE, "workload5_scaled.json"), "w"), indent=1)


def workload3_full():
    """Workload 3 at full size: the oracle's count, algorithmic bytes and
    matched graph ids of all 256 queries (graph ids stored as a sorted list
    and as the SHA-256 of its decimal text, one id per line)."""
    import hashlib
    ds = configs.workload3()
    ogw = O.OracleGraph.of(ds)
    cases = []
    for qt in ds.queries:
        r = ogw.mine(qt, threads=os.cpu_count())
        ids = sorted(r.graph_ids)
        cases.append({"query": qt, "count": r.count, "alg_bytes": r.alg_bytes, "ngraphs": len(ids),
                      "graph_ids_sha256": hashlib.sha256("\n".join(map(str, ids)).encode()).hexdigest()})
    json.dump({"workload": 3, "scale": 1.0, "cases": cases},
              open(os.path.join(HERE, "workload3_full.json"), "w"), indent=1)
    print("workload3 full:", sum(c["count"] for c in cases), "embeddings")


def main():
    if "--only-w3full" in sys.argv:
        return workload3_full()
    if "--only-w5" in sys.argv:
        return workload5()
    if "--only-w4u" in sys.argv:
        return workload4_unconstrained()
    if "--only-full2" in sys.argv:
        return full_workload2()
    if "--only-w4" in sys.argv:
        if "--shards-only" not in sys.argv:
            workload4_scaled()
        return workload4_shards()
    full2 = "--full2" in sys.argv
    if not os.path.exists(O.REF_CPUMINE):
        sys.exit("oracle/_ref/cpumine missing: make -C oracle ref")
    d = tempfile.mkdtemp()
    # a small graph carrying edge labels 1..3 so cpumine can plan any corpus query
    small = G.Dataset.from_arrays([0] + [1, 2, 3, 1, 2, 3, 1, 2],
                                  [(1, 2, 1), (2, 3, 2), (3, 4, 3), (4, 5, 1), (5, 6, 2), (6, 7, 3), (7, 8, 1)])
    sp = write(small, d, "small")
    cases = []
    seen = set()
    for qt in plan_corpus():
        if qt in seen:
            continue
        seen.add(qt)
        plan, _ = ref_run(sp, qt, timeout=60)
        cases.append({"query": qt, "ref_plan": plan})
    json.dump(cases, open(os.path.join(HERE, "plans.json"), "w"), indent=0)
    print("plans:", len(cases))

    # workload 1 at full size: reference counts, oracle counts/graphs/bytes
    c1 = configs.workload1()
    p1 = write(c1, d, "w1")
    og = O.OracleGraph.of(c1)
    out = []
    for qt in c1.queries:
        _, ref = ref_run(p1, qt)
        r = og.mine(qt)
        out.append({"query": qt, "ref_count": ref, "count": r.count,
                    "graph_ids": sorted(r.graph_ids), "alg_bytes": r.alg_bytes})
    json.dump({"workload": 1, "cases": out}, open(os.path.join(HERE, "workload1.json"), "w"), indent=1)
    print("workload1:", [c["count"] for c in out])

    # Enron: the shipped query plus walk queries (edge labels 1..3)
    en = enron()
    pe = write(en, d, "enron")
    with open(os.path.join(HERE, "enron", "Enron_4_6_0.query")) as f:
        eq = [f.read()]
    en.add_walk_queries(12, 3, 5, seed=201)
    eq += en.queries
    oge = O.OracleGraph.of(en)
    out = []
    for qt in eq:
        _, ref = ref_run(pe, qt, timeout=900)
        r = oge.mine(qt)
        if ref is None:
            print("reference timed out on", qt.replace("\n", " | "))
            continue
        out.append({"query": qt, "ref_count": ref, "count": r.count, "alg_bytes": r.alg_bytes})
        assert ref == r.count, (ref, r.count, qt)
    json.dump({"cases": out}, open(os.path.join(HERE, "enron.json"), "w"), indent=1)
    print("enron:", [c["count"] for c in out])

    # scaled-down workloads 2-4 (oracle; reference where cheap)
    wl = []
    for w, scale, kw in ((2, 0.01, {}), (3, 0.01, {"nqueries": 48}), (4, 0.005, {"nqueries": 24})):
        ds = configs.WORKLOADS[w](scale=scale, **kw)
        ogw = O.OracleGraph.of(ds)
        pw = write(ds, d, f"w{w}")
        cases = []
        for qt in ds.queries:
            r = ogw.mine(qt, threads=8)
            c = {"query": qt, "count": r.count, "alg_bytes": r.alg_bytes,
                 "graph_ids": sorted(r.graph_ids)}
            if r.count < 2_000_000 and w == 3:
                _, ref = ref_run(pw, qt, timeout=120)
                if ref is not None:
                    c["ref_count"] = ref
                    assert ref == r.count, (ref, r.count, qt)
            cases.append(c)
        wl.append({"workload": w, "scale": scale, "kw": kw, "cases": cases})
        print(f"workload{w}@{scale}:", [c["count"] for c in cases][:12])
    json.dump({"workloads": wl}, open(os.path.join(HERE, "scaled.json"), "w"), indent=1)

    if full2:
        full_workload2()


if __name__ == "__main__":
    main()
