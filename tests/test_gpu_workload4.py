"""Workload 4 (skewed Chung-Lu graph, exponent 1.9, 4M vertices, hubs of
degree ~5e4, 64 labels; induced 6-8 vertex queries with two or more cycles).

Whole counts at full size are far beyond the oracle, so parity is checked
(1) on the scaled graph for whole queries, (2) at full size on shards of the
start list that the oracle mined (tests/golden/make_golden.py --only-w4), and
(3) by a size-independent property: the shard counts of a partition of the
start list add up to the whole count."""
import json
import os

import pytest

import geneva_amd as G
from geneva_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_workload4_scaled_whole_queries(require_gpu):
    gold = _load("workload4_scaled.json")
    ds = configs.workload4(scale=gold["scale"], **gold["kw"])
    eng = G.Engine(ds)
    for c in gold["cases"]:
        print("workload4 scaled query", c["query"].replace("\n", " "), flush=True)
        r = eng.mine(c["query"])
        assert r.count == c["count"]
        assert r.alg_bytes == c["alg_bytes"]
    eng.close()


def test_workload4_unconstrained_scaled_queries(require_gpu):
    """The configuration's walk queries as they come (pendant vertices on
    hubs included, no two-neighbour condition), scaled: plain path and batch
    engine against the oracle's counts and bytes."""
    gold = _load("workload4_unconstrained_scaled.json")
    ds = configs.workload4(scale=gold["scale"], **gold["kw"])
    eng = G.Engine(ds)
    qs = [c["query"] for c in gold["cases"]]
    for c in gold["cases"]:
        print("workload4 unconstrained query", c["query"].replace("\n", " "), flush=True)
        r = eng.mine(c["query"])
        assert r.count == c["count"]
        assert r.alg_bytes == c["alg_bytes"]
    print("workload4 unconstrained batch", flush=True)
    res, st = eng.mine_batch(qs)
    print("stats", st, flush=True)
    for c, r in zip(gold["cases"], res):
        assert r.count == c["count"]
        assert r.alg_bytes == c["alg_bytes"]
    eng.close()


def test_workload4_shards_partition_the_count(require_gpu):
    gold = _load("workload4_scaled.json")
    ds = configs.workload4(scale=gold["scale"], **gold["kw"])
    eng = G.Engine(ds)
    for c in gold["cases"][:8]:
        q = c["query"]
        print("workload4 partition query", q.replace("\n", " "), flush=True)
        strided = [eng.mine(q, start_begin=b, stride=3) for b in range(3)]
        assert sum(r.count for r in strided) == c["count"]
        assert sum(r.alg_bytes for r in strided) == c["alg_bytes"]
        half = 97
        lo, hi = eng.mine(q, start_end=half), eng.mine(q, start_begin=half)
        assert lo.count + hi.count == c["count"]
    eng.close()


def test_workload4_full_size_shards(require_gpu):
    if not os.path.exists(os.path.join(GOLDEN, "workload4_shards.json")):
        pytest.skip("workload4_shards.json not generated (make_golden.py --only-w4)")
    gold = _load("workload4_shards.json")
    ds = configs.workload4()
    eng = G.Engine(ds)
    for c in gold["cases"]:
        for s in c["shards"]:
            print("workload4 full shard", s["begin"], s["end"], "query",
                  c["query"].replace("\n", " "), flush=True)
            r = eng.mine(c["query"], start_begin=s["begin"], start_end=s["end"])
            assert r.count == s["count"]
            assert r.alg_bytes == s["alg_bytes"]
    eng.close()
