"""GPU parity at the benchmark's full sizes: workload 3 (100,000 graphs, 256
queries; plain and batched) and workload 2 (1M-vertex Chung-Lu, 5 shapes)
against the oracle's full-size fixtures, and the time-limited plain path
(partial counts that are exact for the part mined)."""
import hashlib
import json
import os

import pytest

import geneva_amd as G
from geneva_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _sha(ids):
    return hashlib.sha256("\n".join(map(str, sorted(ids))).encode()).hexdigest()


def test_workload3_full_size_matches_oracle(require_gpu):
    gold = json.load(open(os.path.join(GOLDEN, "workload3_full.json")))["cases"]
    ds = configs.workload3()
    qs = ds.queries
    assert qs == [c["query"] for c in gold], "generator drifted from the fixture"
    eng = G.Engine(ds)
    eng.batch_config(slots=256, slot_bytes=64 << 20, row_cap=1 << 26)
    print("workload 3 full size: plain", flush=True)
    plain = [eng.mine(q) for q in qs]
    print("workload 3 full size: batched", flush=True)
    batched, st = eng.mine_batch(qs)
    print("batch stats", st, flush=True)
    for i, (c, p, b) in enumerate(zip(gold, plain, batched)):
        for r in (p, b):
            assert r.count == c["count"], i
            assert r.alg_bytes == c["alg_bytes"], i
            assert len(r.graph_ids) == c["ngraphs"], i
            assert _sha(r.graph_ids) == c["graph_ids_sha256"], i
    eng.close()


def test_workload2_full_size_matches_oracle(require_gpu):
    gold = {c["query"]: c for c in json.load(open(os.path.join(GOLDEN, "workload2_full.json")))["cases"]}
    ds = configs.workload2()
    eng = G.Engine(ds, level_bytes=1 << 31)
    for q in ds.queries:
        print("workload 2 query", q.replace("\n", " "), flush=True)
        r = eng.mine(q)
        assert r.count == gold[q]["count"]
        assert r.alg_bytes == gold[q]["alg_bytes"]
    eng.close()


def test_time_limit_gives_exact_partial_counts(require_gpu):
    """A query stopped at its time limit reports the part mined; a limit it
    can meet gives the whole result of the unlimited path."""
    ds = configs.workload4(scale=0.01, nqueries=6)
    eng = G.Engine(ds, level_bytes=1 << 20)
    for q in ds.queries:
        print("time limit query", q.replace("\n", " "), flush=True)
        whole = eng.mine(q)
        eng.time_limit(120.0)
        r = eng.mine(q)
        fin, frac = eng.last_progress()
        assert fin and frac == 1.0
        assert (r.count, r.alg_bytes) == (whole.count, whole.alg_bytes)
        eng.time_limit(1e-9)
        p = eng.mine(q)
        fin, frac = eng.last_progress()
        eng.time_limit(0)
        assert 0.0 <= frac <= 1.0
        assert p.count <= whole.count and p.alg_bytes <= whole.alg_bytes
        if not fin:
            assert frac < 1.0
    eng.close()
