"""GPU tests of the batched engine (hipGraph iterations, device work queue over
slots) and of the serving loop: every result must equal the plain path's and
the oracle's, alone and in the mix, through repeated graph replays."""
import json
import os

import numpy as np
import pytest

import geneva_amd as G
from geneva_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def w3(require_gpu):
    gold = json.load(open(os.path.join(GOLDEN, "scaled.json")))
    w = [x for x in gold["workloads"] if x["workload"] == 3][0]
    ds = configs.workload3(scale=w["scale"], **w["kw"])
    eng = G.Engine(ds)
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)
    return ds, eng, w["cases"]


def test_batch_matches_golden_and_plain(w3):
    ds, eng, cases = w3
    qs = [c["query"] for c in cases]
    print("batch of", len(qs), "queries", flush=True)
    res, st = eng.mine_batch(qs)
    print("stats", st, flush=True)
    for c, r in zip(cases, res):
        assert r.count == c["count"]
        assert r.alg_bytes == c["alg_bytes"]
        assert sorted(r.graph_ids) == c["graph_ids"]
    # replay the captured iteration again, with eager plain-path work between
    for rep in range(3):
        _ = eng.mine(qs[rep])
        res2, _ = eng.mine_batch(qs)
        assert [r.count for r in res2] == [r.count for r in res]
        assert [r.graph_ids for r in res2] == [r.graph_ids for r in res]


def test_each_result_same_alone_and_in_the_mix(w3):
    ds, eng, cases = w3
    qs = [c["query"] for c in cases][:16]
    mix, _ = eng.mine_batch(qs)
    for q, m in zip(qs, mix):
        alone, _ = eng.mine_batch([q])
        plain = eng.mine(q)
        assert alone[0].count == m.count == plain.count
        assert alone[0].graph_ids == m.graph_ids == plain.graph_ids


def test_prefix_sharing_equals_unshared_tables(w3):
    """Queries with the same first steps share one table: every result equals
    the fixture with sharing on and off, tables were shared, and under a row
    cap too small for some leaders with their followers those followers
    restart on their own, with the same results."""
    ds, eng, cases = w3
    want = {c["query"]: c for c in cases}
    qs = [c["query"] for c in cases]
    qs = qs + qs[:12]   # admitted again: the same steps as the originals
    for row_cap in (1 << 22, 1 << 14):
        eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=row_cap)
        s0 = eng.batch_share_stats()
        on, _ = eng.mine_batch(qs)
        s1 = eng.batch_share_stats()
        eng.batch_share(False)
        off, _ = eng.mine_batch(qs)
        s2 = eng.batch_share_stats()
        print("row_cap", row_cap, "shared/restarted", s0, s1, s2, flush=True)
        assert s2 == s1
        assert s1[0] > s0[0]
        for q, a, b in zip(qs, on, off):
            c = want[q]
            assert a.count == b.count == c["count"]
            assert a.alg_bytes == b.alg_bytes == c["alg_bytes"]
            assert sorted(a.graph_ids) == sorted(b.graph_ids) == c["graph_ids"]
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)


def test_sizing_groups_equal_separate_sizing(w3):
    """Slots that size one input table in an iteration (one start list or
    label-pair range; a leader and its followers) are sized together: every
    result equals the fixture with the groups on and off, with prefix sharing
    on and off, and rows were sized in groups only when they were on."""
    ds, eng, cases = w3
    want = {c["query"]: c for c in cases}
    qs = [c["query"] for c in cases]
    qs = qs + qs[:12]
    for share in (True, False):
        eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)
        eng.batch_share(share)
        f0 = eng.batch_fused_rows()
        on, _ = eng.mine_batch(qs)
        f1 = eng.batch_fused_rows()
        eng.batch_fuse_sizing(False)
        off, _ = eng.mine_batch(qs)
        f2 = eng.batch_fused_rows()
        print("share", share, "fused rows", f0, f1, f2, flush=True)
        assert f1 > f0 and f2 == f1
        for q, a, b in zip(qs, on, off):
            c = want[q]
            assert a.count == b.count == c["count"]
            assert a.alg_bytes == b.alg_bytes == c["alg_bytes"]
            assert sorted(a.graph_ids) == sorted(b.graph_ids) == c["graph_ids"]
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)


def test_overflow_falls_back_to_plain_path(require_gpu):
    import oracle as O
    ds = configs.workload4(scale=0.005, nqueries=8)
    og = O.OracleGraph.of(ds)
    eng = G.Engine(ds)
    eng.batch_config(slots=8, slot_bytes=1 << 16, row_cap=1 << 20)   # tiny slot tables
    res, st = eng.mine_batch(ds.queries)
    print("stats", st, flush=True)
    assert st["overflowed"] + st["pieces"] >= 1
    for q, r in zip(ds.queries, res):
        o = og.mine(q, threads=4)
        assert r.count == o.count
        assert r.alg_bytes == o.alg_bytes


def test_pieces_of_large_queries_add_up(w3):
    """Queries whose tables outgrow small slots come back as pieces of their
    first table; the pieces' counts, bytes and graph ids add up to the
    fixture's."""
    ds, eng, cases = w3
    qs = [c["query"] for c in cases]
    eng.batch_config(slots=64, slot_bytes=1 << 18, row_cap=1 << 22)
    res, st = eng.mine_batch(qs)
    print("stats", st, flush=True)
    assert st["pieces"] > 0
    for c, r in zip(cases, res):
        assert r.count == c["count"]
        assert r.alg_bytes == c["alg_bytes"]
        assert sorted(r.graph_ids) == c["graph_ids"]
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)


def test_row_cap_defers_slots_without_changing_results(w3):
    ds, eng, cases = w3
    qs = [c["query"] for c in cases][:24]
    want = [c["count"] for c in cases][:24]
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 14)
    res, st = eng.mine_batch(qs)
    assert [r.count for r in res] == want
    # sustained admission under the same row cap: slots that keep waiting
    # age and run ahead of refilled lower slots, and every request finishes
    # with its own result
    rng = np.random.default_rng(9)
    arr = np.cumsum(rng.exponential(1 / 20000.0, size=len(qs))).tolist()
    lat, cnt, st = eng.serve(qs, arr, batched=True)
    print("serve under row cap", st, flush=True)
    assert cnt == want
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)


def test_tunings_agree(w3):
    ds, eng, cases = w3
    qs = [c["query"] for c in cases][:20]
    want = [c["count"] for c in cases][:20]
    for ipt in (1, 2, 4, 8):
        for bpc in (4, 32):
            eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22, ipt=ipt, blocks_per_cu=bpc)
            res, _ = eng.mine_batch(qs)
            assert [r.count for r in res] == want, (ipt, bpc)
    tuned = eng.autotune([c["query"] for c in cases])
    print("autotune", tuned, flush=True)
    res, _ = eng.mine_batch(qs)
    assert [r.count for r in res] == want
    # per-class candidates per lane (the tuning of each query class) change
    # nothing but speed
    assert tuned["classes"], tuned
    assert all(c["ipt"] <= tuned["ipt"] for c in tuned["classes"].values())
    # rows per iteration: the allocation's
    assert tuned["row_cap"] == 1 << 22, tuned
    # first steps from the start tables instead of the label-pair edge tables
    eng.batch_pair_start(False)
    res2, _ = eng.mine_batch(qs)
    eng.batch_pair_start(True)
    assert [r.count for r in res2] == want
    assert [r.alg_bytes for r in res2] == [r.alg_bytes for r in res]


def test_serving_results_match(w3):
    ds, eng, cases = w3
    qs = [c["query"] for c in cases]
    want = [c["count"] for c in cases]
    rng = np.random.default_rng(5)
    arr = np.cumsum(rng.exponential(1 / 2000.0, size=len(qs))).tolist()
    lat, cnt, st = eng.serve(qs, arr, batched=True)
    print("serve stats", st, "p50", np.percentile(lat, 50), "p99", np.percentile(lat, 99), flush=True)
    assert cnt == want
    assert all(l > 0 for l in lat)
    lat2, cnt2, st2 = eng.serve(qs, arr, batched=False)
    assert cnt2 == want


@pytest.fixture(scope="module")
def w5(require_gpu):
    """The serving query mix (random walks of 3-10 vertices) over the scaled
    workload-3 database, with oracle counts, graph ids and bytes (and reference
    cpumine counts) in tests/golden/workload5_scaled.json."""
    gold = json.load(open(os.path.join(GOLDEN, "workload5_scaled.json")))
    ds = configs.workload3(scale=gold["scale"], nqueries=0)
    qs = configs.workload5_queries(ds, gold["nqueries"])
    assert qs == [c["query"] for c in gold["cases"]], "generator drifted from the fixture"
    eng = G.Engine(ds)
    eng.batch_config(slots=64, slot_bytes=8 << 20, row_cap=1 << 22)
    return ds, eng, gold["cases"]


def _check(cases, counts, ids, what):
    for i, c in enumerate(cases):
        assert counts[i] == c["count"], (what, i, counts[i], c["count"])
        assert sorted(ids[i]) == c["graph_ids"], (what, i)
        if "ref_count" in c:
            assert counts[i] == c["ref_count"], (what, i)


def test_workload5_batch_and_plain_match_oracle(w5):
    ds, eng, cases = w5
    qs = [c["query"] for c in cases]
    print("workload 5 mix:", len(qs), "queries", flush=True)
    res, st = eng.mine_batch(qs)
    print("batch stats", st, flush=True)
    _check(cases, [r.count for r in res], [r.graph_ids for r in res], "batch")
    assert [r.alg_bytes for r in res] == [c["alg_bytes"] for c in cases]
    plain = [eng.mine(q) for q in qs]
    _check(cases, [r.count for r in plain], [r.graph_ids for r in plain], "plain")
    assert [r.alg_bytes for r in plain] == [c["alg_bytes"] for c in cases]


def test_workload5_serving_matches_oracle(w5):
    ds, eng, cases = w5
    qs = [c["query"] for c in cases]
    rng = np.random.default_rng(55)
    arr = np.cumsum(rng.exponential(1 / 3000.0, size=len(qs))).tolist()
    stats = {}
    for batched, budget in ((True, 16 << 30), (False, 16 << 30), (True, 1)):
        print("serve batched" if batched else "serve plain", "budget", budget, flush=True)
        lat, cnt, st, ids = eng.serve(qs, arr, batched=batched, mem_budget=budget, with_graphs=True)
        print("  stats", st, "p50 %.3f p99 %.3f" % (np.percentile(lat, 50), np.percentile(lat, 99)),
              flush=True)
        _check(cases, cnt, ids, ("serve", batched, budget))
        assert all(l > 0 for l in lat)
        stats[(batched, budget)] = st
    # an admission bound below every materialising request's estimate: such
    # requests enter the batch engine one at a time (requests whose plan
    # materialises no table have an estimate of 0 and still join), so the
    # same stream takes more iterations, with unchanged results
    tight, loose = stats[(True, 1)], stats[(True, 16 << 30)]
    assert tight["iterations"] > loose["iterations"]
    assert tight["max_batched"] <= loose["max_batched"]
