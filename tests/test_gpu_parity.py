"""GPU parity: the HIP path (through libgeneva_host -> libgeneva_hip) against
the oracle and the golden fixtures.  Counts, matched graph ids and the
algorithmic byte totals must be identical (integer work, no tolerance)."""
import json
import os

import numpy as np
import pytest

import geneva_amd as G
from geneva_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def w1(require_gpu):
    ds = configs.workload1()
    return ds, G.Engine(ds)


def test_workload1_matches_reference_counts(w1):
    ds, eng = w1
    gold = _load("workload1.json")
    for c in gold["cases"]:
        print("workload1 query", c["query"].replace("\n", " "), flush=True)
        for sv in (False, True):
            r = eng.mine(c["query"], single_v=sv)
            assert r.count == c["ref_count"]
            assert sorted(r.graph_ids) == c["graph_ids"]
            assert r.alg_bytes == c["alg_bytes"]


def test_enron_multi_edge_label(require_gpu):
    import gzip
    import shutil
    import tempfile
    d = tempfile.mkdtemp()
    for f in ("Enron.mygraph", "Enron.metadata"):
        with gzip.open(os.path.join(GOLDEN, "enron", f + ".gz"), "rb") as s, \
                open(os.path.join(d, f), "wb") as t:
            shutil.copyfileobj(s, t)
    ds = G.Dataset.load(os.path.join(d, "Enron.mygraph"), os.path.join(d, "Enron.metadata"))
    eng = G.Engine(ds)
    for c in _load("enron.json")["cases"]:
        print("enron query", c["query"].replace("\n", " "), flush=True)
        r = eng.mine(c["query"])
        assert r.count == c["ref_count"]
        assert r.alg_bytes == c["alg_bytes"]


def test_scaled_workloads(require_gpu):
    for w in _load("scaled.json")["workloads"]:
        ds = configs.WORKLOADS[w["workload"]](scale=w["scale"], **w["kw"])
        eng = G.Engine(ds)
        for c in w["cases"]:
            print("workload", w["workload"], "query", c["query"].replace("\n", " "), flush=True)
            r = eng.mine(c["query"])
            assert r.count == c["count"]
            assert r.alg_bytes == c["alg_bytes"]
            if ds.ngraphs > 1:
                assert sorted(r.graph_ids) == c["graph_ids"]
        eng.close()


def test_random_databases_against_live_oracle(require_gpu):
    import oracle as O
    for seed in range(4):
        ds = G.Dataset.database(300, 5, 40, 3.0 + seed, 0.5, 3 + seed, 1.0 * (seed % 2), 1000 + seed)
        ds.add_walk_queries(10, 3, 7, seed=seed)
        ds.add_walk_queries(5, 4, 7, induced=True, seed=seed + 50)
        og = O.OracleGraph.of(ds)
        eng = G.Engine(ds)
        for q in ds.queries:
            print("random db", seed, "query", q.replace("\n", " "), flush=True)
            want = og.mine(q)
            got = eng.mine(q)
            assert got.count == want.count
            assert got.graph_ids == want.graph_ids
            assert got.alg_bytes == want.alg_bytes
        eng.close()


HUB_ISECT_QUERIES = [
    # the counted vertex has two back edges into the parent prefix, so the
    # parents' lists are intersected once per parent (k_isect, k_isect_big):
    # a 4-clique, a triangle with a fourth vertex on two of its corners (the
    # plan places it last, joined to columns 0 and 1) and a 4-clique with a
    # pendant of a fourth label
    "v 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 1 2 1\ne 1 3 1\ne 2 3 1\n",
    "v 0 1\nv 1 2\nv 2 3\nv 3 4\ne 0 1 1\ne 0 2 1\ne 1 2 1\ne 3 0 1\ne 3 1 1\n",
    "v 0 1\nv 1 1\nv 2 2\nv 3 1\nv 4 4\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 1 2 1\ne 1 3 1\ne 2 3 1\n"
    "e 3 4 1\n",
]
HUB_CHAIN_QUERIES = [
    # a triangle and a 4-cycle with a two-edge tail of labels no other query
    # vertex has: counted from per-vertex chain weights (largest degree > 256)
    "v 0 1\nv 1 1\nv 2 1\nv 3 2\nv 4 3\ne 0 1 1\ne 1 2 1\ne 0 2 1\ne 2 3 1\ne 3 4 1\n",
    "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 2\nv 5 3\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 0 1\ne 0 4 1\n"
    "e 4 5 1\n",
]


def test_hub_graph_and_chunked_frontiers(require_gpu):
    """A hub with ~20k neighbours (one label group > 4096, skip index in the
    stored block), a few secondary hubs of ~1,000 neighbours, and a tiny level
    budget that forces many extension chunks: counts must not depend on the
    chunking.  The prefix-shared intersections (with parents of more than 256
    anchor entries and other lists longer than the 4096-entry LDS stage) and
    the pendant-chain counting must run, and agree with the oracle; with the
    tiny budget the intersection buffer overflows and the query is re-run in
    chunks."""
    import oracle as O
    rng = np.random.default_rng(7)
    n = 30000
    vl = np.concatenate([[0], rng.integers(1, 5, size=n)]).astype(np.uint32)
    vl[1] = 1
    # the hub touches every vertex, so the sources form one id interval
    hub = [(1, v, 1) for v in range(2, n + 1)]
    sec = [(int(h), int(v), 1) for h in range(2, 12)
           for v in rng.choice(np.arange(12, n + 1), size=1000, replace=False)]
    rest = rng.integers(2, n + 1, size=(40000, 2))
    rest = [(int(a), int(b), 1) for a, b in rest if a != b]
    ds = G.Dataset.from_arrays(vl, np.array(hub + sec + rest, dtype=np.uint32))
    qs = ["v 0 1\nv 1 2\nv 2 2\ne 0 1 1\ne 0 2 1\n",
          "v 0 1\nv 1 1\nv 2 2\nv 3 3\ne 0 1 1\ne 0 2 1\ne 1 2 1\ne 2 3 1\n",
          "v 0 2\nv 1 1\nv 2 2\nv 3 3\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 0 1\n",
          "v 0 1\nv 1 2\nv 2 3\nv 3 2\nv 4 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 3 4 1\n"]
    og = O.OracleGraph.of(ds)
    big = G.Engine(ds)
    small = G.Engine(ds, level_bytes=1 << 16)
    for q in qs + HUB_ISECT_QUERIES + HUB_CHAIN_QUERIES:
        print("hub query", q.replace("\n", " "), flush=True)
        want = og.mine(q, threads=8)
        big.profile(True, count_bytes=False)
        a = big.mine(q)
        prof = big.profile_read()
        big.profile(False)
        b = small.mine(q)
        assert a.count == b.count == want.count
        assert a.alg_bytes == b.alg_bytes == want.alg_bytes
        assert b.launches >= a.launches
        if q in HUB_ISECT_QUERIES:
            assert prof["k_isect"][0] > 0, prof
        if q in HUB_CHAIN_QUERIES:
            assert prof["k_count_chain"][0] > 0, prof


def test_absent_labels_give_zero(w1):
    ds, eng = w1
    assert eng.mine("v 0 99\nv 1 1\ne 0 1 1\n").count == 0
    assert eng.mine("v 0 1\nv 1 2\ne 0 1 7\n").count == 0


def test_bad_query_raises(w1):
    ds, eng = w1
    with pytest.raises(G.GenevaError):
        eng.mine("v 0 1\nv 1 2\n")          # no edges
    with pytest.raises(G.GenevaError):
        eng.mine("v 0 1\nv 1 2\nv 2 3\ne 0 1 1\n")  # disconnected


ROW_COUNT_QUERIES = [
    # (query, kind): 0-2 the row-count modes of the counted step (one edge; a
    # path whose last vertex hangs off an earlier column; stars whose last two
    # leaves share label and centre), 3 leaf groups counted on their core, 4
    # pendant chains counted from per-vertex weights
    ("v 0 1\nv 1 2\ne 0 1 1\n", 0),
    ("v 0 1\nv 1 1\ne 0 1 1\n", 0),
    ("v 0 1\nv 1 2\nv 2 3\nv 3 1\ne 0 1 1\ne 1 2 1\ne 2 3 1\n", 1),
    ("v 0 1\nv 1 2\nv 2 2\nv 3 2\ne 0 1 1\ne 0 2 1\ne 0 3 1\n", 2),
    ("v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 0 4 1\n", 2),
    ("v 0 1\nv 1 2\nv 2 1\nv 3 1\ne 0 1 1\ne 1 2 1\ne 1 3 1\n", 2),
    # leaf groups on a core: two same-label pairs on one centre, a double
    # star, three same-label leaves, a leaf whose label is the centre's
    ("v 0 1\nv 1 2\nv 2 2\nv 3 3\nv 4 3\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 0 4 1\n", 3),
    ("v 0 1\nv 1 2\nv 2 3\nv 3 3\nv 4 1\nv 5 2\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 1 4 1\ne 1 5 1\n", 3),
    ("v 0 2\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\n", 3),
    ("v 0 1\nv 1 1\nv 2 2\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\n", 3),
    # pendant chains: a triangle and a 4-cycle with a two-edge tail of labels
    # no other vertex has (these graphs have no vertex of degree > 256, so the
    # tail is walked here; test_hub_graph_and_chunked_frontiers runs the same
    # shapes through the chain-weight kernels)
    ("v 0 1\nv 1 1\nv 2 1\nv 3 2\nv 4 3\ne 0 1 1\ne 1 2 1\ne 0 2 1\ne 2 3 1\ne 3 4 1\n", 4),
    ("v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 2\nv 5 3\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 0 1\ne 0 4 1\n"
     "e 4 5 1\n", 4),
]


def test_count_by_rows_equals_candidate_walk_and_oracle(require_gpu):
    """Counting the last step per row (segment arithmetic) gives the counts,
    graph ids and algorithmic bytes of the candidate walk and of the oracle."""
    import oracle as O
    for seed in range(3):
        ds = G.Dataset.database(200, 5, 60, 4.0 + 2 * seed, 0.5, 3 + seed, 0.0, 2000 + seed)
        og = O.OracleGraph.of(ds)
        rows, walk = G.Engine(ds), G.Engine(ds)
        walk.count_by_rows(False)
        rows.profile(True)
        for q, mode in ROW_COUNT_QUERIES:
            print("row count seed", seed, "mode", mode, "query", q.replace("\n", " "), flush=True)
            want = og.mine(q)
            a, b = rows.mine(q), walk.mine(q)
            assert a.count == b.count == want.count
            assert a.graph_ids == b.graph_ids == want.graph_ids
            assert a.alg_bytes == b.alg_bytes == want.alg_bytes
        assert rows.profile_read()["k_count_rows"][0] > 0
        rows.close()
        walk.close()
