"""CPU tests: planner parity, oracle modes, closed forms and golden counts.

The product planner (geneva_amd/csrc/host/genMatchOrder.cpp) and the oracle's
planner (oracle/oracle_planner.cpp) are two restatements of the reference
planner; both are compared with plan dumps of the reference's own cpumine
(tests/golden/plans.json, written by tests/golden/make_golden.py).  The oracle
is pinned by brute force, by closed forms and by the reference cpumine's counts.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

import geneva_amd as G
import oracle as O
from geneva_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _norm(plan_text):
    return "\n".join(l.rstrip() for l in plan_text.strip().splitlines())


# ---------------------------------------------------------------- planner ----

def test_plans_match_reference_cpumine_dumps():
    plans = _load("plans.json")
    assert len(plans) >= 100
    for case in plans:
        q, ref = case["query"], _norm(case["ref_plan"])
        assert _norm(G.plan_dump(q)) == ref, q
        assert _norm(O.plan(q)) == ref, q


def _order_and_restricts(dump):
    order, restr = [], {}
    lines = dump.strip().splitlines()
    for i, l in enumerate(lines):
        if l.startswith("restrict: "):
            k, v = l[len("restrict: "):].split("<")
            restr[int(k)] = int(v)
        if l.startswith("phase") and " type E " in l:
            toks = lines[i + 1].split()
            pairs = [int(t.split(":")[1]) for t in toks]
            for j in range(0, len(pairs), 2):
                s, x = pairs[j], pairs[j + 1]
                if not order:
                    order.append(s)
                if x not in order:
                    order.append(x)
    return order, restr


def test_single_vertex_plan_keeps_order_and_restrictions():
    # the SV planner splits two-vertex extensions and multi-pair reductions but
    # visits vertices in the same order, so the restriction set is the same
    for case in _load("plans.json"):
        q = case["query"]
        dv, sv = G.plan_dump(q), G.plan_dump(q, single_v=True)
        assert _order_and_restricts(dv) == _order_and_restricts(sv), q
        for l in sv.splitlines():
            if l.startswith("svid:"):
                assert len(l.split()) == 2, "SV phases carry one pair"


def test_plan_invariants_on_random_queries():
    ds = configs.workload1(scale=0.2)
    ds.add_walk_queries(60, 3, 10, seed=7)
    ds.add_walk_queries(40, 5, 9, induced=True, seed=8)
    for q in ds.queries:
        dump = G.plan_dump(q)
        order, restr = _order_and_restricts(dump)
        n = sum(1 for l in q.splitlines() if l.startswith("v "))
        assert sorted(order) == list(range(n))
        # a restriction key is always placed after its value
        pos = {v: i for i, v in enumerate(order)}
        for k, v in restr.items():
            if v != 0xFFFFFFFF and k < n:
                assert pos[v] < pos[k]
        # every query edge covered exactly once by an E or R pair
        edges = set()
        for l in q.splitlines():
            if l.startswith("e "):
                a, b = map(int, l.split()[1:3])
                edges.add((min(a, b), max(a, b)))
        covered = []
        for l in dump.splitlines():
            if l.startswith("svid:"):
                toks = [int(t.split(":")[1]) for t in l.split()]
                covered += [(min(toks[j], toks[j + 1]), max(toks[j], toks[j + 1]))
                            for j in range(0, len(toks), 2)]
        assert sorted(covered) == sorted(edges), q


def test_automorphism_closed_forms_through_restrictions():
    # same-label path/cycle/star/clique: the restriction chain must leave
    # count == monomorphisms / |Aut| on a graph with plenty of each
    ds = G.Dataset.chung_lu(300, 2.5, 8.0, 0.0, 1, 0.0, 11)
    og = O.OracleGraph.of(ds)
    shapes = {
        "path4": ([(0, 1), (1, 2), (2, 3)], 2),
        "star4": ([(0, 1), (0, 2), (0, 3)], 6),
        "cycle4": ([(0, 1), (1, 2), (2, 3), (3, 0)], 8),
        "k4": ([(a, b) for a, b in itertools.combinations(range(4), 2)], 24),
        "tri": ([(0, 1), (1, 2), (0, 2)], 6),
    }
    for name, (edges, aut) in shapes.items():
        q = "".join(f"v {i} 1\n" for i in range(4 if name != "tri" else 3))
        q += "".join(f"e {a} {b} 1\n" for a, b in edges)
        unrestricted = og.mine(q.replace("v ", "v "), "brute").count  # restricted count
        mono = _monomorphisms(ds, edges, 4 if name != "tri" else 3)
        assert unrestricted * aut == mono, name


def _monomorphisms(ds, qedges, k):
    e = ds.edges
    adj = {}
    for a, b, _ in e:
        adj.setdefault(int(a), set()).add(int(b))
        adj.setdefault(int(b), set()).add(int(a))
    verts = sorted(adj)
    count = 0

    def rec(f):
        nonlocal count
        u = len(f)
        if u == k:
            count += 1
            return
        cand = verts
        for a, b in qedges:
            if b == u and a < u:
                cand = [x for x in cand if x in adj[f[a]]]
            if a == u and b < u:
                cand = [x for x in cand if x in adj[f[b]]]
        for x in cand:
            if x not in f:
                rec(f + [x])

    rec([])
    return count


# ----------------------------------------------------------------- oracle ----

def test_oracle_modes_agree_and_match_brute_force():
    ds = G.Dataset.database(60, 6, 14, 3.5, 0.5, 3, 0.0, 5)
    ds.add_walk_queries(12, 3, 5, seed=3)
    ds.add_walk_queries(6, 4, 5, induced=True, seed=4)
    og = O.OracleGraph.of(ds)
    for q in ds.queries:
        c = og.mine(q, "count")
        f = og.mine(q, "faithful")
        b = og.mine(q, "brute")
        assert c.count == f.count == b.count, q
        assert c.graph_ids == f.graph_ids == b.graph_ids, q


def test_oracle_closed_forms_same_label():
    ds = G.Dataset.chung_lu(2000, 2.3, 10.0, 0.0, 1, 0.0, 21)
    og = O.OracleGraph.of(ds)
    e = ds.edges
    n = ds.nv
    deg = np.bincount(np.concatenate([e[:, 0], e[:, 1]]).astype(np.int64), minlength=n + 1)
    # edges = m, 2-paths = sum C(d,2), stars K1,3 = sum C(d,3)
    q_edge = "v 0 1\nv 1 1\ne 0 1 1\n"
    assert og.mine(q_edge).count == len(e)
    q_p3 = "v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 1 2 1\n"
    assert og.mine(q_p3).count == int(sum(math.comb(int(d), 2) for d in deg))
    q_s3 = "v 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\n"
    assert og.mine(q_s3).count == int(sum(math.comb(int(d), 3) for d in deg))
    # triangles = trace(A^3)/6
    import scipy.sparse as sp
    A = sp.coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n + 1, n + 1))
    A = ((A + A.T) > 0).astype(np.int64).tocsr()
    tri = int((A @ A).multiply(A).sum() // 6)
    q_tri = "v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 1 2 1\ne 0 2 1\n"
    assert og.mine(q_tri).count == tri


def test_oracle_label_filtered_closed_form():
    # 2-path a-b-c with three distinct labels: sum over centres of |N_a| * |N_c|
    ds = G.Dataset.database(200, 10, 30, 4.0, 0.5, 3, 0.0, 9)
    og = O.OracleGraph.of(ds)
    vl, e = ds.vlabel, ds.edges
    nb = {}
    for a, b, _ in e:
        nb.setdefault(int(a), []).append(int(b))
        nb.setdefault(int(b), []).append(int(a))
    expect = 0
    for v, ns in nb.items():
        if vl[v] != 2:
            continue
        expect += sum(1 for x in ns if vl[x] == 1) * sum(1 for x in ns if vl[x] == 3)
    q = "v 0 1\nv 1 2\nv 2 3\ne 0 1 1\ne 1 2 1\n"
    assert og.mine(q).count == expect


def test_oracle_algorithmic_bytes_formula_on_one_edge():
    # one edge query u0-u1 (different labels): B = sum over start vertices of
    # 4 + 8 + 4 |N_l(u1)(v)|
    ds = G.Dataset.database(50, 5, 9, 3.0, 0.0, 2, 0.0, 12)
    og = O.OracleGraph.of(ds)
    vl, e = ds.vlabel, ds.edges
    starts = [v for v in range(1, ds.nv + 1) if vl[v] == 1]
    nbr2 = {v: 0 for v in starts}
    for a, b, _ in e:
        for x, y in ((int(a), int(b)), (int(b), int(a))):
            if vl[x] == 1 and vl[y] == 2:
                nbr2[x] += 1
    expect = sum(4 + 8 + 4 * nbr2[v] for v in starts)
    q = "v 0 1\nv 1 2\ne 0 1 1\n"
    assert og.mine(q).alg_bytes == expect


def _plan_order_and_restricts(q):
    """Column order (vertices as the plan's E phases introduce them) and the
    restriction map of the planner (restrict: u<v means f(u) < f(v))."""
    order, res = [], {}
    for line in G.plan_dump(q).splitlines():
        if line.startswith("restrict: "):
            a, b = line[len("restrict: "):].split("<")
            if int(b) != 0xffffffff:
                res[int(a)] = int(b)
        elif line.startswith("svid:"):
            toks = line.replace("svid:", " ").replace("evid:", " ").split()
            for i in range(0, len(toks), 2):
                s_, e_ = int(toks[i]), int(toks[i + 1])
                if not order:
                    order.append(s_)
                if e_ not in order:
                    order.append(e_)
    return order, res


def test_oracle_rows_match_brute_force_embeddings():
    """The faithful mode's final rows (the table the row-multiset comparisons
    use) equal every map of Appendix-B semantics enumerated here directly:
    injective, label-preserving, every query edge on a data edge of its label,
    the planner's restrictions."""
    ds = G.Dataset.database(40, 5, 10, 3.5, 0.5, 3, 0.0, 17)
    ds.add_walk_queries(10, 3, 5, seed=8)
    ds.add_walk_queries(4, 4, 5, induced=True, seed=9)
    og = O.OracleGraph.of(ds)
    vl = ds.vlabel
    adj = {}
    for a, b, l in ds.edges:
        adj.setdefault(int(a), set()).add((int(b), int(l)))
        adj.setdefault(int(b), set()).add((int(a), int(l)))
    for q in ds.queries:
        qv = {int(l.split()[1]): int(l.split()[2]) for l in q.splitlines() if l.startswith("v ")}
        qe = [tuple(map(int, l.split()[1:4])) for l in q.splitlines() if l.startswith("e ")]
        k = len(qv)
        _, res = _plan_order_and_restricts(q)
        found = []

        def rec(f):
            u = len(f)
            if u == k:
                if all(f[a] < f[b] for a, b in res.items()):
                    found.append(tuple(f))
                return
            for x in range(1, ds.nv + 1):
                if vl[x] != qv[u] or x in f:
                    continue
                if all((f[b if a == u else a], l) in adj.get(x, ()) for a, b, l in qe
                       if (a == u and b < u) or (b == u and a < u)):
                    rec(f + [x])

        rec([])
        got = og.rows(q)
        assert sorted(found) == [tuple(int(v) for v in r) for r in got], q


def test_oracle_algorithmic_bytes_two_edge_path():
    """B on a two-edge path of three labels, counted by hand here: the first
    extension reads each start row (4 B), the row-pointer pair and segment of
    its list (8 + 4|N|) and writes every survivor as a two-column row
    (8 B each); the last extension reads each two-column row (8 B) and the
    segments of its back neighbours, and writes nothing."""
    ds = G.Dataset.database(120, 6, 20, 4.0, 0.5, 3, 0.0, 31)
    og = O.OracleGraph.of(ds)
    vl = ds.vlabel
    nb = {}
    for a, b, _ in ds.edges:
        nb.setdefault(int(a), []).append(int(b))
        nb.setdefault(int(b), []).append(int(a))
    q = "v 0 1\nv 1 2\nv 2 3\ne 0 1 1\ne 1 2 1\n"
    order, res = _plan_order_and_restricts(q)
    assert not res
    lab = {0: 1, 1: 2, 2: 3}
    qadj = {0: {1}, 1: {0, 2}, 2: {1}}

    def seg(v, l):
        return [x for x in nb.get(v, []) if vl[x] == l]

    u0, u1, u2 = order
    expect = 0
    for v in range(1, ds.nv + 1):
        if vl[v] != lab[u0]:
            continue
        n1 = seg(v, lab[u1])
        expect += 4 + 8 + 4 * len(n1) + 8 * len(n1)       # read row, read list, write survivors
        for x in n1:
            f = {u0: v, u1: x}
            expect += 8                                    # read the two-column row
            for w in qadj[u2]:
                expect += 8 + 4 * len(seg(f[w], lab[u2]))  # each back neighbour's segment
    assert og.mine(q).alg_bytes == expect


def test_dv_phase_count_matches_reference_plan_dumps():
    """gh_dv_phase_count (tools/classifyExtPhase.cpp:27-38: E phases with two
    extension pairs) equals the count read off the reference cpumine's plan
    dump of every corpus query, and is 0 exactly when the SV plan has as many
    E phases as the DV plan."""
    cases = json.load(open(os.path.join(GOLDEN, "plans.json")))
    seen_dv = 0
    for c in cases:
        lines = c["ref_plan"].splitlines()
        want = 0
        for i, l in enumerate(lines):
            if l.startswith("phase ") and " type E " in l:
                pairs = lines[i + 1].count("svid:")
                want += pairs == 2
        got = G.dv_phase_count(c["query"])
        assert got == want, c["query"]
        seen_dv += got > 0
        ne = lambda sv: sum(1 for l in G.plan_dump(c["query"], single_v=sv).splitlines()
                            if l.startswith("phase ") and " type E " in l)
        assert (ne(True) == ne(False)) == (got == 0), c["query"]
    assert seen_dv > 50


def test_golden_counts_workload1_match_reference_cpumine():
    gold = _load("workload1.json")
    ds = configs.workload1()
    og = O.OracleGraph.of(ds)
    assert ds.queries == [c["query"] for c in gold["cases"]]
    for c in gold["cases"]:
        r = og.mine(c["query"])
        assert r.count == c["ref_count"] == c["count"]
        assert sorted(r.graph_ids) == c["graph_ids"]
        assert r.alg_bytes == c["alg_bytes"]


def test_golden_counts_enron_match_reference_cpumine():
    gold = _load("enron.json")
    ds = _enron()
    og = O.OracleGraph.of(ds)
    for c in gold["cases"]:
        assert og.mine(c["query"]).count == c["ref_count"], c["query"]


def test_golden_scaled_workloads_oracle():
    gold = _load("scaled.json")
    for w in gold["workloads"]:
        ds = configs.WORKLOADS[w["workload"]](scale=w["scale"], **w.get("kw", {}))
        og = O.OracleGraph.of(ds)
        for c in w["cases"][:8]:
            r = og.mine(c["query"], threads=4)
            assert r.count == c["count"], (w["workload"], c["query"])


# ---------------------------------------------------------------- formats ----

def _enron():
    import gzip
    import shutil
    import tempfile
    d = tempfile.mkdtemp()
    for f in ("Enron.mygraph", "Enron.metadata"):
        with gzip.open(os.path.join(GOLDEN, "enron", f + ".gz"), "rb") as src, \
                open(os.path.join(d, f), "wb") as dst:
            shutil.copyfileobj(src, dst)
    return G.Dataset.load(os.path.join(d, "Enron.mygraph"), os.path.join(d, "Enron.metadata"))


def test_enron_loader_decodes_the_text_graph():
    """loadDataBase + skip-index decoding reproduce Enron.g's labelled edge set."""
    import gzip
    ds = _enron()
    meta = {}
    with gzip.open(os.path.join(GOLDEN, "enron", "Enron.metadata.gz"), "rt") as f:
        for line in f:
            new, lab, old = map(int, line.split())
            meta[old] = (new, lab)
    want = set()
    vlab_txt = {}
    with gzip.open(os.path.join(GOLDEN, "enron", "Enron.g.gz"), "rt") as f:
        for line in f:
            if line.startswith("v "):
                _, v, l = line.split()[:3]
                vlab_txt[int(v)] = int(l)
            elif line.startswith("e "):
                a, b, l = map(int, line.split()[1:4])
                a, b = meta[a][0], meta[b][0]
                want.add((min(a, b), max(a, b), l))
    got = {(int(a), int(b), int(l)) for a, b, l in ds.edges}
    assert got == want
    vl = ds.vlabel
    for old, (new, lab) in meta.items():
        assert vl[new] == lab
        if old in vlab_txt:
            assert vlab_txt[old] == lab
    assert len(vlab_txt) > 0


def test_mygraph_writer_round_trip(tmp_path):
    ds = G.Dataset.database(300, 5, 40, 5.0, 0.5, 6, 1.0, 33)
    # hubs with long label groups exercise the skip index (>128 and >4096)
    vl = ds.vlabel
    hub_edges = [(1, v, 1) for v in range(2, min(ds.nv, 6000)) if True]
    e = np.concatenate([ds.edges, np.array(hub_edges, dtype=np.uint32)])
    big = G.Dataset.from_arrays(vl, e)
    prefix = str(tmp_path / "rt")
    big.write(prefix)
    back = G.Dataset.load(prefix + ".mygraph", prefix + ".metadata")
    norm = lambda E: {(min(int(a), int(b)), max(int(a), int(b)), int(l)) for a, b, l in E}
    assert norm(back.edges) == norm(e)
    assert np.array_equal(back.vlabel[1:], vl[1:])


def test_reorder_intervals_fits_scattered_multi_label_sources(tmp_path):
    """A graph whose eight edge labels give every partition thousands of
    scattered source runs: the partition block writer refuses it (more than
    256 intervals, tools/genDatabase.cpp:454-460), the Gray-code renumbering
    brings every partition within 256 runs, the renumbered files load back to
    the same graph under the recorded id map, and mining the loaded files
    gives the oracle's counts."""
    rng = np.random.default_rng(44)
    n = 3000
    vl = np.concatenate([[0], rng.integers(1, 4, size=n)]).astype(np.uint32)
    m = 12000
    a = rng.integers(1, n + 1, size=m)
    b = rng.integers(1, n + 1, size=m)
    lab = rng.integers(1, 9, size=m)
    keep = a != b
    edges = np.stack([a[keep], b[keep], lab[keep]], axis=1).astype(np.uint32)
    ds = G.Dataset.from_arrays(vl, edges)
    before = ds.partition_runs()
    assert max(before.values()) > 256
    with pytest.raises(G.GenevaError):
        ds.write(str(tmp_path / "raw"))
    worst = ds.reorder_intervals()
    after = ds.partition_runs()
    assert worst == max(after.values()) <= 256
    assert sorted(after) == sorted(before)
    old = ds.old_ids
    assert sorted(old[1:].tolist()) == list(range(1, n + 1))
    # same graph under the id map: labels and the edge set
    assert np.array_equal(ds.vlabel[1:], vl[old[1:]])
    new_of = np.zeros(n + 1, dtype=np.int64)
    new_of[old] = np.arange(n + 1)
    want = {(min(new_of[x], new_of[y]), max(new_of[x], new_of[y]), l) for x, y, l in edges.tolist()}
    got = {(min(x, y), max(x, y), l) for x, y, l in ds.edges.tolist()}
    assert got == want
    # the written files: headers within 256 intervals, a round trip through
    # loadDataBase, the .metadata's third column the original id
    ds.write(str(tmp_path / "re"))
    raw = np.fromfile(str(tmp_path / "re.mygraph"), dtype=np.uint32)
    off, nblk = 1, 0
    while off < len(raw):
        assert raw[off + 2] <= 256
        off += raw[off]
        nblk += 1
    assert nblk == len(after)
    meta = np.loadtxt(str(tmp_path / "re.metadata"), dtype=np.int64)
    assert np.array_equal(meta[:, 2], old[1:])
    back = G.Dataset.load(str(tmp_path / "re.mygraph"), str(tmp_path / "re.metadata"))
    assert {(min(x, y), max(x, y), l) for x, y, l in back.edges.tolist()} == got
    og_a, og_b = O.OracleGraph.of(ds), O.OracleGraph.of(back)
    for q in ("v 0 1\nv 1 2\ne 0 1 3\n", "v 0 1\nv 1 2\nv 2 3\ne 0 1 1\ne 1 2 2\n",
              "v 0 1\nv 1 1\nv 2 2\ne 0 1 4\ne 1 2 5\ne 0 2 6\n"):
        assert og_a.mine(q).count == og_b.mine(q).count
