"""GPU tests of the reference-shaped launchers (include/geneva_launchers.h)
driven by the reference's phase loop (geneva_amd/csrc/host/phase_driver.cpp,
the shape of src/gpumine.cpp and src/gpumine_singleV.cpp) on the reference's
device layout: counts equal the reference cpumine's, and the final rows, as a
multiset, equal the oracle's faithful rows (the comparison
scripts/compareCPUandGPUemb.py makes between two runs)."""
import gzip
import json
import os
import shutil
import tempfile

import numpy as np
import pytest

import geneva_amd as G
from geneva_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu
ROWS_LIMIT = 400_000     # compare row multisets up to this many rows


def _query_file(d, i, text):
    p = os.path.join(d, f"q{i}.query")
    with open(p, "w") as f:
        f.write(text)
    return p


@pytest.fixture(scope="module")
def w1_files(require_gpu):
    ds = configs.workload1()
    d = tempfile.mkdtemp()
    ds.write(os.path.join(d, "w1"))
    yield ds, d
    shutil.rmtree(d, ignore_errors=True)


def test_workload1_phase_driver_counts_and_rows(w1_files):
    import oracle as O
    ds, d = w1_files
    og = O.OracleGraph.of(ds)
    gold = json.load(open(os.path.join(GOLDEN, "workload1.json")))
    for i, c in enumerate(gold["cases"]):
        qf = _query_file(d, i, c["query"])
        want = og.rows(c["query"])
        assert len(want) == c["ref_count"]
        auto = G.phase_mine(os.path.join(d, "w1.mygraph"), os.path.join(d, "w1.metadata"), qf,
                            single_v=None, keep_rows=False)
        assert auto.count == c["ref_count"]
        assert auto.single_v == (G.dv_phase_count(c["query"]) == 0)
        for sv in (True, False):
            # the reference bound, then a bound of 16 rows that makes every
            # phase overflow, spill and resume with isContinue
            for cap in (0, 16):
                print("workload1 query", i, "single_v", sv, "max_rows", cap, flush=True)
                r = G.phase_mine(os.path.join(d, "w1.mygraph"), os.path.join(d, "w1.metadata"), qf,
                                 single_v=sv, max_rows=cap, arena_bytes=64 << 20)
                assert r.count == c["ref_count"], (i, sv, cap, r)
                assert np.array_equal(G.sorted_rows(r.rows), want), (i, sv, cap)
                if cap:
                    assert r.spills > 0


def test_enron_phase_driver(require_gpu):
    import oracle as O
    d = tempfile.mkdtemp()
    try:
        for f in ("Enron.mygraph", "Enron.metadata"):
            with gzip.open(os.path.join(GOLDEN, "enron", f + ".gz"), "rb") as s, \
                    open(os.path.join(d, f), "wb") as t:
                shutil.copyfileobj(s, t)
        mg, md = os.path.join(d, "Enron.mygraph"), os.path.join(d, "Enron.metadata")
        ds = G.Dataset.load(mg, md)
        og = O.OracleGraph.of(ds)
        for i, c in enumerate(json.load(open(os.path.join(GOLDEN, "enron.json")))["cases"]):
            qf = _query_file(d, i, c["query"])
            small = c["ref_count"] <= ROWS_LIMIT
            for sv in (True, False):
                print("enron query", i, "single_v", sv, "rows", c["ref_count"], flush=True)
                r = G.phase_mine(mg, md, qf, single_v=sv, keep_rows=small, arena_bytes=2 << 30)
                assert r.count == c["ref_count"], (i, sv, r)
                if small:
                    assert np.array_equal(G.sorted_rows(r.rows), og.rows(c["query"])), (i, sv)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_copy_partial_emb_overlapping_and_disjoint(require_gpu):
    """gv_copy_partial_emb / gv_copyPartialEmb (src/reduction.cu:280-296):
    after the copy, dst holds what src held, for destinations above and below
    an overlapping source and for disjoint ranges (numpy's copy semantics)."""
    import ctypes as C
    hip = C.CDLL(G.HIP_LIB)
    ctx = C.c_void_p()
    assert hip.gv_ctx_create(0, C.byref(ctx)) == 0
    n_all = 1 << 20
    buf = C.c_void_p()
    assert hip.gv_malloc(ctx, C.c_uint64(4 * n_all), C.byref(buf)) == 0
    try:
        rng = np.random.default_rng(3)
        base = rng.integers(1, 1 << 31, size=n_all, dtype=np.uint32)
        cases = [(1000, 1500, 300_000), (1500, 1000, 300_000), (0, 400_000, 300_000),
                 (400_000, 0, 300_000), (10, 11, 500_000), (11, 10, 500_000), (5, 5, 100), (7, 9, 0)]
        for fn in ("gv_copy_partial_emb", "gv_copyPartialEmb"):
            f = getattr(hip, fn)
            for s, d, n in cases:
                print(fn, "src", s, "dst", d, "n", n, flush=True)
                assert hip.gv_memcpy_h2d(ctx, buf, base.ctypes.data_as(C.c_void_p), C.c_uint64(4 * n_all)) == 0
                src = C.c_void_p(buf.value + 4 * s)
                dst = C.c_void_p(buf.value + 4 * d)
                if fn == "gv_copy_partial_emb":
                    rc = f(ctx, src, dst, C.c_uint64(n))
                else:
                    rc = f(ctx, src, dst, C.c_uint32(n))
                assert rc == 0
                out = np.zeros(n_all, dtype=np.uint32)
                assert hip.gv_memcpy_d2h(ctx, out.ctypes.data_as(C.c_void_p), buf, C.c_uint64(4 * n_all)) == 0
                want = base.copy()
                want[d:d + n] = base[s:s + n]
                assert np.array_equal(out, want), (fn, s, d, n)
    finally:
        hip.gv_free(ctx, buf)
        hip.gv_ctx_destroy(ctx)
