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
