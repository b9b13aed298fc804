"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
product planner, loader, writers and renumbering, and the oracle, built from
their sources with g++ -fsanitize=address,undefined (no GPU code) and run on
the plan corpus and the Enron fixture (tests/native/sanitize_main.cpp)."""
import gzip
import json
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
HOST = os.path.join(ROOT, "geneva_amd", "csrc", "host")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_is_clean_under_asan_and_ubsan():
    d = tempfile.mkdtemp()
    try:
        exe = os.path.join(d, "sanitize_main")
        srcs = [os.path.join(ROOT, "tests", "native", "sanitize_main.cpp"),
                os.path.join(HOST, "genMatchOrder.cpp"), os.path.join(HOST, "loadDataBase.cpp"),
                os.path.join(ROOT, "oracle", "oracle.cpp"), os.path.join(ROOT, "oracle", "oracle_planner.cpp")]
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-o", exe] + srcs
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        qf = os.path.join(d, "queries.txt")
        cases = json.load(open(os.path.join(GOLDEN, "plans.json")))
        with open(qf, "w") as f:
            f.write("\n%%\n".join(c["query"].rstrip("\n") for c in cases) + "\n")
        for name in ("Enron.mygraph", "Enron.metadata"):
            with gzip.open(os.path.join(GOLDEN, "enron", name + ".gz"), "rb") as s, \
                    open(os.path.join(d, name), "wb") as t:
                shutil.copyfileobj(s, t)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                   UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
        r = subprocess.run([exe, qf, os.path.join(d, "Enron.mygraph"), os.path.join(d, "Enron.metadata")],
                           capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
        assert "sanitize ok" in r.stdout
        print(r.stdout.strip())
    finally:
        shutil.rmtree(d, ignore_errors=True)
