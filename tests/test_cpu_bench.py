"""CPU tests of bench.py's contract and of its traffic accounting: a request
for several GPUs ends with a message and status 2 before any device work, and
the PMC summary scales FETCH_SIZE / WRITE_SIZE by the copy calibration and
keeps only the accounting iterations when asked."""
import csv
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _csv(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for d, k, c, v in rows:
            w.writerow({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single MI355X" in r.stderr
    assert r.stdout.strip() == ""


def test_pmc_summary_scales_by_the_copy_calibration(tmp_path):
    bench = _bench()
    bench.ROOT = str(tmp_path)
    (tmp_path / "profiles").mkdir()
    prep = "(anonymous namespace)::k_b_prep((anonymous namespace)::BArgs)"
    size = "(anonymous namespace)::k_b_size((anonymous namespace)::BArgs)"
    copy = "(anonymous namespace)::k_copy_back(unsigned int*, unsigned long)"
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    cf, cw = tmp_path / "cf.csv", tmp_path / "cw.csv"
    # two iterations: k_b_prep, k_b_size; the second is the accounting one
    _csv(f, [(1, prep, "FETCH_SIZE", 1), (2, size, "FETCH_SIZE", 100), (3, prep, "FETCH_SIZE", 1),
             (4, size, "FETCH_SIZE", 300)])
    _csv(w, [(1, prep, "WRITE_SIZE", 1), (2, size, "WRITE_SIZE", 10), (3, prep, "WRITE_SIZE", 1),
             (4, size, "WRITE_SIZE", 50)])
    # a 1 MiB calibration moves 524,288 bytes each way per launch; FETCH_SIZE
    # reads half of them (256 KiB: scale 2), WRITE_SIZE all (512 KiB: scale 1)
    _csv(cf, [(1, copy, "FETCH_SIZE", 256), (2, copy, "FETCH_SIZE", 256)])
    _csv(cw, [(1, copy, "WRITE_SIZE", 512), (2, copy, "WRITE_SIZE", 512)])
    bench.pmc_summary([str(f), str(w), str(cf), str(cw)], 7, 1, last=1)
    out = json.load(open(tmp_path / "profiles" / "traffic_workload7.json"))
    assert out["fetch_scale"] == 2.0 and out["write_scale"] == 1.0
    assert out["k_b_size"]["launches"] == 1
    assert out["k_b_size"]["hbm_bytes_per_launch"] == 300 * 1024 * 2 + 50 * 1024
    # every dispatch: the mean per launch
    bench.pmc_summary([str(f), str(w), str(cf), str(cw)], 7, 1, last=0)
    out = json.load(open(tmp_path / "profiles" / "traffic_workload7.json"))
    assert out["k_b_size"]["launches"] == 2
    assert out["k_b_size"]["hbm_bytes_per_launch"] == 200 * 1024 * 2 + 30 * 1024
