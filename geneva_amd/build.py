"""Build the native libraries in-tree.

    python -m geneva_amd.build            # product libraries + oracle
    python -m geneva_amd.build --ref      # also the reference cpumine (needs the
                                          # reference checkout; checker only)

Outputs
    geneva_amd/lib/libgeneva_hip.so   gfx950 kernels + C ABI (hipcc)
    geneva_amd/lib/libgeneva_host.so  loader, planner, drivers (g++, links the above)
    oracle/liboracle.so               CPU oracle (test infrastructure)
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "geneva_amd")
LIB = os.path.join(PKG, "lib")
HIP_SRC = [os.path.join(PKG, "csrc", "hip", f) for f in ("geneva_hip.hip", "batch.hip", "launchers.hip")]
HOST_SRC = [os.path.join(PKG, "csrc", "host", f) for f in
            ("genMatchOrder.cpp", "loadDataBase.cpp", "gen.cpp", "miner.cpp",
             "batch.cpp", "serve.cpp", "phase_driver.cpp", "host_capi.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GENEVA_ARCH", "gfx950")


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        print(out)
        raise RuntimeError("build step failed: " + cmd[0])
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libgeneva_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
          "-Wno-unused-result", "-Wno-unused-value",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-o", out + ".tmp"] + HIP_SRC)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    out = os.path.join(LIB, "libgeneva_host.so")
    srcs = [s for s in HOST_SRC if os.path.exists(s)]
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          "-Wno-sign-compare", "-o", out + ".tmp"] + srcs +
         ["-L" + LIB, "-lgeneva_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle(ref=False):
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "all"])
    if ref:
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build(ref=None):
    build_hip()
    build_host()
    if ref is None:
        ref = os.path.isdir("/root/reference/src")
    build_oracle(ref=ref)


if __name__ == "__main__":
    build(ref=True if "--ref" in sys.argv else None)
