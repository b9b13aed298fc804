"""CPU tests of the C-ABI boundary: both libraries load without a GPU and
export every function their headers declare; argument checks reject bad calls
without touching a device."""
import ctypes as C
import os
import re

import geneva_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(g[vh]_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _exports(path):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_hip_library_exports_every_declared_symbol():
    names = _declared("geneva_hip.h")
    assert len(names) >= 25
    exp = _exports(G.HIP_LIB)
    missing = [n for n in names if n not in exp]
    assert not missing, missing
    lib = C.CDLL(G.HIP_LIB)
    for n in names:
        getattr(lib, n)


def test_launcher_abi_exports_every_declared_symbol():
    """include/geneva_launchers.h: one entry per reference launcher
    (include/miningPattern.h:1-31, include/reduction.h:1-4), and the shim
    include/geneva_launchers.hpp defines the reference's names."""
    names = _declared("geneva_launchers.h")
    ref = ["initEmb_3V_1and2_NoHash", "initEmb_2V_NoHash", "extEmb_2V_2src_1and2_sameLabel_NoHash",
           "extEmb_2V_2src_1and2_notSameLabel_NoHash", "extEmb_2V_1src_1and2_sameLabel_NoHash",
           "extEmb_2V_1src_1and2_notSameLabel_NoHash", "extEmb_1V_NoHash", "reductionPhase",
           "copyPartialEmb"]
    assert sorted("gv_" + r for r in ref) == sorted(
        n for n in names if n not in ("gv_launchers_reserve", "gv_extEmb_1V_NoHash_noalloc"))
    exp = _exports(G.HIP_LIB)
    assert not [n for n in names if n not in exp]
    shim = open(os.path.join(ROOT, "include", "geneva_launchers.hpp")).read()
    for r in ref:
        assert re.search(r"inline void " + r + r"\(", shim), r


def test_launcher_argument_checks_without_a_device():
    lib = C.CDLL(G.HIP_LIB)
    # no context, no partition: rejected before any device work
    assert lib.gv_extEmb_1V_NoHash(None, None, None, None, None, None, None, None, None, 1, 0, 3,
                                   False, False, False, False, False, 10) == -1
    assert lib.gv_reductionPhase(None, None, None, None, None, None, None, None, 1, 0, 3, False) == -1
    assert lib.gv_initEmb_2V_NoHash(None, None, None, None, None, None, None, 300, 1, 1, 1, 0, 0,
                                    False, 10) == -1   # more than 256 intervals


def test_host_library_exports_every_declared_symbol():
    names = _declared("geneva_host.h")
    assert len(names) >= 20
    exp = _exports(G.HOST_LIB)
    missing = [n for n in names if n not in exp]
    assert not missing, missing


def test_argument_checks_without_a_device():
    lib = C.CDLL(G.HIP_LIB)
    lib.gv_status_string.restype = C.c_char_p
    assert lib.gv_status_string(0) == b"ok"
    assert lib.gv_ctx_create(0, None) == -1                    # null out pointer
    assert lib.gv_graph_destroy(None) == -1
    assert lib.gv_frontier_create(None, 10, None) == -1
    assert lib.gv_size_rows(None, None, None, None, 1, 1, None, None) == -1
    # batch engine entries on no batch: rejected, nothing touched
    rows = C.c_uint64(0)
    assert lib.gv_batch_fuse_sizing(None, 1) == -1
    assert lib.gv_batch_fused_rows(None, C.byref(rows)) == -1
    assert lib.gv_batch_share(None, 1) == -1
    assert lib.gv_batch_run(None, 1) == -1
    n = C.c_int(-1)
    assert lib.gv_device_count(C.byref(n)) == 0
    assert n.value >= 0


def test_engine_requires_a_gpu_and_never_falls_back():
    import pytest
    ds = G.Dataset.database(5, 5, 8, 2.0, 0.0, 2, 0.0, 1)
    n = C.c_int(0)
    C.CDLL(G.HIP_LIB).gv_device_count(C.byref(n))
    if n.value == 0:
        with pytest.raises(G.GenevaError):
            G.Engine(ds)


def test_count_rows_mode_selection():
    """gv_count_rows_mode: which last steps are counted per row."""
    import ctypes as C
    hip = C.CDLL(G.HIP_LIB)

    class Step(C.Structure):
        _fields_ = [("label", C.c_uint32), ("col", C.c_uint32), ("nback", C.c_uint8),
                    ("back_col", C.c_uint8 * 16), ("back_part", C.c_uint8 * 16),
                    ("restrict_col", C.c_uint8), ("ncoll", C.c_uint8), ("coll_col", C.c_uint8 * 16)]

    def step(label, back, restrict=0xff, coll=()):
        s = Step()
        s.label, s.nback, s.restrict_col, s.ncoll = label, len(back), restrict, len(coll)
        for i, b in enumerate(back):
            s.back_col[i] = b
        for i, c in enumerate(coll):
            s.coll_col[i] = c
        return s

    f = hip.gv_count_rows_mode
    f.argtypes = [C.POINTER(Step), C.POINTER(Step), C.c_uint32]
    f.restype = C.c_int
    L = 3
    st = step(2, [0], coll=(1,))
    assert f(C.byref(st), None, L) == 0
    assert f(C.byref(step(2, [0, 1])), None, L) == -1                          # two back edges
    assert f(C.byref(st), C.byref(step(5, [1])), L) == 1                       # tail off column 1
    assert f(C.byref(st), C.byref(step(5, [0], restrict=2)), L) == 1
    assert f(C.byref(st), C.byref(step(5, [L])), L) == -1                      # tail off v
    assert f(C.byref(st), C.byref(step(2, [0], restrict=L, coll=(1, L))), L) == 2
    assert f(C.byref(st), C.byref(step(2, [0], coll=(1, L))), L) == -1         # no restriction to v
    assert f(C.byref(st), C.byref(step(2, [1], restrict=L, coll=(1, L))), L) == -1  # other segment
    assert f(C.byref(st), C.byref(step(2, [0], restrict=L, coll=(L,))), L) == -1    # columns differ
