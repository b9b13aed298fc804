"""ORACLE -- test infrastructure only.

CPU restatement of the reference's cpumine (src/cpumine.cpp) and planner
(src/genMatchOrderCPU.cpp), built into oracle/liboracle.so.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only to check or time the product path; the product never calls it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_CPUMINE = os.path.join(_HERE, "_ref", "cpumine")

MODES = {"count": 0, "faithful": 1, "brute": 2}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so missing; run `make -C oracle`")
        L = C.CDLL(LIB)
        u32p = C.POINTER(C.c_uint32)
        L.oracle_graph_create.restype = C.c_void_p
        L.oracle_graph_create.argtypes = [C.c_uint32, u32p, C.c_uint64, u32p]
        L.oracle_graph_destroy.argtypes = [C.c_void_p]
        L.oracle_plan.restype = C.c_uint64
        L.oracle_plan.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64]
        L.oracle_mine.restype = C.c_int
        L.oracle_mine.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                  C.c_uint64, u32p, C.c_uint32, C.POINTER(C.c_uint8), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_double)]
        L.oracle_rows.restype = C.c_int64
        L.oracle_rows.argtypes = [C.c_void_p, C.c_char_p, u32p, C.c_uint64, C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def plan(query_text):
    buf = C.create_string_buffer(1 << 16)
    lib().oracle_plan(query_text.encode(), buf, len(buf))
    return buf.value.decode()


class OracleResult:
    __slots__ = ("count", "alg_bytes", "seconds", "graph_ids", "nstart")

    def __init__(self, count, alg_bytes, seconds, graph_ids, nstart):
        self.count, self.alg_bytes, self.seconds = count, alg_bytes, seconds
        self.graph_ids, self.nstart = graph_ids, nstart

    def __repr__(self):
        return f"OracleResult(count={self.count}, alg_bytes={self.alg_bytes}, s={self.seconds:.3f})"


class OracleGraph:
    def __init__(self, vlabel, edges, graph_of=None, ngraphs=1):
        self.vlabel = np.ascontiguousarray(vlabel, dtype=np.uint32)
        self.edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 3)
        u32p = C.POINTER(C.c_uint32)
        self._h = lib().oracle_graph_create(len(self.vlabel) - 1, self.vlabel.ctypes.data_as(u32p),
                                            len(self.edges), self.edges.ctypes.data_as(u32p))
        self.graph_of = None if graph_of is None else np.ascontiguousarray(graph_of, dtype=np.uint32)
        self.ngraphs = ngraphs

    @staticmethod
    def of(ds):
        return OracleGraph(ds.vlabel, ds.edges, ds.graph_of if ds.ngraphs > 1 else None, ds.ngraphs)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.oracle_graph_destroy(h)

    def rows(self, query_text):
        """The faithful mode's embeddings as an (n, k) uint32 array, column u =
        data vertex of query vertex u, rows sorted."""
        k = C.c_uint32(0)
        n = lib().oracle_rows(self._h, query_text.encode(), None, 0, C.byref(k))
        if n < 0:
            raise ValueError("oracle: malformed query")
        out = np.zeros((n, k.value), dtype=np.uint32)
        if n:
            lib().oracle_rows(self._h, query_text.encode(), out.ctypes.data_as(C.POINTER(C.c_uint32)),
                              out.size, C.byref(k))
        return out

    def mine(self, query_text, mode="count", threads=0, start_begin=0, start_end=0, stride=1):
        cnt, by, ns = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        secs = C.c_double(0)
        bits = None
        gop = None
        if self.graph_of is not None and self.ngraphs > 1:
            bits = (C.c_uint8 * self.ngraphs)()
            gop = self.graph_of.ctypes.data_as(C.POINTER(C.c_uint32))
        rc = lib().oracle_mine(self._h, query_text.encode(), MODES[mode], threads, start_begin,
                               start_end, stride, gop, self.ngraphs if bits is not None else 0, bits,
                               C.byref(cnt), C.byref(by), C.byref(ns), C.byref(secs))
        if rc != 0:
            raise ValueError("oracle: malformed query")
        ids = frozenset(i for i in range(self.ngraphs) if bits[i]) if bits is not None else frozenset()
        return OracleResult(cnt.value, by.value, secs.value, ids, ns.value)
