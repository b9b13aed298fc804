"""geneva_amd: MI355X-native embedding-extension mining (subgraph matching on
labelled graphs), with the reference's loader / planner interfaces on the host
and hand-written gfx950 kernels behind a C ABI.

Python binds the host library (libgeneva_host.so, include/geneva_host.h) with
ctypes.  Mining always runs the HIP kernels; there is no CPU fallback, and an
Engine cannot be created when the native libraries or a GPU are missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libgeneva_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libgeneva_hip.so")

GV_OK = 0
STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "device allocation failed",
          -4: "output capacity exceeded", -5: "malformed partition block", -6: "no HIP device"}

_lib = None


class GenevaError(RuntimeError):
    pass


def _u32p():
    return C.POINTER(C.c_uint32)


def lib():
    """Load libgeneva_host.so (and with it libgeneva_hip.so); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HOST_LIB) or not os.path.exists(HIP_LIB):
        raise GenevaError(f"native libraries missing under {LIB_DIR}; run `python -m geneva_amd.build`")
    L = C.CDLL(HOST_LIB)
    vp, u32, u64, i32, dbl, cs = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_char_p
    sig = {
        "gh_last_error": (cs, []),
        "gh_ds_gen_database": (vp, [u32, u32, u32, dbl, dbl, u32, u32, dbl, u64]),
        "gh_ds_gen_chung_lu": (vp, [u32, dbl, dbl, dbl, u32, dbl, u64]),
        "gh_ds_load": (vp, [cs, cs]),
        "gh_ds_from_arrays": (vp, [u32, _u32p(), u64, _u32p(), _u32p(), u32]),
        "gh_ds_free": (None, [vp]),
        "gh_ds_add_walk_queries": (i32, [vp, u32, u32, u32, i32, u32, u64, u32]),
        "gh_ds_add_shape_queries": (i32, [vp, u64]),
        "gh_ds_add_query": (i32, [vp, cs]),
        "gh_ds_nv": (u32, [vp]),
        "gh_ds_vlabel": (_u32p(), [vp]),
        "gh_ds_nedges": (u64, [vp]),
        "gh_ds_edges": (_u32p(), [vp]),
        "gh_ds_graph_of": (_u32p(), [vp]),
        "gh_ds_ngraphs": (u32, [vp]),
        "gh_ds_nqueries": (u32, [vp]),
        "gh_ds_query": (cs, [vp, u32]),
        "gh_ds_write": (i32, [vp, cs, i32]),
        "gh_ds_reorder_intervals": (i32, [vp]),
        "gh_ds_partition_runs": (i32, [vp, _u32p(), _u32p(), u32]),
        "gh_ds_old_ids": (_u32p(), [vp]),
        "gh_plan_dump": (u64, [cs, i32, C.c_char_p, u64]),
        "gh_dv_phase_count": (i32, [cs]),
        "gh_engine_create": (vp, [vp, i32, u64, C.POINTER(i32)]),
        "gh_engine_free": (None, [vp]),
        "gh_engine_graph_bytes": (u64, [vp]),
        "gh_engine_mine": (i32, [vp, cs, i32, C.POINTER(u64), C.POINTER(u64), C.POINTER(dbl),
                                 _u32p(), u32, C.POINTER(u64)]),
        "gh_engine_mine_shard": (i32, [vp, cs, i32, u64, u64, u64, C.POINTER(u64), C.POINTER(u64),
                                       C.POINTER(dbl), _u32p(), u32, C.POINTER(u64)]),
        "gh_engine_count_by_rows": (i32, [vp, i32]),
        "gh_engine_time_limit": (i32, [vp, dbl]),
        "gh_engine_last_progress": (i32, [vp, C.POINTER(i32), C.POINTER(dbl)]),
        "gh_engine_batch_config": (i32, [vp, u32, u64, u64, u32, u32]),
        "gh_engine_batch_mine": (i32, [vp, C.POINTER(cs), u32, i32, C.POINTER(u64), C.POINTER(u64),
                                       _u32p(), u32, C.POINTER(dbl)]),
        "gh_engine_autotune": (i32, [vp, C.POINTER(cs), u32, C.POINTER(u32), C.POINTER(u32),
                                     C.POINTER(dbl)]),
        "gh_engine_autotune_classes": (u64, [vp, C.c_char_p, u64]),
        "gh_engine_serve": (i32, [vp, C.POINTER(cs), C.POINTER(dbl), u32, i32, u64, C.POINTER(dbl),
                                  C.POINTER(u64), _u32p(), u32, C.POINTER(dbl)]),
        "gh_phase_mine": (i32, [cs, cs, cs, i32, i32, u64, u32, i32, C.POINTER(vp)]),
        "gh_phase_count": (u64, [vp]),
        "gh_phase_width": (u32, [vp]),
        "gh_phase_order": (_u32p(), [vp]),
        "gh_phase_rows": (_u32p(), [vp]),
        "gh_phase_launches": (u64, [vp]),
        "gh_phase_spills": (u64, [vp]),
        "gh_phase_single_v": (i32, [vp]),
        "gh_phase_free": (None, [vp]),
        "gh_engine_batch_profile": (i32, [vp, i32]),
        "gh_engine_batch_profile_bytes": (i32, [vp, C.POINTER(u64), C.POINTER(u64)]),
        "gh_engine_batch_pair_start": (i32, [vp, i32]),
        "gh_engine_batch_row_cap": (i32, [vp, C.POINTER(u64)]),
        "gh_engine_batch_share": (i32, [vp, i32]),
        "gh_engine_batch_share_stats": (i32, [vp, C.POINTER(u64), C.POINTER(u64)]),
        "gh_engine_batch_fuse_sizing": (i32, [vp, i32]),
        "gh_engine_batch_fused_rows": (i32, [vp, C.POINTER(u64)]),
        "gh_engine_batch_history": (i32, [vp, C.POINTER(u64), C.POINTER(u64), u32, C.POINTER(u32)]),
        "gh_engine_profile_records": (i32, [vp, C.POINTER(i32), C.POINTER(C.c_float), u64, C.POINTER(u64)]),
        "gh_engine_profile": (i32, [vp, i32]),
        "gh_engine_profile_reset": (i32, [vp]),
        "gh_engine_profile_read": (i32, [vp, i32, C.POINTER(u64), C.POINTER(dbl), C.POINTER(u64)]),
        "gh_engine_profile_units": (i32, [vp, i32, C.POINTER(u64)]),
        "gh_engine_profile_alg_bytes": (i32, [vp, i32, C.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _err(prefix, rc=None):
    msg = lib().gh_last_error().decode(errors="replace")
    if rc is not None:
        msg = f"{STATUS.get(rc, rc)}: {msg}"
    return GenevaError(f"{prefix}: {msg}")


class Dataset:
    """A labelled data graph (ids 1..nv), owner graph per vertex and queries."""

    def __init__(self, handle):
        if not handle:
            raise _err("dataset")
        self._h = handle

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.gh_ds_free(h)

    @staticmethod
    def database(ngraphs, nmin, nmax, avg_deg, ba_fraction, nlabels, zipf_s, seed, ba_m=2):
        return Dataset(lib().gh_ds_gen_database(ngraphs, nmin, nmax, avg_deg, ba_fraction, ba_m,
                                                nlabels, zipf_s, seed))

    @staticmethod
    def chung_lu(nv, gamma, mean_deg, wmax, nlabels, zipf_s, seed):
        return Dataset(lib().gh_ds_gen_chung_lu(nv, gamma, mean_deg, wmax, nlabels, zipf_s, seed))

    @staticmethod
    def load(mygraph, metadata):
        return Dataset(lib().gh_ds_load(mygraph.encode(), metadata.encode()))

    @staticmethod
    def from_arrays(vlabel, edges, graph_of=None, ngraphs=0):
        vlabel = np.ascontiguousarray(vlabel, dtype=np.uint32)
        edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 3)
        go = None if graph_of is None else np.ascontiguousarray(graph_of, dtype=np.uint32)
        p = lambda a: a.ctypes.data_as(_u32p()) if a is not None else None
        return Dataset(lib().gh_ds_from_arrays(len(vlabel) - 1, p(vlabel), len(edges), p(edges),
                                               p(go), ngraphs if go is not None else 0))

    def add_walk_queries(self, nq, kmin, kmax, induced=False, min_extra=0, seed=0, min_deg=0):
        rc = lib().gh_ds_add_walk_queries(self._h, nq, kmin, kmax, int(induced), min_extra, seed,
                                          min_deg)
        if rc:
            raise _err("add_walk_queries", rc)
        return self

    def add_shape_queries(self, seed):
        lib().gh_ds_add_shape_queries(self._h, seed)
        return self

    def add_query(self, text):
        lib().gh_ds_add_query(self._h, text.encode())
        return self

    @property
    def nv(self):
        return lib().gh_ds_nv(self._h)

    @property
    def vlabel(self):
        return np.ctypeslib.as_array(lib().gh_ds_vlabel(self._h), shape=(self.nv + 1,)).copy()

    @property
    def edges(self):
        m = lib().gh_ds_nedges(self._h)
        if m == 0:
            return np.zeros((0, 3), dtype=np.uint32)
        return np.ctypeslib.as_array(lib().gh_ds_edges(self._h), shape=(m, 3)).copy()

    @property
    def graph_of(self):
        return np.ctypeslib.as_array(lib().gh_ds_graph_of(self._h), shape=(self.nv + 1,)).copy()

    @property
    def ngraphs(self):
        return lib().gh_ds_ngraphs(self._h)

    @property
    def queries(self):
        L = lib()
        return [L.gh_ds_query(self._h, i).decode() for i in range(L.gh_ds_nqueries(self._h))]

    def reorder_intervals(self):
        """Renumber vertices so every edge-label partition's sources form at
        most 256 id runs; returns the largest run count."""
        rc = lib().gh_ds_reorder_intervals(self._h)
        if rc < 0:
            raise _err("reorder_intervals", rc)
        return rc

    def partition_runs(self):
        """{edge label: source id runs its partition block needs}."""
        n = lib().gh_ds_partition_runs(self._h, None, None, 0)
        lab = (C.c_uint32 * max(n, 1))()
        run = (C.c_uint32 * max(n, 1))()
        lib().gh_ds_partition_runs(self._h, lab, run, n)
        return {lab[i]: run[i] for i in range(n)}

    @property
    def old_ids(self):
        p = lib().gh_ds_old_ids(self._h)
        if not p:
            return np.arange(self.nv + 1, dtype=np.uint32)
        return np.ctypeslib.as_array(p, shape=(self.nv + 1,)).copy()

    def write(self, prefix, skip_index=False):
        rc = lib().gh_ds_write(self._h, prefix.encode(), int(skip_index))
        if rc:
            raise _err("write", rc)


def plan_dump(query_text, single_v=False):
    """Planner output in the reference cpumine print layout."""
    L = lib()
    n = L.gh_plan_dump(query_text.encode(), int(single_v), None, 0)
    if n == 0:
        raise _err("plan")
    buf = C.create_string_buffer(n + 1)
    L.gh_plan_dump(query_text.encode(), int(single_v), buf, n + 1)
    return buf.value.decode()


def dv_phase_count(query_text):
    n = lib().gh_dv_phase_count(query_text.encode())
    if n < 0:
        raise _err("dv_phase_count", n)
    return n


class PhaseRun:
    """Result of the reference-shaped phase driver: count, the final rows as an
    (n, k) uint32 array with column u = data vertex of query vertex u (None
    unless rows were kept), launcher calls and spills to host memory."""
    __slots__ = ("count", "rows", "launches", "spills", "single_v")

    def __init__(self, count, rows, launches, spills, single_v):
        self.count, self.rows, self.launches, self.spills = count, rows, launches, spills
        self.single_v = single_v

    def __repr__(self):
        return f"PhaseRun(count={self.count}, launches={self.launches}, spills={self.spills})"


def phase_mine(mygraph, metadata, query_file, single_v=False, device=0, arena_bytes=0, max_rows=0,
               keep_rows=True):
    """Mine the query in query_file over a .mygraph/.metadata pair the way the
    reference drivers do (src/gpumine.cpp, src/gpumine_singleV.cpp), through
    the reference-named launchers (include/geneva_launchers.hpp) on the
    reference's device layout.  single_v: False (DV plan), True (SV plan) or
    None (DV when the DV plan has a two-vertex phase).  max_rows caps
    maxRowNum (forces the overflow protocol on small inputs)."""
    L = lib()
    h = C.c_void_p()
    sv = -1 if single_v is None else int(bool(single_v))
    rc = L.gh_phase_mine(mygraph.encode(), metadata.encode(), query_file.encode(), sv, device,
                         arena_bytes, max_rows, int(keep_rows), C.byref(h))
    if rc:
        raise _err("phase_mine", rc)
    try:
        n, k = L.gh_phase_count(h), L.gh_phase_width(h)
        rows = None
        if keep_rows:
            rows = np.zeros((n, k), dtype=np.uint32)
            if n:
                order = np.ctypeslib.as_array(L.gh_phase_order(h), shape=(k,)).copy()
                raw = np.ctypeslib.as_array(L.gh_phase_rows(h), shape=(n, k))
                rows[:, order] = raw
        return PhaseRun(n, rows, L.gh_phase_launches(h), L.gh_phase_spills(h),
                        bool(L.gh_phase_single_v(h)))
    finally:
        L.gh_phase_free(h)


def sorted_rows(rows):
    """Rows sorted lexicographically (for multiset comparisons)."""
    if rows is None or len(rows) == 0:
        return rows
    return rows[np.lexsort(rows.T[::-1])]


class MineResult:
    """count (uint64), alg_bytes, device ms, launches, and the matched graph ids
    (decoded from the bitset on first access)."""
    __slots__ = ("count", "alg_bytes", "ms", "launches", "_words", "_ngraphs", "_ids")

    def __init__(self, count, alg_bytes, ms, words, launches, ngraphs=0):
        self.count, self.alg_bytes, self.ms, self.launches = count, alg_bytes, ms, launches
        self._words, self._ngraphs, self._ids = words, ngraphs, None

    @property
    def graph_ids(self):
        if self._ids is None:
            self._ids = _bits_to_ids(self._words, self._ngraphs) if self._words is not None \
                else frozenset()
        return self._ids

    def __repr__(self):
        return (f"MineResult(count={self.count}, alg_bytes={self.alg_bytes}, ms={self.ms:.3f}, "
                f"graphs={len(self.graph_ids)}, launches={self.launches})")


def _bits_to_ids(words, ngraphs):
    if ngraphs <= 1:
        return frozenset()
    bits = np.unpackbits(np.asarray(words, dtype=np.uint32).view(np.uint8), bitorder="little")
    return frozenset(np.nonzero(bits[:ngraphs])[0].tolist())


class Engine:
    """A data set's graph resident in HBM on one GPU, with the plain mining path."""

    def __init__(self, ds, device=0, level_bytes=1 << 30):
        st = C.c_int(0)
        self._h = lib().gh_engine_create(ds._h, device, level_bytes, C.byref(st))
        if not self._h:
            raise _err("engine", st.value)
        self.ngraphs = ds.ngraphs
        self._ds = ds

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.gh_engine_free(h)

    def __del__(self):
        self.close()

    @property
    def graph_bytes(self):
        return lib().gh_engine_graph_bytes(self._h)

    KERNELS = {0: "k_size", 1: "k_scan", 2: "k_walk<extend>", 3: "k_walk<count>",
               4: "batch_iteration", 5: "k_count_rows", 6: "k_isect", 7: "k_count_chain",
               8: "ref_launchers"}

    def batch_config(self, slots=0, slot_bytes=0, row_cap=0, ipt=0, blocks_per_cu=0):
        """Configure the batch engine (0 keeps a setting); rebuilt on next use."""
        rc = lib().gh_engine_batch_config(self._h, slots, slot_bytes, row_cap, ipt, blocks_per_cu)
        if rc:
            raise _err("batch_config", rc)

    @staticmethod
    def _cstrs(queries):
        arr = (C.c_char_p * len(queries))(*[q.encode() for q in queries])
        return arr

    def mine_batch(self, queries, single_v=False):
        """Mine all queries together on the batch engine.  Returns (results, stats)."""
        n = len(queries)
        # numpy buffers (calloc'd): a ctypes array of n * nwords words is
        # zeroed element by element, about 2 ms for 256 queries on 100,000 graphs
        counts = np.zeros(max(n, 1), dtype=np.uint64)
        ab = np.zeros(max(n, 1), dtype=np.uint64)
        nwords = (self.ngraphs + 31) // 32 if self.ngraphs > 1 else 0
        words = np.zeros(max(n * nwords, 1), dtype=np.uint32)
        st = (C.c_double * 8)()
        u64p = C.POINTER(C.c_uint64)
        rc = lib().gh_engine_batch_mine(self._h, self._cstrs(queries), n, int(single_v),
                                        counts.ctypes.data_as(u64p), ab.ctypes.data_as(u64p),
                                        words.ctypes.data_as(_u32p()), nwords, st)
        if rc:
            raise _err("mine_batch", rc)
        rows = words[:n * nwords].reshape(n, nwords) if nwords else [None] * n
        cl, al = counts.tolist(), ab.tolist()
        res = [MineResult(cl[i], al[i], 0.0, rows[i] if nwords else None, 0, self.ngraphs)
               for i in range(n)]
        stats = dict(zip(("batched", "overflowed", "graph_launches", "device_ms", "wall_ms", "compile_ms",
                          "fallback_ms", "pieces"), list(st)))
        return res, stats

    def autotune(self, queries):
        ipt, bpc, ms = C.c_uint32(0), C.c_uint32(0), C.c_double(0)
        rc = lib().gh_engine_autotune(self._h, self._cstrs(queries), len(queries), C.byref(ipt),
                                      C.byref(bpc), C.byref(ms))
        if rc:
            raise _err("autotune", rc)
        n = lib().gh_engine_autotune_classes(self._h, None, 0)
        buf = C.create_string_buffer(n + 1)
        lib().gh_engine_autotune_classes(self._h, buf, n + 1)
        classes = {}
        for ent in buf.value.decode().split(";"):
            if ent:
                k, i, m = ent.split(":")
                classes[k] = {"ipt": int(i), "device_ms": round(float(m), 4)}
        cap = C.c_uint64(0)
        lib().gh_engine_batch_row_cap(self._h, C.byref(cap))
        return {"ipt": ipt.value, "blocks_per_cu": bpc.value, "row_cap": cap.value, "device_ms": ms.value,
                "classes": classes}

    def serve(self, queries, arrival_s, batched=True, mem_budget=16 << 30, with_graphs=False):
        """Serve requests arriving at arrival_s (seconds).  Returns (latency_ms,
        counts, stats), and with with_graphs the matched graph ids of every
        request as a fourth item."""
        n = len(queries)
        arr = (C.c_double * max(n, 1))(*arrival_s)
        lat = (C.c_double * max(n, 1))()
        cnt = (C.c_uint64 * max(n, 1))()
        st = (C.c_double * 6)()
        nwords = (self.ngraphs + 31) // 32 if (with_graphs and self.ngraphs > 1) else 0
        words = np.zeros(max(n * nwords, 1), dtype=np.uint32)
        rc = lib().gh_engine_serve(self._h, self._cstrs(queries), arr, n, 0 if batched else 1,
                                   mem_budget, lat, cnt, words.ctypes.data_as(_u32p()) if nwords else None,
                                   nwords, st)
        if rc:
            raise _err("serve", rc)
        stats = dict(zip(("wall_s", "iterations", "max_concurrent", "overflowed", "direct",
                          "max_batched"), list(st)))
        if not with_graphs:
            return list(lat)[:n], list(cnt)[:n], stats
        ids = [_bits_to_ids(words[i * nwords:(i + 1) * nwords], self.ngraphs) if nwords else frozenset()
               for i in range(n)]
        return list(lat)[:n], list(cnt)[:n], stats, ids

    def batch_profile(self, on=True):
        """Per-kernel algorithmic bytes of the batch engine (re-captures its
        iteration); with profile() on, batch iterations run kernel by kernel
        under HIP events."""
        rc = lib().gh_engine_batch_profile(self._h, int(on))
        if rc:
            raise _err("batch_profile", rc)

    def batch_profile_bytes(self):
        """(B charged to k_b_size, B charged to k_b_walk) since batch_profile(True)."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        rc = lib().gh_engine_batch_profile_bytes(self._h, C.byref(a), C.byref(b))
        if rc:
            raise _err("batch_profile_bytes", rc)
        return a.value, b.value

    def time_limit(self, seconds):
        """Stop plain-path queries after `seconds` of wall time (0 = no limit);
        see last_progress()."""
        rc = lib().gh_engine_time_limit(self._h, float(seconds))
        if rc:
            raise _err("time_limit", rc)

    def last_progress(self):
        """(finished, fraction of the first extension's candidates mined) of the
        last plain-path query."""
        f, x = C.c_int(0), C.c_double(0)
        lib().gh_engine_last_progress(self._h, C.byref(f), C.byref(x))
        return bool(f.value), x.value

    def batch_pair_start(self, on=True):
        """Start qualifying batch queries on the label-pair edge tables."""
        rc = lib().gh_engine_batch_pair_start(self._h, int(on))
        if rc:
            raise _err("batch_pair_start", rc)

    def batch_share(self, on=True):
        """Let batch queries with the same first steps share one table
        (applies to queries admitted from now on)."""
        rc = lib().gh_engine_batch_share(self._h, int(on))
        if rc:
            raise _err("batch_share", rc)

    def batch_share_stats(self):
        """(tables shared, followers restarted on their own) since the batch
        engine was made."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        rc = lib().gh_engine_batch_share_stats(self._h, C.byref(a), C.byref(b))
        if rc:
            raise _err("batch_share_stats", rc)
        return a.value, b.value

    def batch_fuse_sizing(self, on=True):
        """Size batch slots that read one input table together (on by
        default; the results are the same either way)."""
        rc = lib().gh_engine_batch_fuse_sizing(self._h, int(on))
        if rc:
            raise _err("batch_fuse_sizing", rc)

    def batch_fused_rows(self):
        """Rows sized as part of another slot's sizing group since the batch
        engine was made."""
        a = C.c_uint64(0)
        rc = lib().gh_engine_batch_fused_rows(self._h, C.byref(a))
        if rc:
            raise _err("batch_fused_rows", rc)
        return a.value

    def batch_history(self):
        """[(rows sized, candidates walked)] of each batch iteration run while
        profiling was on; clears the history."""
        n = C.c_uint32(0)
        lib().gh_engine_batch_history(self._h, None, None, 0, C.byref(n))
        r = (C.c_uint64 * max(n.value, 1))()
        c = (C.c_uint64 * max(n.value, 1))()
        lib().gh_engine_batch_history(self._h, r, c, n.value, C.byref(n))
        return list(zip(list(r)[:n.value], list(c)[:n.value]))

    def profile_records(self):
        """[(kernel kind name, device ms)] of every profiled launch, in order."""
        n = C.c_uint64(0)
        lib().gh_engine_profile_records(self._h, None, None, 0, C.byref(n))
        k = (C.c_int * max(n.value, 1))()
        m = (C.c_float * max(n.value, 1))()
        lib().gh_engine_profile_records(self._h, k, m, n.value, C.byref(n))
        return [(self.KERNELS.get(k[i], str(k[i])), m[i]) for i in range(n.value)]

    def count_by_rows(self, on=True):
        """Count single-back-edge last steps per row (default) or by a candidate walk."""
        rc = lib().gh_engine_count_by_rows(self._h, int(on))
        if rc:
            raise _err("count_by_rows", rc)

    def profile(self, on=True, count_bytes=True):
        """Per-kernel profiling: HIP events around every launch, and with
        count_bytes the kernels' requested bytes and work units (counting
        variants of the kernels)."""
        rc = lib().gh_engine_profile(self._h, (2 if count_bytes else 1) if on else 0)
        if rc:
            raise _err("profile", rc)
        lib().gh_engine_profile_reset(self._h)

    def profile_read(self):
        """{kernel: (launches, total device ms, requested bytes, units, algorithmic
        bytes B)} since the last reset; units are rows (k_size, k_scan,
        k_count_rows) or candidates (k_walk)."""
        out = {}
        for k, name in self.KERNELS.items():
            n, ms, rq = C.c_uint64(0), C.c_double(0), C.c_uint64(0)
            rc = lib().gh_engine_profile_read(self._h, k, C.byref(n), C.byref(ms), C.byref(rq))
            if rc:
                raise _err("profile_read", rc)
            u, b = C.c_uint64(0), C.c_uint64(0)
            lib().gh_engine_profile_units(self._h, k, C.byref(u))
            lib().gh_engine_profile_alg_bytes(self._h, k, C.byref(b))
            out[name] = (n.value, ms.value, rq.value, u.value, b.value)
        return out

    def mine(self, query_text, single_v=False, start_begin=0, start_end=0, stride=1):
        """Mine one query on the plain path.  start_begin/start_end/stride pick a
        shard of the first vertex's candidates (indices into the id-sorted list
        of vertices with its label; start_end 0 = all); shard counts of a
        partition add up to the whole count."""
        count, ab, launches = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        ms = C.c_double(0)
        nwords = (self.ngraphs + 31) // 32 if self.ngraphs > 1 else 0
        bits = np.zeros(max(nwords, 1), dtype=np.uint32)
        rc = lib().gh_engine_mine_shard(self._h, query_text.encode(), int(single_v), start_begin,
                                        start_end, stride, C.byref(count), C.byref(ab),
                                        C.byref(ms), bits.ctypes.data_as(_u32p()), nwords,
                                        C.byref(launches))
        if rc:
            raise _err("mine", rc)
        return MineResult(count.value, ab.value, ms.value, bits[:nwords] if nwords else None,
                          launches.value, self.ngraphs)
