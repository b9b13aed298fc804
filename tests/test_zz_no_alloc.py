"""GPU test of gv_extEmb_1V_NoHash_noalloc (the extension launcher whose
working arrays live in a fixed device array): through the C ABI, on the
reference's device layout, its rows and counters equal gv_extEmb_1V_NoHash's
on two calls in a row, the second with a smaller input, and a third call past
its row limit is refused."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import geneva_amd as G
from geneva_amd import configs

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, hip, ctx):
        self.hip, self.ctx, self.ptrs = hip, ctx, []

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.uint32)
        p = C.c_void_p()
        assert self.hip.gv_malloc(self.ctx, C.c_uint64(max(4 * arr.size, 4)), C.byref(p)) == 0
        if arr.size:
            assert self.hip.gv_memcpy_h2d(self.ctx, p, arr.ctypes.data_as(C.c_void_p), C.c_uint64(4 * arr.size)) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, n):
        out = np.zeros(n, dtype=np.uint32)
        if n:
            assert self.hip.gv_memcpy_d2h(self.ctx, out.ctypes.data_as(C.c_void_p), p, C.c_uint64(4 * n)) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.gv_free(self.ctx, p)


def test_no_alloc_extension_equals_the_allocating_one(require_gpu):
    hip = C.CDLL(G.HIP_LIB)
    ctx = C.c_void_p()
    assert hip.gv_ctx_create(0, C.byref(ctx)) == 0
    dev = Dev(hip, ctx)
    try:
        ds = configs.workload1()
        d = tempfile.mkdtemp()
        ds.write(os.path.join(d, "w1"))
        raw = np.fromfile(os.path.join(d, "w1.mygraph"), dtype=np.uint32)
        blk = raw[1:1 + raw[1]]                     # the one partition block (edge label 1)
        n_int, u4 = int(blk[2]), int(blk[4])
        vl = ds.vlabel.copy()
        vl[0] = ds.nv
        part = dev.put(blk[5:])
        a3 = C.c_void_p(part.value + 4 * u4)
        vlab = dev.put(vl)
        e = ds.edges
        rows = np.concatenate([e[:, :2], e[:, [1, 0]]]).astype(np.uint32)   # every (v, u) edge
        lab = int(np.bincount(vl[1:]).argmax())
        # new vertex: label `lab` from column 1's list, below column 0, unlike columns 0 and 1
        aux = np.zeros(64, dtype=np.uint32)
        aux[0] = aux[1] = lab
        aux[2] = aux[3] = 0
        aux[4] = aux[5] = 1
        aux[6] = aux[7] = 2
        aux[8], aux[9] = 0, 1
        auxd = dev.put(aux)
        ind = dev.put(rows.reshape(-1))
        cap = 4 * len(rows) * 8
        out_a, out_b = dev.put(np.zeros(cap * 3, np.uint32)), dev.put(np.zeros(cap * 3, np.uint32))
        cnt_a, cnt_b = dev.put(np.zeros(2, np.uint32)), dev.put(np.zeros(2, np.uint32))
        zero2 = np.zeros(2, np.uint32)
        u32 = C.c_uint32
        for nrows in (len(rows), len(rows) // 3):
            print("no-alloc extension, input rows", nrows, flush=True)
            outs = []
            for fn, out, cnt in ((hip.gv_extEmb_1V_NoHash, out_a, cnt_a),
                                 (hip.gv_extEmb_1V_NoHash_noalloc, out_b, cnt_b)):
                assert hip.gv_memcpy_h2d(ctx, cnt, zero2.ctypes.data_as(C.c_void_p), C.c_uint64(8)) == 0
                rc = fn(ctx, vlab, cnt, part, a3, None, auxd, out, ind, u32(n_int), u32(nrows), u32(3),
                        C.c_bool(True), C.c_bool(False), C.c_bool(False), C.c_bool(False), C.c_bool(False),
                        u32(cap))
                assert rc == 0, fn
                c = dev.get(cnt, 2)
                outs.append((c, G.sorted_rows(dev.get(out, int(c[0]) * 3).reshape(-1, 3))))
            (ca, ra), (cb, rb) = outs
            assert ca[0] == cb[0] and ca[1] == cb[1] == 0
            assert np.array_equal(ra, rb)
            assert len(ra) > 0
            # every row: the input row and a label-`lab` neighbour of column 1,
            # below column 0 and unlike both columns
            assert (ra[:, 2] < ra[:, 0]).all() and (ra[:, 2] != ra[:, 1]).all()
            assert (vl[ra[:, 2]] == lab).all()
        # past the fixed array's row limit: refused, nothing written
        big = 2_000_001
        rc = hip.gv_extEmb_1V_NoHash_noalloc(ctx, vlab, cnt_b, part, a3, None, auxd, out_b, ind, u32(n_int),
                                             u32(big), u32(3), C.c_bool(True), C.c_bool(False), C.c_bool(False),
                                             C.c_bool(False), C.c_bool(False), u32(cap))
        assert rc == -4
    finally:
        dev.free()
        hip.gv_ctx_destroy(ctx)
