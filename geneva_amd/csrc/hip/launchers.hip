// The reference's kernel launchers (include/miningPattern.h:1-31,
// include/reduction.h:1-4) with their parameter lists, on the reference's
// device layout: a partition block as the reference driver uploads it
// (array0 = interval starts and prefix counts, array1 = per-source offsets,
// array3 = neighbour lists grouped by neighbour label, each group of more than
// 130 entries led by its skip index), vLabel indexed by vertex id, auxArray as
// src/gpumine.cpp:40-61 lays it out, and row tables in which a row holding 0
// is invalid.
//
// One launch is, whatever the launcher:
//   k_ref_size  one thread per input row (per source index for the init
//               launchers): the source's list by interval lookup, the label
//               group of each new vertex by binary search over the labels of
//               the list, the skip index dropped by the reader rule of
//               src/miningPattern.cu:3289-3290, the restriction cut by binary
//               search; candidates = group length (one new vertex) or the
//               product of the two group lengths (two new vertices);
//   scan        inclusive uint64 scan of the candidate counts;
//   k_ref_emit  one thread per candidate of [c0, c1): its row by binary search
//               over the scan, the new vertex (or pair) by index, injectivity
//               against the same-label columns and between the pair, the
//               pair's restriction; survivors reserved with one atomic per
//               wave on totWriteRowNum[0] and written at that row index.
// Rows the reference would write with a 0 (a candidate equal to a mapped
// same-label vertex, src/miningPattern.cu:190-194) are not written at all:
// the next phase skips such rows anyway, and the last phase compacts them out
// (src/miningPattern.cu:172-187).  Capacity follows the reference protocol
// (src/gpumine_singleV.cpp:258-279): a launch never writes past row
// maxRowNum - 1; when candidates remain, totWriteRowNum[0] is raised to at
// least maxRowNum and totWriteRowNum[1] holds the difference to the rows
// actually written, and the same call with isContinue = true goes on from the
// first candidate not yet written (the resume point is kept in the context
// instead of records behind the table).  recordPos caches are written as the
// reference writes them (offset of the label group from the list start, group
// length without the skip index; 1, 0 for an empty group) and never read:
// every launch finds its groups itself, which gives the same rows.
#include "../../../include/geneva_launchers.h"
#include "common.h"

namespace {

constexpr uint32_t kNoCol = 0xffffffffu;
constexpr int kRefMaxIntervals = 256;
constexpr int kRScanItems = 8;
constexpr uint64_t kRScanTile = (uint64_t)kBlock * kRScanItems;

struct RefGraph {
    const uint32_t* vlab;   // vLabel[v]
    const uint32_t* a0;     // array0: starts[n], pre[n+1], one slot
    const uint32_t* a1;     // array1: offsets of each source's list in a3 (+1 terminal)
    const uint32_t* a3;     // neighbour lists
    uint32_t nint;
};

struct RefNew {             // one new column
    uint32_t label;
    uint32_t src;           // column of its source vertex in the input row
    uint32_t rcol;          // new < row[rcol]; kNoCol = none
    uint32_t ncoll;         // same-label columns it must differ from
    uint32_t coll[GV_MAXK];
    uint32_t* rec;          // recordPos of its label (nullable)
};

struct RefArgs {
    RefGraph g;
    const uint32_t* in;     // input rows (nullptr: the rows are the source indices)
    uint64_t nrows;
    uint32_t in_w;          // columns of an input row (1 for the init launchers)
    uint32_t init_label;    // init: label of the start vertex
    uint32_t nnew;          // 1 or 2
    RefNew nv[2];
    uint32_t e2_below_e1;   // two new vertices: second < first
    uint32_t distinct;      // two new vertices: first != second
    uint32_t* gb;           // [2][nrows] first real entry of each new vertex's group (index in a3)
    uint32_t* gl;           // [2][nrows] entries after the restriction cut
    uint64_t* cnt;          // candidates per row
    uint64_t* csum;         // their inclusive scan
    uint32_t* out;          // output rows of in_w + nnew columns
    uint32_t* counter;      // totWriteRowNum
    uint64_t c0, c1;        // candidate range of this emit
};

__device__ __forceinline__ void stage_array0(const RefGraph& g, uint32_t* s) {
    for (uint32_t t = threadIdx.x; t < 2 * g.nint + 2; t += blockDim.x) s[t] = g.a0[t];
    __syncthreads();
}

// array1 index of vertex v: v in [start_j, start_j + pre_{j+1} - pre_j)
// gives pre_j + v - start_j (src/miningPattern.cu:3254-3266); -1 if none
__device__ __forceinline__ int64_t ref_index(const uint32_t* a0s, uint32_t n, uint32_t v) {
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t st = a0s[j], len = a0s[n + j + 1] - a0s[n + j];
        if (v >= st && v - st < len) return (int64_t)a0s[n + j] + (v - st);
    }
    return -1;
}

// vertex of array1 index i: i in [pre_j, pre_{j+1}) gives start_j + i - pre_j
__device__ __forceinline__ uint32_t ref_vertex(const uint32_t* a0s, uint32_t n, uint64_t i) {
    for (uint32_t j = 0; j < n; ++j)
        if (i >= a0s[n + j] && i < a0s[n + j + 1]) return a0s[j] + (uint32_t)(i - a0s[n + j]);
    return 0;
}

// first k in [b, e) with vlab[a3[k]] >= l (labels ascend along a list)
__device__ __forceinline__ uint32_t label_lower(const RefGraph& g, uint32_t b, uint32_t e,
                                                uint32_t l) {
    uint32_t n = e - b;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (g.vlab[g.a3[b + h]] < l) {
            b += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return b;
}

// skip-index entries at the head of a label group of L entries
// (src/miningPattern.cu:3289-3290; tools/genDatabase.cpp:415-436)
__device__ __forceinline__ uint32_t skip_entries(uint32_t L) {
    if (L > 130 && L <= 4128) return (L + 128) / 129;
    if (L > 4128) return 32;
    return 0;
}

// Real entries of v's label-l group: [*beg, *beg + *len); the record of
// array1 index idx is written when rec is non-null.
__device__ __forceinline__ void ref_group(const RefGraph& g, uint64_t idx, uint32_t l, uint32_t* rec,
                                          uint32_t& beg, uint32_t& len) {
    const uint32_t lb = g.a1[idx], le = g.a1[idx + 1];
    const uint32_t gb = label_lower(g, lb, le, l);
    const uint32_t ge = label_lower(g, gb, le, l + 1);
    const uint32_t L = ge - gb;
    const uint32_t skip = skip_entries(L);
    beg = gb + skip;
    len = L - skip;
    if (rec) {
        rec[2 * idx] = L ? gb - lb : 1u;
        rec[2 * idx + 1] = len;
    }
}

__global__ __launch_bounds__(kBlock) void k_ref_size(RefArgs a) {
    __shared__ uint32_t a0s[2 * kRefMaxIntervals + 2];
    stage_array0(a.g, a0s);
    const uint32_t n = a.g.nint;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nrows; i += stride) {
        uint32_t row[GV_MAXK];
        bool ok = true;
        if (!a.in) {
            row[0] = ref_vertex(a0s, n, i);
            ok = row[0] != 0 && a.g.vlab[row[0]] == a.init_label;
        } else {
            for (uint32_t j = 0; j < a.in_w; ++j) {
                row[j] = a.in[i * a.in_w + j];
                ok = ok && row[j] != 0;
            }
        }
        uint64_t c = ok ? 1 : 0;
        for (uint32_t t = 0; t < a.nnew && ok; ++t) {
            const RefNew& nv = a.nv[t];
            const int64_t idx = a.in ? ref_index(a0s, n, row[nv.src]) : (int64_t)i;
            uint32_t beg = 0, len = 0;
            if (idx >= 0) ref_group(a.g, (uint64_t)idx, nv.label, nv.rec, beg, len);
            if (len && nv.rcol != kNoCol)
                len = lower_bound_u32(a.g.a3, beg, beg + len, row[nv.rcol]) - beg;
            a.gb[t * a.nrows + i] = beg;
            a.gl[t * a.nrows + i] = len;
            c *= len;
            ok = len != 0;
        }
        a.cnt[i] = ok ? c : 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_ref_emit(RefArgs a) {
    __shared__ uint32_t a0s[2 * kRefMaxIntervals + 2];
    if (!a.in) stage_array0(a.g, a0s);
    const unsigned lane = lane_id();
    const uint32_t ow = a.in_w + a.nnew;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = a.c0 + (uint64_t)blockIdx.x * blockDim.x; base < a.c1; base += stride) {
        const uint64_t c = base + threadIdx.x;
        bool keep = false;
        uint32_t row[GV_MAXK + 2];
        if (c < a.c1) {
            // row r: first with csum[r] > c
            uint64_t lo = 0, n = a.nrows;
            while (n > 0) {
                const uint64_t h = n >> 1;
                if (a.csum[lo + h] <= c) {
                    lo += h + 1;
                    n -= h + 1;
                } else {
                    n = h;
                }
            }
            const uint64_t r = lo;
            const uint64_t k = c - (r ? a.csum[r - 1] : 0);
            if (a.in) {
                for (uint32_t j = 0; j < a.in_w; ++j) row[j] = a.in[r * a.in_w + j];
            } else {
                row[0] = ref_vertex(a0s, a.g.nint, r);
            }
            uint32_t e1, e2 = 0;
            if (a.nnew == 1) {
                e1 = a.g.a3[a.gb[r] + k];
            } else {
                const uint64_t l2 = a.gl[a.nrows + r];
                e1 = a.g.a3[a.gb[r] + (uint32_t)(k / l2)];
                e2 = a.g.a3[a.gb[a.nrows + r] + (uint32_t)(k % l2)];
            }
            keep = true;
            for (uint32_t j = 0; j < a.nv[0].ncoll; ++j) keep = keep && row[a.nv[0].coll[j]] != e1;
            if (a.nnew == 2) {
                for (uint32_t j = 0; j < a.nv[1].ncoll; ++j) keep = keep && row[a.nv[1].coll[j]] != e2;
                if (a.distinct) keep = keep && e1 != e2;
                if (a.e2_below_e1) keep = keep && e2 < e1;
            }
            row[a.in_w] = e1;
            row[a.in_w + 1] = e2;
        }
        const uint64_t m = __ballot(keep);
        uint32_t pos = 0;
        if (m) {
            if (lane == 0) pos = atomicAdd(a.counter, (uint32_t)__popcll(m));
            pos = __shfl(pos, 0, 64);
        }
        if (keep) {
            uint32_t* dst = a.out + (uint64_t)(pos + lanes_below(m)) * ow;
            for (uint32_t j = 0; j < ow; ++j) dst[j] = row[j];
        }
    }
}

// ---- uint64 inclusive scan (three launches) --------------------------------

__device__ uint64_t block_exscan_u64(uint64_t v, uint64_t& total) {
    __shared__ uint64_t ws[kWavesPerBlock];
    const unsigned lane = lane_id(), w = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, 64);
        if (lane >= (unsigned)o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint64_t off = 0;
    total = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
        if (i < (int)w) off += ws[i];
        total += ws[i];
    }
    __syncthreads();
    return off + x - v;
}

__global__ __launch_bounds__(kBlock) void k_rscan_reduce(const uint64_t* in, uint64_t n, uint64_t* bsum) {
    const uint64_t b0 = (uint64_t)blockIdx.x * kRScanTile;
    uint64_t s = 0;
    for (int k = 0; k < kRScanItems; ++k) {
        const uint64_t i = b0 + (uint64_t)k * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint64_t tot;
    block_exscan_u64(s, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// exclusive scan of nb tile sums in place, one workgroup
__global__ __launch_bounds__(kBlock) void k_rscan_tiles(uint64_t* bsum, uint64_t nb) {
    uint64_t carry = 0;
    for (uint64_t b0 = 0; b0 < nb; b0 += kBlock) {
        const uint64_t i = b0 + threadIdx.x;
        const uint64_t v = i < nb ? bsum[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_exscan_u64(v, tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_rscan_final(const uint64_t* in, uint64_t n, const uint64_t* bsum,
                                                        uint64_t* out) {
    const uint64_t b0 = (uint64_t)blockIdx.x * kRScanTile + (uint64_t)threadIdx.x * kRScanItems;
    uint64_t v[kRScanItems];
    uint64_t s = 0;
    for (int k = 0; k < kRScanItems; ++k) {
        v[k] = b0 + k < n ? in[b0 + k] : 0;
        s += v[k];
    }
    uint64_t tot;
    uint64_t run = bsum[blockIdx.x] + block_exscan_u64(s, tot);
    for (int k = 0; k < kRScanItems; ++k) {
        run += v[k];
        if (b0 + k < n) out[b0 + k] = run;
    }
}

// ---- reductionPhase -----------------------------------------------------------

struct RedArgs {
    RefGraph g;
    uint32_t* rows;         // partialEmb, in place
    uint64_t nrows;
    uint32_t w;             // embLen
    uint32_t npairs;
    uint32_t spos[16], epos[16], label[16];
    uint64_t* flag;         // 1 = the row survives
    const uint64_t* csum;
    uint32_t* scratch;      // survivors, then copied back over rows
    uint32_t* counter;
    uint32_t last;
};

__global__ __launch_bounds__(kBlock) void k_ref_red_flag(RedArgs a) {
    __shared__ uint32_t a0s[2 * kRefMaxIntervals + 2];
    stage_array0(a.g, a0s);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nrows; i += stride) {
        const uint32_t* row = a.rows + i * a.w;
        bool ok = true;
        for (uint32_t j = 0; j < a.w; ++j) ok = ok && row[j] != 0;
        for (uint32_t p = 0; p < a.npairs && ok; ++p) {
            const uint32_t s = row[a.spos[p]], e = row[a.epos[p]];
            const int64_t idx = ref_index(a0s, a.g.nint, s);
            uint32_t beg = 0, len = 0;
            if (idx >= 0) ref_group(a.g, (uint64_t)idx, a.label[p], nullptr, beg, len);
            const uint32_t at = lower_bound_u32(a.g.a3, beg, beg + len, e);
            ok = at < beg + len && a.g.a3[at] == e;
        }
        a.flag[i] = ok ? 1 : 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_ref_red_scatter(RedArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nrows; i += stride) {
        if (!a.flag[i]) continue;
        const uint64_t pos = a.csum[i] - 1;
        for (uint32_t j = 0; j < a.w; ++j) a.scratch[pos * a.w + j] = a.rows[i * a.w + j];
    }
}

// survivors back over the input rows; totWriteRowNum[0] += survivors in the
// last phase (rows are reserved by atomicAdd there, src/reduction.cu:244-247)
// and = max(itself, survivors) otherwise (src/reduction.cu:250-260)
__global__ __launch_bounds__(kBlock) void k_ref_red_copy(RedArgs a) {
    const uint64_t surv = a.nrows ? a.csum[a.nrows - 1] : 0;
    const uint64_t words = surv * a.w;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
        a.rows[i] = a.scratch[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (a.last) a.counter[0] += (uint32_t)surv;
        else a.counter[0] = a.counter[0] > surv ? a.counter[0] : (uint32_t)surv;
    }
}

// a launch that stopped before its last candidate: totWriteRowNum[0] becomes
// M = max(maxRowNum, rows written) and totWriteRowNum[1] = M - rows written,
// so the caller's [0] - [1] are the rows written (src/gpumine_singleV.cpp:264)
__global__ void k_ref_overflow(uint32_t* counter, uint32_t max_rows) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t w = counter[0];
    const uint32_t m = w > max_rows ? w : max_rows;
    counter[0] = m;
    counter[1] = m - w;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

// Resume point of the last launch that stopped before its last candidate.
struct RefResume {
    bool active = false;
    int kind = -1;
    const void* in = nullptr;
    const void* out = nullptr;
    uint64_t nrows = 0;
    uint64_t total = 0;
    uint64_t next = 0;
};

struct RefState {
    uint64_t cap = 0;         // rows the workspace holds
    uint32_t* gb = nullptr;
    uint32_t* gl = nullptr;
    uint64_t* cnt = nullptr;
    uint64_t* csum = nullptr;
    uint64_t* bsum = nullptr;
    uint32_t* scratch = nullptr;
    uint64_t scratch_cap = 0; // words
    uint64_t* hval = nullptr; // pinned: total candidates
    uint32_t* hcnt = nullptr; // pinned: totWriteRowNum
    RefResume res;
};

void ref_state_free(gv_ctx* c) {
    RefState* r = c->ref;
    if (!r) return;
    void* ptrs[] = {r->gb, r->gl, r->cnt, r->csum, r->bsum, r->scratch};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (r->hval) (void)hipHostFree(r->hval);
    if (r->hcnt) (void)hipHostFree(r->hcnt);
    delete r;
    c->ref = nullptr;
}

namespace {

enum RefKind { kInit2V, kInit3V, kExt1V, kExt2V2S1L, kExt2V2S2L, kExt2V1S1L, kExt2V1S2L };

int ref_state(gv_ctx* c, RefState*& r) {
    if (!c->ref) {
        c->ref = new RefState();
        if (hipHostMalloc((void**)&c->ref->hval, 16, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&c->ref->hcnt, 16, hipHostMallocDefault) != hipSuccess) {
            ref_state_free(c);
            return GV_ERR_NOMEM;
        }
    }
    r = c->ref;
    return GV_OK;
}

int ref_reserve_rows(gv_ctx* c, RefState* r, uint64_t n) {
    if (n <= r->cap) return GV_OK;
    HIPCHK(hipStreamSynchronize(c->stream));   // the old buffers may be in use
    void* ptrs[] = {r->gb, r->gl, r->cnt, r->csum, r->bsum};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    r->gb = r->gl = nullptr;
    r->cnt = r->csum = r->bsum = nullptr;
    r->cap = 0;
    const uint64_t cap = std::max<uint64_t>(n, 4096);
    const uint64_t nb = (cap + kRScanTile - 1) / kRScanTile + 1;
    if (hipMalloc(&r->gb, 8 * cap) != hipSuccess || hipMalloc(&r->gl, 8 * cap) != hipSuccess ||
        hipMalloc(&r->cnt, 8 * cap) != hipSuccess || hipMalloc(&r->csum, 8 * cap) != hipSuccess ||
        hipMalloc(&r->bsum, 8 * nb) != hipSuccess)
        return GV_ERR_NOMEM;
    r->cap = cap;
    return GV_OK;
}

int ref_scan(gv_ctx* c, uint64_t* bsum, uint64_t* csum, const uint64_t* in, uint64_t n) {
    if (n == 0) return GV_OK;
    const uint64_t nb = (n + kRScanTile - 1) / kRScanTile;
    hipLaunchKernelGGL(k_rscan_reduce, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, in, n, bsum);
    hipLaunchKernelGGL(k_rscan_tiles, dim3(1), dim3(kBlock), 0, c->stream, bsum, nb);
    hipLaunchKernelGGL(k_rscan_final, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, in, n, bsum, csum);
    HIPCHK(hipGetLastError());
    return GV_OK;
}

bool ref_graph(RefGraph& g, uint32_t* vLabel, uint32_t* edgeLabelPartition, uint32_t* neighborsData,
               uint32_t intervalNum) {
    if (!vLabel || !edgeLabelPartition || !neighborsData || intervalNum < 1 ||
        intervalNum > kRefMaxIntervals)
        return false;
    g.vlab = vLabel;
    g.a0 = edgeLabelPartition;
    g.a1 = edgeLabelPartition + 2 * intervalNum + 2;
    g.a3 = neighborsData;
    g.nint = intervalNum;
    return true;
}

void new_col(RefNew& nv, uint32_t label, uint32_t src, uint32_t rcol, uint32_t* rec) {
    memset(&nv, 0, sizeof(nv));
    nv.label = label;
    nv.src = src;
    nv.rcol = rcol;
    nv.rec = rec;
}

// Same-label columns of the new vertices from auxArray[6..] (host copy):
// aux[6] of them for the first, aux[7] - aux[6] after those for the second
// (the first's when both carry one label).
int aux_colls(const uint32_t* aux, uint32_t in_w, bool same, RefNew& n1, RefNew* n2) {
    const uint32_t a6 = aux[6], a7 = aux[7];
    if (a6 > GV_MAXK || a6 > 56 || (n2 && (a7 < a6 || a7 > 56 || a7 - a6 > GV_MAXK))) return GV_ERR_ARG;
    n1.ncoll = a6;
    for (uint32_t j = 0; j < a6; ++j) {
        n1.coll[j] = aux[8 + j];
        if (n1.coll[j] >= in_w) return GV_ERR_ARG;
    }
    if (n2) {
        if (same) {
            n2->ncoll = n1.ncoll;
            memcpy(n2->coll, n1.coll, sizeof(n1.coll));
        } else {
            n2->ncoll = a7 - a6;
            for (uint32_t j = 0; j < n2->ncoll; ++j) {
                n2->coll[j] = aux[8 + a6 + j];
                if (n2->coll[j] >= in_w) return GV_ERR_ARG;
            }
        }
    }
    return GV_OK;
}

// Working arrays of one launch and the host words its sizes are read into.
struct RefWork {
    uint32_t* gb;
    uint32_t* gl;
    uint64_t* cnt;
    uint64_t* csum;
    uint64_t* bsum;
    uint64_t* hval;   // total candidates
    uint32_t* hcnt;   // totWriteRowNum[0]
    RefResume* res;
};

// One launch of an extension or init launcher: size and scan on a fresh
// call, then emit the candidates that fit below maxRowNum.  The workspace
// must hold a.nrows rows.
int ref_extend_on(gv_ctx* c, RefArgs& a, int kind, uint32_t* counter, uint32_t maxRowNum, bool isContinue,
                  const RefWork& w) {
    RefResume* r = w.res;
    a.gb = w.gb;
    a.gl = w.gl;
    a.cnt = w.cnt;
    a.csum = w.csum;
    if (!isContinue) {
        r->active = false;
        r->total = 0;
        if (a.nrows) {
            hipLaunchKernelGGL(k_ref_size, dim3(grid_for(a.nrows, kBlock, c->num_cu * 16)), dim3(kBlock), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
            int rc = ref_scan(c, w.bsum, w.csum, w.cnt, a.nrows);
            if (rc != GV_OK) return rc;
            HIPCHK(hipMemcpyAsync(w.hval, w.csum + a.nrows - 1, 8, hipMemcpyDeviceToHost, c->stream));
        } else {
            *w.hval = 0;
        }
        HIPCHK(hipMemcpyAsync(w.hcnt, counter, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        r->total = *w.hval;
        r->next = 0;
        r->kind = kind;
        r->in = a.in;
        r->out = a.out;
        r->nrows = a.nrows;
    } else {
        if (!r->active || r->kind != kind || r->in != a.in || r->out != a.out || r->nrows != a.nrows)
            return GV_ERR_ARG;
        HIPCHK(hipMemcpyAsync(w.hcnt, counter, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    const uint64_t written = *w.hcnt;
    const uint64_t room = (uint64_t)maxRowNum > written + 1 ? maxRowNum - 1 - written : 0;
    const uint64_t end = std::min<uint64_t>(r->total, r->next + room);
    if (end > r->next) {
        a.counter = counter;
        a.c0 = r->next;
        a.c1 = end;
        hipLaunchKernelGGL(k_ref_emit, dim3(grid_for(end - r->next, kBlock, c->num_cu * 16)), dim3(kBlock), 0,
                           c->stream, a);
        HIPCHK(hipGetLastError());
    }
    r->next = end;
    r->active = end < r->total;
    if (r->active) {
        hipLaunchKernelGGL(k_ref_overflow, dim3(1), dim3(64), 0, c->stream, counter, maxRowNum);
        HIPCHK(hipGetLastError());
        if (room == 0) return GV_ERR_CAPACITY;   // no row fits: the caller would loop forever
    }
    return GV_OK;
}

// The launcher entries: the context's workspace, grown to the input.
int ref_extend(gv_ctx* c, RefArgs& a, int kind, uint32_t* counter, uint32_t maxRowNum, bool isContinue) {
    if (!c || !counter || !a.out || (a.in_w + a.nnew) > GV_MAXK + 1) return GV_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    RefState* r = nullptr;
    int rc = ref_state(c, r);
    if (rc != GV_OK) return rc;
    if (!isContinue) {
        r->res.active = false;
        rc = ref_reserve_rows(c, r, a.nrows);
        if (rc != GV_OK) return rc;
    }
    ProfScope ps(c, kKindRef);
    const RefWork w{r->gb, r->gl, r->cnt, r->csum, r->bsum, r->hval, r->hcnt, &r->res};
    return ref_extend_on(c, a, kind, counter, maxRowNum, isContinue, w);
}

// ---- the extension launcher without allocations --------------------------
// The working arrays of gv_extEmb_1V_NoHash_noalloc live in one file-scope
// device array per GPU, `scratch`, of kNoAllocBytes: 32 bytes per input row
// (two group starts, two lengths, a count and its scan) and the scan's tile
// sums, so at most kNoAllocRows input rows per call.  Its address on a
// device comes from hipGetSymbolAddress at the first call on that device.
constexpr uint64_t kNoAllocRows = 2000000;
constexpr uint64_t kNoAllocTiles = kNoAllocRows / kRScanTile + 2;
constexpr uint64_t kNoAllocBytes = 32 * kNoAllocRows + 8 * kNoAllocTiles;
__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
    uint8_t* ptr = nullptr;     // this device's copy of scratch
    RefResume res;
};
NoAllocDevice g_noalloc[kNoAllocDevices];
std::mutex g_noalloc_mu;

int noalloc_device(int device, NoAllocDevice*& out) {
    if (device < 0 || device >= kNoAllocDevices) return GV_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice& d = g_noalloc[device];
    if (!d.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_last_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return GV_ERR_HIP;
        }
        d.ptr = static_cast<uint8_t*>(p);
    }
    out = &d;
    return GV_OK;
}

}  // namespace

extern "C" {

int gv_initEmb_2V_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                         uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                         uint32_t* recordPos_dev, uint32_t* newEmb_dev, uint32_t intervalNum,
                         uint32_t totSrcNum, uint32_t svidlabel, uint32_t evidlabel, uint32_t isRestrict,
                         uint32_t isRecord, bool isContinue, uint32_t maxRowNum) {
    RefArgs a;
    memset(&a, 0, sizeof(a));
    if (!ref_graph(a.g, vLabel_dev, edgeLabelPartition_dev, neighborsData_dev, intervalNum))
        return GV_ERR_ARG;
    a.in = nullptr;
    a.nrows = totSrcNum;
    a.in_w = 1;
    a.init_label = svidlabel;
    a.nnew = 1;
    new_col(a.nv[0], evidlabel, 0, isRestrict ? 0 : kNoCol, isRecord ? recordPos_dev : nullptr);
    a.out = newEmb_dev;
    return ref_extend(c, a, kInit2V, totWriteRowNum_dev, maxRowNum, isContinue);
}

int gv_initEmb_3V_1and2_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                               uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                               uint32_t* recordPos1_dev, uint32_t* recordPos2_dev, uint32_t* newEmb_dev,
                               uint32_t intervalNum, uint32_t totSrcNum, uint32_t svidlabel,
                               uint32_t evid1label, uint32_t evid2label, bool isExt1Restrict,
                               bool isExt2Restrict, bool isExt1and2Restrict, bool isExt1and2SameLabel,
                               bool isRecord1, bool isRecord2, bool isContinue, uint32_t maxRowNum) {
    RefArgs a;
    memset(&a, 0, sizeof(a));
    if (!ref_graph(a.g, vLabel_dev, edgeLabelPartition_dev, neighborsData_dev, intervalNum))
        return GV_ERR_ARG;
    if (isExt1and2SameLabel != (evid1label == evid2label)) return GV_ERR_ARG;
    a.in = nullptr;
    a.nrows = totSrcNum;
    a.in_w = 1;
    a.init_label = svidlabel;
    a.nnew = 2;
    new_col(a.nv[0], evid1label, 0, isExt1Restrict ? 0 : kNoCol, isRecord1 ? recordPos1_dev : nullptr);
    new_col(a.nv[1], evid2label, 0, isExt2Restrict ? 0 : kNoCol,
            isRecord2 && !isExt1and2SameLabel ? recordPos2_dev : nullptr);
    a.distinct = isExt1and2SameLabel;
    a.e2_below_e1 = isExt1and2SameLabel && isExt1and2Restrict;
    a.out = newEmb_dev;
    return ref_extend(c, a, kInit3V, totWriteRowNum_dev, maxRowNum, isContinue);
}

int gv_extEmb_1V_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                        uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev, uint32_t* recordPos_dev,
                        uint32_t* auxArray_dev, uint32_t* newEmb_dev, uint32_t* partialEmb_dev,
                        uint32_t intervalNum, uint32_t partialRowNum, uint32_t embLen, bool isExtRestrict,
                        bool isRecord, bool useRecord, bool isLastPhase, bool isContinue,
                        uint32_t maxRowNum) {
    (void)useRecord;
    (void)isLastPhase;   // survivors are compacted in every phase
    RefArgs a;
    memset(&a, 0, sizeof(a));
    if (!c || !ref_graph(a.g, vLabel_dev, edgeLabelPartition_dev, neighborsData_dev, intervalNum) ||
        !auxArray_dev || embLen < 2 || embLen > GV_MAXK + 1 || (partialRowNum && !partialEmb_dev))
        return GV_ERR_ARG;
    uint32_t aux[64];
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(aux, auxArray_dev, sizeof(aux), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t in_w = embLen - 1;
    if (aux[4] >= in_w || (isExtRestrict && aux[2] >= in_w)) return GV_ERR_ARG;
    a.in = partialEmb_dev;
    a.nrows = partialRowNum;
    a.in_w = in_w;
    a.nnew = 1;
    new_col(a.nv[0], aux[0], aux[4], isExtRestrict ? aux[2] : kNoCol, isRecord ? recordPos_dev : nullptr);
    int rc = aux_colls(aux, in_w, false, a.nv[0], nullptr);
    if (rc != GV_OK) return rc;
    a.out = newEmb_dev;
    return ref_extend(c, a, kExt1V, totWriteRowNum_dev, maxRowNum, isContinue);
}

// The four two-vertex extension launchers: the first new vertex from
// row[aux[4]]'s aux[0]-labelled group, the second from row[aux[5]]'s
// aux[1]-labelled group (aux[4] == aux[5] for one source).
static int ext2v(gv_ctx_t c, int kind, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                 uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev, uint32_t* rec1,
                 uint32_t* rec2, uint32_t* auxArray_dev, uint32_t* newEmb_dev, uint32_t* partialEmb_dev,
                 uint32_t intervalNum, uint32_t partialRowNum, uint32_t embLen, bool r1, bool r2, bool r12,
                 bool sameLabel, bool oneSrc, bool isContinue, uint32_t maxRowNum) {
    RefArgs a;
    memset(&a, 0, sizeof(a));
    if (!c || !ref_graph(a.g, vLabel_dev, edgeLabelPartition_dev, neighborsData_dev, intervalNum) ||
        !auxArray_dev || embLen < 3 || embLen > GV_MAXK + 1 || (partialRowNum && !partialEmb_dev))
        return GV_ERR_ARG;
    uint32_t aux[64];
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(aux, auxArray_dev, sizeof(aux), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t in_w = embLen - 2;
    if (aux[4] >= in_w || aux[5] >= in_w || (r1 && aux[2] >= in_w) || (r2 && aux[3] >= in_w))
        return GV_ERR_ARG;
    if (sameLabel != (aux[0] == aux[1]) || oneSrc != (aux[4] == aux[5])) return GV_ERR_ARG;
    a.in = partialEmb_dev;
    a.nrows = partialRowNum;
    a.in_w = in_w;
    a.nnew = 2;
    new_col(a.nv[0], aux[0], aux[4], r1 ? aux[2] : kNoCol, rec1);
    new_col(a.nv[1], aux[1], aux[5], r2 ? aux[3] : kNoCol, sameLabel ? nullptr : rec2);
    int rc = aux_colls(aux, in_w, sameLabel, a.nv[0], &a.nv[1]);
    if (rc != GV_OK) return rc;
    a.distinct = sameLabel;
    a.e2_below_e1 = sameLabel && r12;
    a.out = newEmb_dev;
    return ref_extend(c, a, kind, totWriteRowNum_dev, maxRowNum, isContinue);
}

int gv_extEmb_2V_2src_1and2_sameLabel_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                             uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                             uint32_t* recordPos_dev, uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                             uint32_t* partialEmb_dev, uint32_t intervalNum,
                                             uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                             bool isExt2Restrict, bool isExt1and2Restrict, bool isRecord,
                                             bool useRecord, bool isContinue, uint32_t maxRowNum) {
    (void)useRecord;
    return ext2v(c, kExt2V2S1L, vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
                 isRecord ? recordPos_dev : nullptr, nullptr, auxArray_dev, newEmb_dev, partialEmb_dev,
                 intervalNum, partialRowNum, embLen, isExt1Restrict, isExt2Restrict, isExt1and2Restrict, true,
                 false, isContinue, maxRowNum);
}

int gv_extEmb_2V_2src_1and2_notSameLabel_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                                uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                                uint32_t* recordPos1_dev, uint32_t* recordPos2_dev,
                                                uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                                uint32_t* partialEmb_dev, uint32_t intervalNum,
                                                uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                                bool isExt2Restrict, bool isRecord1, bool isRecord2,
                                                bool useRecord1, bool useRecord2, bool isContinue,
                                                uint32_t maxRowNum) {
    (void)useRecord1;
    (void)useRecord2;
    return ext2v(c, kExt2V2S2L, vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
                 isRecord1 ? recordPos1_dev : nullptr, isRecord2 ? recordPos2_dev : nullptr, auxArray_dev,
                 newEmb_dev, partialEmb_dev, intervalNum, partialRowNum, embLen, isExt1Restrict, isExt2Restrict,
                 false, false, false, isContinue, maxRowNum);
}

int gv_extEmb_2V_1src_1and2_sameLabel_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                             uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                             uint32_t* recordPos_dev, uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                             uint32_t* partialEmb_dev, uint32_t intervalNum,
                                             uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                             bool isExt2Restrict, bool isExt1and2Restrict, bool isRecord,
                                             bool useRecord, bool isContinue, uint32_t maxRowNum) {
    (void)useRecord;
    return ext2v(c, kExt2V1S1L, vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
                 isRecord ? recordPos_dev : nullptr, nullptr, auxArray_dev, newEmb_dev, partialEmb_dev,
                 intervalNum, partialRowNum, embLen, isExt1Restrict, isExt2Restrict, isExt1and2Restrict, true,
                 true, isContinue, maxRowNum);
}

int gv_extEmb_2V_1src_1and2_notSameLabel_NoHash(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                                uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                                uint32_t* recordPos1_dev, uint32_t* recordPos2_dev,
                                                uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                                uint32_t* partialEmb_dev, uint32_t intervalNum,
                                                uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                                bool isExt2Restrict, bool isRecord1, bool isRecord2,
                                                bool useRecord1, bool useRecord2, bool isContinue,
                                                uint32_t maxRowNum) {
    (void)useRecord1;
    (void)useRecord2;
    return ext2v(c, kExt2V1S2L, vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
                 isRecord1 ? recordPos1_dev : nullptr, isRecord2 ? recordPos2_dev : nullptr, auxArray_dev,
                 newEmb_dev, partialEmb_dev, intervalNum, partialRowNum, embLen, isExt1Restrict, isExt2Restrict,
                 false, false, true, isContinue, maxRowNum);
}

// auxArray of a reduction (src/gpumine.cpp:40-50, src/gpumine_singleV.cpp:599-633):
// [0] words of packed pairs, [1..6] the pairs, three per word at bits 10i
// (end column) and 10i + 5 (start column), [7..22] the end vertex's label per
// pair, [23..38] record offsets, [39] record flags.  A pair of two zero
// columns ends the list (a real pair joins two different columns).
int gv_reductionPhase(gv_ctx_t c, uint32_t* vLabel, uint32_t* totWriteRowNum, uint32_t* edgeLabelPartition,
                      uint32_t* neighborsData, uint32_t* baseRecordPos, uint32_t* auxArray,
                      uint32_t* partialEmb, uint32_t intervalNum, uint32_t partialRowNum, uint32_t embLen,
                      bool isLastPhase) {
    (void)baseRecordPos;   // record caches are not read (see the top of this file)
    RedArgs a;
    memset(&a, 0, sizeof(a));
    if (!c || !totWriteRowNum || !auxArray || embLen < 2 || embLen > 32 || (partialRowNum && !partialEmb) ||
        !ref_graph(a.g, vLabel, edgeLabelPartition, neighborsData, intervalNum))
        return GV_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    uint32_t aux[40];
    HIPCHK(hipMemcpyAsync(aux, auxArray, sizeof(aux), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (aux[0] < 1 || aux[0] > 6) return GV_ERR_ARG;
    for (uint32_t p = 0; p < 3 * aux[0] && p < 16; ++p) {
        const uint32_t w = aux[1 + p / 3] >> (10 * (p % 3));
        const uint32_t e = w & 31, s = (w >> 5) & 31;
        if (e == 0 && s == 0) break;
        if (e >= embLen || s >= embLen || e == s) return GV_ERR_ARG;
        a.epos[a.npairs] = e;
        a.spos[a.npairs] = s;
        a.label[a.npairs] = aux[7 + p];
        ++a.npairs;
    }
    if (a.npairs == 0) return GV_ERR_ARG;
    RefState* r = nullptr;
    int rc = ref_state(c, r);
    if (rc != GV_OK) return rc;
    r->res.active = false;
    rc = ref_reserve_rows(c, r, partialRowNum);
    if (rc != GV_OK) return rc;
    const uint64_t words = (uint64_t)partialRowNum * embLen;
    if (words > r->scratch_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (r->scratch) (void)hipFree(r->scratch);
        r->scratch = nullptr;
        r->scratch_cap = 0;
        if (hipMalloc(&r->scratch, 4 * words) != hipSuccess) return GV_ERR_NOMEM;
        r->scratch_cap = words;
    }
    a.rows = partialEmb;
    a.nrows = partialRowNum;
    a.w = embLen;
    a.flag = r->cnt;
    a.csum = r->csum;
    a.scratch = r->scratch;
    a.counter = totWriteRowNum;
    a.last = isLastPhase ? 1 : 0;
    ProfScope ps(c, kKindRef);
    if (partialRowNum) {
        hipLaunchKernelGGL(k_ref_red_flag, dim3(grid_for(partialRowNum, kBlock, c->num_cu * 16)), dim3(kBlock),
                           0, c->stream, a);
        HIPCHK(hipGetLastError());
        rc = ref_scan(c, r->bsum, r->csum, r->cnt, partialRowNum);
        if (rc != GV_OK) return rc;
        hipLaunchKernelGGL(k_ref_red_scatter, dim3(grid_for(partialRowNum, kBlock, c->num_cu * 16)),
                           dim3(kBlock), 0, c->stream, a);
    }
    hipLaunchKernelGGL(k_ref_red_copy, dim3(grid_for(words, kBlock, c->num_cu * 16)), dim3(kBlock), 0,
                       c->stream, a);
    HIPCHK(hipGetLastError());
    return GV_OK;
}

int gv_copyPartialEmb(gv_ctx_t c, uint32_t* src_dev, uint32_t* dst_dev, uint32_t numUints) {
    return gv_copy_partial_emb(c, src_dev, dst_dev, numUints);
}

int gv_extEmb_1V_NoHash_noalloc(gv_ctx_t c, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                uint32_t* recordPos_dev, uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                uint32_t* partialEmb_dev, uint32_t intervalNum, uint32_t partialRowNum,
                                uint32_t embLen, bool isExtRestrict, bool isRecord, bool useRecord,
                                bool isLastPhase, bool isContinue, uint32_t maxRowNum) {
    (void)useRecord;
    (void)isLastPhase;
    RefArgs a;
    memset(&a, 0, sizeof(a));
    if (!c || !totWriteRowNum_dev || !newEmb_dev ||
        !ref_graph(a.g, vLabel_dev, edgeLabelPartition_dev, neighborsData_dev, intervalNum) || !auxArray_dev ||
        embLen < 2 || embLen > GV_MAXK + 1 || (partialRowNum && !partialEmb_dev))
        return GV_ERR_ARG;
    if (partialRowNum > kNoAllocRows) {
        g_last_error = "gv_extEmb_1V_NoHash_noalloc: " + std::to_string(partialRowNum) +
                       " input rows, at most " + std::to_string(kNoAllocRows);
        return GV_ERR_CAPACITY;
    }
    HIPCHK(hipSetDevice(c->device));
    NoAllocDevice* d = nullptr;
    int rc = noalloc_device(c->device, d);
    if (rc != GV_OK) return rc;
    uint32_t aux[64];
    HIPCHK(hipMemcpyAsync(aux, auxArray_dev, sizeof(aux), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t in_w = embLen - 1;
    if (aux[4] >= in_w || (isExtRestrict && aux[2] >= in_w)) return GV_ERR_ARG;
    a.in = partialEmb_dev;
    a.nrows = partialRowNum;
    a.in_w = in_w;
    a.nnew = 1;
    new_col(a.nv[0], aux[0], aux[4], isExtRestrict ? aux[2] : kNoCol, isRecord ? recordPos_dev : nullptr);
    rc = aux_colls(aux, in_w, false, a.nv[0], nullptr);
    if (rc != GV_OK) return rc;
    a.out = newEmb_dev;
    // the working arrays in this device's scratch (fixed layout)
    uint8_t* p = d->ptr;
    uint64_t hval = 0;
    uint32_t hcnt = 0;
    const RefWork w{reinterpret_cast<uint32_t*>(p), reinterpret_cast<uint32_t*>(p + 8 * kNoAllocRows),
                    reinterpret_cast<uint64_t*>(p + 16 * kNoAllocRows),
                    reinterpret_cast<uint64_t*>(p + 24 * kNoAllocRows),
                    reinterpret_cast<uint64_t*>(p + 32 * kNoAllocRows), &hval, &hcnt, &d->res};
    return ref_extend_on(c, a, kExt1V, totWriteRowNum_dev, maxRowNum, isContinue, w);
}

int gv_launchers_reserve(gv_ctx_t c, uint64_t rows, uint64_t reduction_words) {
    if (!c) return GV_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    RefState* r = nullptr;
    int rc = ref_state(c, r);
    if (rc != GV_OK) return rc;
    rc = ref_reserve_rows(c, r, rows);
    if (rc != GV_OK) return rc;
    if (reduction_words > r->scratch_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (r->scratch) (void)hipFree(r->scratch);
        r->scratch = nullptr;
        r->scratch_cap = 0;
        if (hipMalloc(&r->scratch, 4 * reduction_words) != hipSuccess) return GV_ERR_NOMEM;
        r->scratch_cap = reduction_words;
    }
    return GV_OK;
}

}  // extern "C"
