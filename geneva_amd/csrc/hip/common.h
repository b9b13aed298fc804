// Internal definitions shared by the translation units of libgeneva_hip.so:
// device helpers, the context / graph / frontier objects and launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/geneva_hip.h"

namespace gvi {
extern thread_local std::string g_last_error;
}
using gvi::g_last_error;

namespace {

#define HIPCHK(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            g_last_error = std::string(#call) + ": " + hipGetErrorString(e_);     \
            return GV_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

constexpr int kBlock = 256;             // 4 wave64 per workgroup
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kIPT = 4;                 // candidates per lane per tile
constexpr uint64_t kTile = 64 * kIPT;   // candidates per wave tile
constexpr int kScanItems = 16;          // per thread in the scan
constexpr int kScanTile = kBlock * kScanItems;

struct DevGraph {
    const uint32_t* loff;   // [npart][nv1][nl1]
    const uint32_t* nbr;
    const uint32_t* gof;    // owner graph per vertex (nullable)
    uint64_t part_stride;   // nv1 * nl1
    uint32_t nl1;           // dense labels + 1
};

__device__ __forceinline__ void seg(const DevGraph& g, uint32_t part, uint32_t v, uint32_t l,
                                    uint32_t& b, uint32_t& e) {
    const uint64_t base = (uint64_t)part * g.part_stride + (uint64_t)v * g.nl1 + l;
    b = g.loff[base];
    e = g.loff[base + 1];
}

// first index in [b, e) with nbr[i] >= x
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t b, uint32_t e,
                                                    uint32_t x) {
    uint32_t n = e - b;
    while (n > 0) {
        uint32_t h = n >> 1;
        if (a[b + h] < x) {
            b += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return b;
}

__device__ __forceinline__ bool contains_u32(const uint32_t* a, uint32_t b, uint32_t e,
                                             uint32_t x) {
    uint32_t p = lower_bound_u32(a, b, e, x);
    return p < e && a[p] == x;
}

// first r in [lo, hi) with csum[r] > c
__device__ __forceinline__ uint64_t find_row(const uint64_t* csum, uint64_t lo, uint64_t hi,
                                             uint64_t c) {
    uint64_t n = hi - lo;
    while (n > 0) {
        uint64_t h = n >> 1;
        if (csum[lo + h] <= c) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

// entries a binary search over n sorted entries reads: floor(log2 n) + 1
__device__ __forceinline__ uint32_t probes(uint32_t n) { return n ? 32u - __clz(n) : 0u; }

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

struct gv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    // accumulators: [0] count, [1] bytes, [2] rows written, [3] overflow
    unsigned long long* acc = nullptr;
    uint32_t* gbits = nullptr;
    uint32_t gbits_words = 0;
    // scan scratch
    uint64_t* tiles = nullptr;
    uint64_t tiles_cap = 0;
    int num_cu = 256;
    // profiling: HIP events around every launch, per kernel kind
    bool prof = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    unsigned long long* preq = nullptr;   // requested bytes per kind (device)
    uint64_t scan_req = 0;                // scan bytes (computed on the host)
};

struct gv_graph {
    gv_ctx* ctx = nullptr;
    uint64_t nv1 = 0;
    uint32_t nlabels = 0;
    uint32_t nparts = 0;
    uint64_t nnbr = 0;
    uint64_t dev_bytes = 0;
    std::vector<uint32_t> label_values;       // dense -> original
    std::vector<uint32_t> part_edge_label;    // partition -> edge label
    std::vector<uint64_t> lvoff;              // host copy of start-table offsets
    std::vector<uint64_t> pair_edges;         // [part][NL][NL] directed edges by label pair
    uint32_t* d_loff = nullptr;
    uint32_t* d_nbr = nullptr;
    uint32_t* d_lverts = nullptr;
    uint32_t* d_gof = nullptr;
    uint32_t ngraphs = 0;
};

struct gv_frontier {
    gv_ctx* ctx = nullptr;
    uint64_t cap = 0;
    uint32_t* anc_beg = nullptr;
    uint32_t* cnt = nullptr;
    uint8_t* anc_i = nullptr;
    uint64_t* csum = nullptr;
    uint64_t nrows = 0;
    uint64_t total = 0;   // candidates of the last sizing
};


namespace {

DevGraph dev_view(const gv_graph* g) {
    DevGraph d;
    d.loff = g->d_loff;
    d.nbr = g->d_nbr;
    d.gof = g->d_gof;
    d.nl1 = g->nlabels + 1;
    d.part_stride = g->nv1 * (uint64_t)(g->nlabels + 1);
    return d;
}

enum { kKindSize = 0, kKindScan = 1, kKindExt = 2, kKindCount = 3, kKindBatch = 4, kKinds = 5 };

hipEvent_t pool_get(gv_ctx* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Bracket launches with events when profiling is on.
struct ProfScope {
    gv_ctx* c;
    int kind;
    hipEvent_t a = nullptr;
    ProfScope(gv_ctx* c_, int k) : c(c_), kind(k) {
        if (!c->prof) return;
        a = pool_get(c);
        if (a) (void)hipEventRecord(a, c->stream);
    }
    ~ProfScope() {
        if (!c->prof || !a) return;
        hipEvent_t b = pool_get(c);
        if (!b) return;
        (void)hipEventRecord(b, c->stream);
        c->recs.push_back({kind, a, b});
    }
};

unsigned long long* req_ptr(gv_ctx* c, int kind) {
    return (c->prof && c->preq) ? c->preq + kind : nullptr;
}

int grid_for(uint64_t items, uint64_t per_block, int cap_blocks) {
    uint64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > (uint64_t)cap_blocks) b = cap_blocks;
    return (int)b;
}

bool valid_step(const gv_step_t* s, uint32_t row_len, const gv_graph* g) {
    if (!s || s->nback < 1 || s->nback >= GV_MAXK || s->ncoll >= GV_MAXK) return false;
    if (s->label >= g->nlabels) return false;
    if (row_len + 1 > GV_MAXK) return false;
    for (int i = 0; i < s->nback; ++i)
        if (s->back_col[i] >= row_len || s->back_part[i] >= g->nparts) return false;
    for (int i = 0; i < s->ncoll; ++i)
        if (s->coll_col[i] >= row_len) return false;
    if (s->restrict_col != GV_NONE && s->restrict_col >= row_len) return false;
    return true;
}

// the tail step reads columns 0..row_len (row_len = the fused survivor)
bool valid_tail(const gv_step_t* s, uint32_t row_len, const gv_graph* g) {
    if (!s || s->nback != 1 || s->ncoll >= GV_MAXK) return false;
    if (s->label >= g->nlabels || row_len + 2 > GV_MAXK) return false;
    if (s->back_col[0] > row_len || s->back_part[0] >= g->nparts) return false;
    for (int i = 0; i < s->ncoll; ++i)
        if (s->coll_col[i] > row_len) return false;
    if (s->restrict_col != GV_NONE && s->restrict_col > row_len) return false;
    return true;
}

}  // namespace
