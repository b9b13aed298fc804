// Internal definitions shared by the translation units of libgeneva_hip.so:
// device helpers, the context / graph / frontier objects and launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
// candidates per lane of a whole query's extension walk over a step with one
// back list: with no membership test a candidate costs little, and smaller
// tiles spread a launch of a few million candidates over more waves
constexpr int kSingleIPT = 2;
// candidates per lane of a whole query's counting walk (measured: 2 beats 4
// and 8 on the 5-cycle and matches 4 elsewhere)
constexpr int kCountIPT = 2;
constexpr int kScanItems = 16;          // per thread in the scan
constexpr int kSizeRows = 2;            // rows per thread in a sizing pass
constexpr uint32_t kWin = 256;          // LDS window of a back-neighbour list per wave
constexpr uint64_t kSizeTile = (uint64_t)kBlock * kSizeRows;   // rows per sizing tile
constexpr int kScanTile = kBlock * kScanItems;
// The scan reads its LDS tile thread-major (thread t takes words 16t..16t+15),
// so one pad word after every 16 puts the 64 lanes of a wave on 64 different
// banks (17t + i is distinct mod 64 and mod 32).
__host__ __device__ constexpr int lds_pad16(int e) { return e + (e >> 4); }
constexpr int kScanTilePadded = kScanTile + kScanTile / 16;

struct DevGraph {
    const uint32_t* loff;   // [npart][nv1][nl1]
    const uint32_t* nbr;
    // start planes: sp[part][l][pos] = loff[part][lverts[pos]][l], so a start
    // table's rows (consecutive positions of lverts) read their label
    // offsets as consecutive words of two planes
    const uint32_t* sp;
    uint64_t sp_n;          // entries of lverts (plane length)
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

// The same search done by a whole wave (all lanes call it with the same
// arguments; the result is uniform): each round probes 64 evenly spaced rows
// with one load per lane and keeps the interval between the last false and
// the first true probe, so n rows take ceil(log64 n) dependent loads instead
// of log2 n.  Requires csum[hi - 1] > c.
__device__ __forceinline__ uint64_t wave_find_row(const uint64_t* csum, uint64_t lo, uint64_t hi,
                                                  uint64_t c) {
    const unsigned lane = __lane_id();
    while (hi - lo > 1) {
        const uint64_t n = hi - lo;
        const uint64_t step = (n + 63) >> 6;
        const uint64_t x = lo + (uint64_t)lane * step;
        const bool p = x < hi && csum[x] > c;
        const uint64_t m = __ballot(p);
        if (m == 0) {
            // every probe below hi is false: the answer lies past the last one
            const uint64_t last = (n - 1) / step < 63 ? (n - 1) / step : 63;
            lo = lo + last * step + 1;
            continue;
        }
        const unsigned j = __builtin_ctzll(m);
        if (j == 0) return lo;
        hi = lo + (uint64_t)j * step + 1;   // answer <= probe j
        lo = lo + (uint64_t)(j - 1) * step + 1;   // answer > probe j - 1
    }
    return lo;
}

// First index i in [lo, hi) of the sorted a with a[i] >= x (hi if none), found
// by the whole wave (uniform arguments and result): each round probes 64
// evenly spaced entries with one load per lane.
__device__ __forceinline__ uint32_t wave_lower_bound(const uint32_t* a, uint32_t lo, uint32_t hi,
                                                     uint32_t x) {
    const unsigned lane = __lane_id();
    while (lo < hi) {   // the answer lies in [lo, hi]
        const uint32_t n = hi - lo;
        const uint32_t step = (n + 63) >> 6;
        const uint32_t p = lo + lane * step;
        const uint64_t m = __ballot(lane * step < n && a[p] >= x);
        if (m == 0) {
            const uint32_t last = (n - 1) / step < 63 ? (n - 1) / step : 63;
            lo = lo + last * step + 1;
            continue;
        }
        const unsigned j = __builtin_ctzll(m);
        if (j == 0) return lo;
        hi = lo + j * step;
        lo = lo + (j - 1) * step + 1;
    }
    return lo;
}


// Multi-row tiles of at most kRowStage rows stage kRowWords words per row in
// the window space of the walk (2 * kWin words).
constexpr uint32_t kRowWords = 9;
constexpr uint32_t kRowStage = 2 * kWin / kRowWords;

// First index in [lo, n) of the sorted w (LDS) with w[i] >= x (n if none).
__device__ __forceinline__ uint32_t lds_lower_bound(const uint32_t* w, uint32_t lo, uint32_t n,
                                                    uint32_t x) {
    uint32_t len = n - lo;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (w[lo + half] < x) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// ok[it] &= (v[it] is in the sorted window w[0, n)), for a lane's IPT candidates
// in ascending order (those with it < nv): the first is found by a binary
// search, each later one by advancing a cursor from the previous position (a
// few steps, then a search of the rest), a merge of the lane's run of
// candidates with the window.
template <int IPT>
__device__ __forceinline__ void window_member(const uint32_t* w, uint32_t n, const uint32_t (&v)[IPT],
                                              uint32_t nv, bool (&ok)[IPT]) {
    uint32_t p = 0;
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
        if ((uint32_t)it >= nv) break;
        const uint32_t x = v[it];
        if (it == 0) {
            p = lds_lower_bound(w, 0, n, x);
        } else {
            for (int s = 0; s < 4 && p < n && w[p] < x; ++s) ++p;
            if (p < n && w[p] < x) p = lds_lower_bound(w, p, n, x);
        }
        ok[it] = ok[it] && p < n && w[p] == x;
    }
}

// Count of the last vertex when it has a single back edge (step `t`, rows of
// width L extended by candidate v in column L): the label segment of its
// neighbour, cut below the restriction bound, minus mapped vertices inside.
// The part that does not involve v is cached per row: for most last steps
// (the neighbour and the bound are earlier columns) every candidate of a row
// shares it, and only v's own collision is checked per candidate.
// Per-lane cache of the fused last step's row-dependent part.  The last query
// vertex u has one back edge, to column back_col[0] of the row (or to the
// candidate v of the step being walked, column L), an optional restriction
// f(u) < f(row[restrict_col]) and same-label columns it must differ from.
struct TailCache {
    uint64_t r = ~0ull;
    uint32_t b = 0, e = 0, cut = 0, bound = 0;
    uint64_t cnt = 0;
    // bound = v: same-label row values that lie in the segment (else ~0u),
    // so that the count below v is the rank of v minus those below it
    uint32_t x0 = ~0u, x1 = ~0u, x2 = ~0u;
};

// entries a binary search over n sorted entries reads: floor(log2 n) + 1
__device__ __forceinline__ uint32_t probes(uint32_t n) { return n ? 32u - __clz(n) : 0u; }

// Count of the last vertex's candidates for one survivor v (at row r) of the
// second-to-last step.  same_seg: the last vertex's segment is the one v was
// drawn from (same anchor column, label and partition), and v sits at index
// idx of it, so the rank of v needs no search.
__device__ __forceinline__ uint64_t tail_count(const DevGraph& g, const gv_step_t& t, uint32_t L,
                                               const uint32_t* row, uint32_t v, uint64_t r,
                                               bool same_seg, uint32_t idx, TailCache& tc,
                                               uint64_t& rq) {
    const bool segdep = t.back_col[0] == L;
    const bool bnddep = t.restrict_col == L;
    uint32_t nx = 0;   // same-label columns other than L
    for (uint32_t j = 0; j < t.ncoll; ++j) nx += t.coll_col[j] != L;
    if (segdep || (bnddep && nx > 3)) {
        // everything depends on v: segment of v, cut, collisions, per candidate
        const uint32_t w = segdep ? v : row[t.back_col[0]];
        uint32_t b, e;
        seg(g, t.back_part[0], w, t.label, b, e);
        const uint32_t bound =
            t.restrict_col == GV_NONE ? 0xffffffffu : (bnddep ? v : row[t.restrict_col]);
        const uint32_t cut = bound == 0xffffffffu ? e : lower_bound_u32(g.nbr, b, e, bound);
        rq += 8ull + (bound != 0xffffffffu ? 4ull * probes(e - b) : 0ull);
        uint64_t n = cut - b;
        for (uint32_t j = 0; j < t.ncoll; ++j) {
            const uint32_t x = t.coll_col[j] == L ? v : row[t.coll_col[j]];
            if (x < bound) {
                rq += 4ull * probes(cut - b);
                if (contains_u32(g.nbr, b, cut, x)) --n;
            }
        }
        tc.b = b;
        tc.e = e;
        tc.r = ~0ull;
        return n;
    }
    if (r != tc.r) {
        seg(g, t.back_part[0], row[t.back_col[0]], t.label, tc.b, tc.e);
        rq += 8ull;
        if (!bnddep) {
            tc.bound = t.restrict_col == GV_NONE ? 0xffffffffu : row[t.restrict_col];
            // ids are below 0xffffffff, so an absent bound cuts nothing
            tc.cut = tc.bound == 0xffffffffu ? tc.e : lower_bound_u32(g.nbr, tc.b, tc.e, tc.bound);
            rq += tc.bound != 0xffffffffu ? 4ull * probes(tc.e - tc.b) : 0ull;
            tc.cnt = tc.cut - tc.b;
            for (uint32_t j = 0; j < t.ncoll; ++j) {
                if (t.coll_col[j] == L) continue;
                const uint32_t x = row[t.coll_col[j]];
                if (x < tc.bound) {
                    rq += 4ull * probes(tc.cut - tc.b);
                    if (contains_u32(g.nbr, tc.b, tc.cut, x)) --tc.cnt;
                }
            }
        } else {
            uint32_t k = 0;
            tc.x0 = tc.x1 = tc.x2 = ~0u;
            for (uint32_t j = 0; j < t.ncoll; ++j) {
                if (t.coll_col[j] == L) continue;
                const uint32_t x = row[t.coll_col[j]];
                rq += 4ull * probes(tc.e - tc.b);
                const uint32_t y = contains_u32(g.nbr, tc.b, tc.e, x) ? x : ~0u;
                if (k == 0) tc.x0 = y;
                else if (k == 1) tc.x1 = y;
                else tc.x2 = y;
                ++k;
            }
        }
        tc.r = r;
    }
    if (!bnddep) {
        uint64_t n = tc.cnt;
        // v itself is excluded where u shares its label (v is in the segment
        // when it was drawn from it)
        for (uint32_t j = 0; j < t.ncoll; ++j) {
            if (t.coll_col[j] != L || v >= tc.bound) continue;
            if (!same_seg) rq += 4ull * probes(tc.cut - tc.b);
            if (same_seg || contains_u32(g.nbr, tc.b, tc.cut, v)) --n;
        }
        return n;
    }
    // bound = v: candidates of u are the segment entries below v (v itself is
    // never below v), minus the same-label row values among them
    uint32_t cut;
    if (same_seg) {
        cut = tc.b + idx;
    } else {
        cut = lower_bound_u32(g.nbr, tc.b, tc.e, v);
        rq += 4ull * probes(tc.e - tc.b);
    }
    return (uint64_t)(cut - tc.b) - (tc.x0 < v) - (tc.x1 < v) - (tc.x2 < v);
}


// ---- counting a last step per row ----------------------------------------
//
// When the last materialised step has a single back edge, a row's candidates
// are one label segment cut at the restriction, and its survivors are that
// range minus the row's same-label values found in it (s of them).  The row's
// count is then, by mode:
//   0  the step is the last:                                    s
//   1  the last vertex depends on the row only (its back edge,
//      restriction and same-label columns all lie in the row):  s * T(row)
//   2  the last vertex is drawn from the survivor's own segment,
//      has the same label and must lie below the survivor:      s (s - 1) / 2
// -1: not applicable (the candidates are walked).
__host__ __device__ inline int count_rows_mode(const gv_step_t* st, const gv_step_t* tail,
                                               uint32_t L) {
    if (!st || st->nback != 1) return -1;
    if (!tail) return 0;
    if (tail->nback != 1) return -1;
    bool collL = false;
    for (uint32_t j = 0; j < tail->ncoll; ++j) collL |= tail->coll_col[j] == L;
    if (tail->back_col[0] != L && tail->restrict_col != L && !collL) return 1;
    // same segment and label: the last vertex's same-label columns are then
    // exactly the survivor's own and the survivor
    if (tail->back_col[0] == st->back_col[0] && tail->back_part[0] == st->back_part[0] &&
        tail->label == st->label && tail->restrict_col == L && tail->ncoll == st->ncoll + 1) {
        for (uint32_t j = 0; j < st->ncoll; ++j)
            if (tail->coll_col[j] != st->coll_col[j]) return -1;
        return tail->coll_col[st->ncoll] == L ? 2 : -1;
    }
    return -1;
}

// Count of one row whose step segment is nbr[b, e) (the single back edge's
// label segment).  Adds the fused last vertex's algorithmic bytes to `by`
// (the caller adds the row's own) and the requested bytes to `rq`.
__device__ __forceinline__ uint64_t row_count(const DevGraph& g, const gv_step_t& st,
                                              const gv_step_t& tail, int mode, uint32_t L,
                                              const uint32_t* row, uint32_t b, uint32_t e,
                                              uint64_t& by, uint64_t& rq) {
    const uint32_t W = L + 1;
    const uint32_t bound = st.restrict_col != GV_NONE ? row[st.restrict_col] : 0xffffffffu;
    // ids are below 0xffffffff, so an absent bound cuts nothing
    const uint32_t cut = bound == 0xffffffffu ? e : lower_bound_u32(g.nbr, b, e, bound);
    rq += bound != 0xffffffffu ? 4ull * probes(e - b) : 0ull;
    uint64_t s = cut - b;
    for (uint32_t j = 0; j < st.ncoll; ++j) {
        const uint32_t x = row[st.coll_col[j]];
        if (x >= bound) continue;
        rq += 4ull * probes(cut - b);
        if (contains_u32(g.nbr, b, cut, x)) --s;
    }
    if (mode == 1 && s) {
        uint32_t tb, te;
        seg(g, tail.back_part[0], row[tail.back_col[0]], tail.label, tb, te);
        const uint32_t tbound = tail.restrict_col != GV_NONE ? row[tail.restrict_col] : 0xffffffffu;
        const uint32_t tcut = tbound == 0xffffffffu ? te : lower_bound_u32(g.nbr, tb, te, tbound);
        rq += 8ull + (tbound != 0xffffffffu ? 4ull * probes(te - tb) : 0ull);
        uint64_t t = tcut - tb;
        for (uint32_t j = 0; j < tail.ncoll; ++j) {
            const uint32_t x = row[tail.coll_col[j]];
            if (x >= tbound) continue;
            rq += 4ull * probes(tcut - tb);
            if (contains_u32(g.nbr, tb, tcut, x)) --t;
        }
        by += s * (8ull * W + 8ull + 4ull * (te - tb));
        return s * t;
    }
    if (mode == 2) {
        by += s * (8ull * W + 8ull + 4ull * (e - b));
        return s ? s * (s - 1) / 2 : 0;
    }
    return s;
}

// x / d for a d >= 1 that divides x: the power of two 2^k of d by a shift, its
// odd part o by multiplying with the inverse of o modulo 2^64 (exact division
// by multiplication, Granlund & Montgomery, PLDI 1994, section 9).  The
// inverse comes from Newton's iteration inv <- inv (2 - o inv), which doubles
// the number of correct low bits, starting from 3 (o o = 1 mod 8 for odd o).
// Exact when x / d < 2^(64 - k).
__host__ __device__ __forceinline__ uint64_t div_exact(uint64_t x, uint32_t d) {
    const uint32_t k = __builtin_ctz(d);
    const uint64_t o = d >> k;
    uint64_t inv = o;
#pragma unroll
    for (int i = 0; i < 5; ++i) inv *= 2 - o * inv;
    return (x >> k) * inv;
}

// One core row of a leaf-group query (gv_leaves_t): the row's embeddings,
// prod_g C(s_g, m_g); adds to `by` the algorithmic bytes of every row the
// enumeration would hold from the core onwards (with j_g leaves of group g
// placed there are prod_g C(s_g, j_g) rows), to `rq` the requested bytes.
__device__ __forceinline__ uint64_t leaf_row_count(const DevGraph& g, const gv_leaves_t& d,
                                                   const uint32_t* row, uint64_t& by,
                                                   uint64_t& rq) {
    constexpr int G = GV_MAX_LEAF_GROUPS;
    const uint32_t L0 = d.width, ng = d.ngroups, nleaf = d.k - d.width;
    uint32_t sv[G], len[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        sv[q] = 0;
        len[q] = 0;
        if (q >= (int)ng) continue;
        const gv_leaf_group_t& gr = d.group[q];
        uint32_t b, e;
        seg(g, gr.part, row[gr.anchor_col], gr.label, b, e);
        const uint32_t bound = gr.restrict_col != GV_NONE ? row[gr.restrict_col] : 0xffffffffu;
        const uint32_t cut = bound == 0xffffffffu ? e : lower_bound_u32(g.nbr, b, e, bound);
        rq += 8ull + (bound != 0xffffffffu ? 4ull * probes(e - b) : 0ull);
        uint32_t s = cut - b;
        for (uint32_t j = 0; j < gr.ncoll; ++j) {
            const uint32_t x = row[gr.coll_col[j]];
            if (x >= bound) continue;
            rq += 4ull * probes(cut - b);
            if (contains_u32(g.nbr, b, cut, x)) --s;
        }
        sv[q] = s;
        len[q] = e - b;
    }
    // c[q] = C(s_q, t_q) with t_q leaves of group q placed, updated one leaf at
    // a time as C(s, t + 1) = C(s, t) (s - t) / (t + 1).  The product before
    // the division is C(s, t + 1) (t + 1) <= s^(t + 1) < 2^63 (the groups are
    // checked against the largest degree), so it does not wrap and the quotient
    // is below 2^(63 - k) for t + 1 = 2^k o.
    uint64_t c[G];
    uint32_t t[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        c[q] = 1;
        t[q] = 0;
    }
    uint64_t rows_cur = 1;
    for (uint32_t j = 0; j < nleaf; ++j) {
        const uint32_t L = L0 + j;
        const uint32_t gj = d.group_of[j];
        uint64_t next = 1;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            if (q >= (int)ng) continue;
            if (q == (int)gj) {
                by += rows_cur * (4ull * L + 8ull + 4ull * len[q]);
                c[q] = sv[q] > t[q] ? div_exact(c[q] * (sv[q] - t[q]), t[q] + 1) : 0ull;
                t[q] += 1;
            }
            next *= c[q];
        }
        if (L + 1 < d.k) by += 4ull * (L + 1) * next;
        rows_cur = next;
    }
    return rows_cur;
}

// Sizing of R rows per thread (rows ri[k] of a table of width L, those with
// ok[k]), their memory accesses interleaved so that R independent loads are in
// flight at each level (row values, label-segment offsets, restriction
// probes).  Per row: the shortest back-neighbour segment (anchor index ai,
// start beg, length len) and its length below the restriction bound (cnt).
// Adds the algorithmic bytes of the row's step (row read and every required
// list) to `by` and the requested bytes to `rq`.
// start_pos >= 0: the table is the start table at lverts position start_pos
// (width 1), whose row r reads its offsets from the start planes.
template <int R>
__device__ __forceinline__ void size_rows(const DevGraph& g, const gv_step_t& st,
                                          const uint32_t* rows, uint32_t L, const uint64_t (&ri)[R],
                                          const bool (&ok)[R], uint32_t (&beg)[R],
                                          uint32_t (&len)[R], uint32_t (&cnt)[R],
                                          uint32_t (&ai)[R], uint64_t& by, uint64_t& rq,
                                          int64_t start_pos) {
    const bool planes = start_pos >= 0 && L == 1 && g.sp;
    uint32_t bound[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        beg[k] = 0;
        len[k] = 0xffffffffu;
        ai[k] = 0;
        bound[k] = 0xffffffffu;
    }
    if (st.restrict_col != GV_NONE) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (ok[k]) bound[k] = rows[ri[k] * L + st.restrict_col];
    }
    for (uint32_t i = 0; i < st.nback; ++i) {
        uint32_t b[R], e[R];
        if (planes) {
            const uint32_t* p0 =
                g.sp + ((uint64_t)st.back_part[i] * g.nl1 + st.label) * g.sp_n + (uint64_t)start_pos;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                b[k] = ok[k] ? p0[ri[k]] : 0u;
                e[k] = ok[k] ? p0[g.sp_n + ri[k]] : 0u;
            }
        } else {
            uint32_t w[R];
#pragma unroll
            for (int k = 0; k < R; ++k) w[k] = ok[k] ? rows[ri[k] * L + st.back_col[i]] : 0u;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                b[k] = e[k] = 0;
                if (ok[k]) seg(g, st.back_part[i], w[k], st.label, b[k], e[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (!ok[k]) continue;
            const uint32_t l = e[k] - b[k];
            by += 8ull + 4ull * l;
            if (l < len[k]) {
                len[k] = l;
                beg[k] = b[k];
                ai[k] = i;
            }
        }
    }
    // lower bound of the restriction in the anchor segment, R searches in step
    uint32_t lo[R], n[R];
    bool any = false;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        lo[k] = beg[k];
        n[k] = (ok[k] && bound[k] != 0xffffffffu) ? len[k] : 0u;
        any |= n[k] > 0;
        if (ok[k]) {
            by += 4ull * L;
            rq += 4ull * L + 8ull * st.nback + (bound[k] != 0xffffffffu ? 4ull * probes(len[k]) : 0ull) + 9ull;
        }
    }
    while (any) {
        any = false;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (n[k] == 0) continue;
            const uint32_t h = n[k] >> 1;
            if (g.nbr[lo[k] + h] < bound[k]) {
                lo[k] += h + 1;
                n[k] -= h + 1;
            } else {
                n[k] = h;
            }
            any |= n[k] > 0;
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) cnt[k] = bound[k] == 0xffffffffu ? len[k] : lo[k] - beg[k];
}

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

// Spread accumulators.  Atomics from thousands of waves at one address are
// applied one after another, and a query's embedding count and algorithmic
// bytes used to take one such atomic per wave: a third of the plain path's
// device time on the 1M-vertex Chung-Lu graph.  Each of these sums now lives
// on kAccSlots cache lines (kAccStride words apart); a block adds its total to
// line blockIdx % kAccSlots, and k_fold_acc folds the lines into one word
// before the host reads it.
constexpr uint32_t kAccSlots = 128;
constexpr uint32_t kAccStride = 16;   // 128 B
// The context's accumulator block: words 0..7 (count, bytes, output rows,
// overflow / stop, ...), then the lines from word kAccLine0: word 0 of a line
// sums the count, word 1 the algorithmic bytes.
constexpr uint32_t kAccLine0 = 16;
// words 5 and 6: parents with rows listed for the prefix-shared step (small,
// large)
constexpr uint32_t kAccParSmall = 5, kAccParBig = 6;
constexpr uint32_t kAccWords = kAccLine0 + kAccSlots * kAccStride;

// Adds the block's sums of v0 and v1 (one value per thread) to line
// blockIdx % kAccSlots of the spread accumulators p0 and p1 (nullable).  Every
// thread of the block (kBlock threads) calls it.
__device__ __forceinline__ void block_acc_add2(unsigned long long* p0, uint64_t v0,
                                               unsigned long long* p1, uint64_t v1) {
    __shared__ uint64_t part[kWavesPerBlock][2];
    v0 = wave_sum_u64(v0);
    v1 = wave_sum_u64(v1);
    if (__lane_id() == 0) {
        part[threadIdx.x >> 6][0] = v0;
        part[threadIdx.x >> 6][1] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s0 = 0, s1 = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; ++i) {
            s0 += part[i][0];
            s1 += part[i][1];
        }
        const uint64_t off = (uint64_t)(blockIdx.x % kAccSlots) * kAccStride;
        if (p0 && s0) atomicAdd(p0 + off, (unsigned long long)s0);
        if (p1 && s1) atomicAdd(p1 + off, (unsigned long long)s1);
    }
}


// ---- one tile of the candidate walk, shared by the plain and batch kernels --
//
// A tile is up to 64*IPT consecutive candidate indices of one extension step.
// When the tile lies in a single row (hub rows: almost always), the row's
// data is loaded once and held in scalars (up to 3 same-label values and 2
// non-anchor neighbour segments) and the lanes stream consecutive anchor
// entries; otherwise each lane finds its row by binary search in the scan.

// Candidate prefix of a table's rows: row r's candidates are [begin(r),
// end(r)) of the step's candidate space.  Flat: csum = inclusive prefix per
// row (cnt = the row's count).  Two-level (csum null): the rows form tiles of
// 2^kPTileShift rows, tpre = exclusive prefix of the tiles' totals and loc =
// inclusive prefix of each row within its tile (32 bits).
constexpr uint32_t kPTileShift = 10;
struct Prefix {
    const uint64_t* csum = nullptr;
    const uint32_t* cnt = nullptr;
    const uint64_t* tpre = nullptr;
    const uint32_t* loc = nullptr;
    __device__ __forceinline__ uint64_t end(uint64_t r) const {
        return csum ? csum[r] : tpre[r >> kPTileShift] + loc[r];
    }
    __device__ __forceinline__ uint64_t begin(uint64_t r) const {
        if (csum) return csum[r] - cnt[r];
        const uint64_t t = tpre[r >> kPTileShift];
        return (r & ((1u << kPTileShift) - 1)) ? t + loc[r - 1] : t;
    }
};

// first r in [lo, hi) with end(r) > c
__device__ __forceinline__ uint64_t find_row(const Prefix& p, uint64_t lo, uint64_t hi, uint64_t c) {
    uint64_t n = hi - lo;
    while (n > 0) {
        uint64_t h = n >> 1;
        if (p.end(lo + h) <= c) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

// wave_find_row over a Prefix (see wave_find_row above).
__device__ __forceinline__ uint64_t wave_find_row(const Prefix& pf, uint64_t lo, uint64_t hi, uint64_t c) {
    const unsigned lane = __lane_id();
    while (hi - lo > 1) {
        const uint64_t n = hi - lo;
        const uint64_t step = (n + 63) >> 6;
        const uint64_t x = lo + (uint64_t)lane * step;
        const bool p = x < hi && pf.end(x) > c;
        const uint64_t m = __ballot(p);
        if (m == 0) {
            const uint64_t last = (n - 1) / step < 63 ? (n - 1) / step : 63;
            lo = lo + last * step + 1;
            continue;
        }
        const unsigned j = __builtin_ctzll(m);
        if (j == 0) return lo;
        hi = lo + (uint64_t)j * step + 1;
        lo = lo + (uint64_t)(j - 1) * step + 1;
    }
    return lo;
}

struct TileIn {
    const uint32_t* rows;    // table of the step's input rows (row r at rows + (r - row_base) * L)
    // anchor entries: a row's candidates are anc[anc_beg[r] ..]; the graph's
    // neighbour array, or the parents' intersections of a prefix-shared step
    const uint32_t* anc;
    uint64_t row_base;
    uint32_t L;
    Prefix pf;               // the rows' candidate ranges
    const uint32_t* anc_beg;
    const uint8_t* anc_i;
    uint64_t row_lo, row_hi; // rows searched for the tile's candidates
};

struct TileOut {             // extension output
    uint32_t* out;
    uint64_t cap;            // rows
    unsigned long long* out_n;
    unsigned long long* overflow;
    uint32_t* par = nullptr;  // nullable: input row of every output row
    // ~0: survivors are compacted through reservations; otherwise every
    // candidate survives (one back list, no same-label column) and candidate
    // c is output row c - origin, so no reservation is made
    uint64_t origin = ~0ull;
};

struct TileAcc {
    uint64_t count = 0, bytes = 0, rq = 0;
    uint32_t last_gid = 0xffffffffu;
    TailCache tc;
};

// Output reservation for the survivors of a wave's tile (returns the first
// output row; any value when tot is 0).  Returning atomics on one output
// counter are applied one after another, so a reservation per block round
// instead of per wave tile matters: the waves of a block call
// BlockReserve::operator() together, once per round each (tot 0 and out_n
// null for a wave without an extension tile), and waves of the block that
// write to the same counter share one atomic.
struct BlockReserve {
    uint64_t* sh;   // LDS, 3 * kWavesPerBlock words: tot, counter, base per wave
    __device__ __forceinline__ uint64_t operator()(unsigned long long* out_n, uint32_t tot) const {
        const unsigned w = threadIdx.x >> 6;
        if (__lane_id() == 0) {
            sh[w] = tot;
            sh[kWavesPerBlock + w] = tot ? (uint64_t)out_n : 0ull;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t done = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                if (!sh[i] || ((done >> i) & 1u)) continue;
                const uint64_t p = sh[kWavesPerBlock + i];
                uint64_t sum = 0;
                for (int k = i; k < kWavesPerBlock; ++k)
                    if (sh[k] && sh[kWavesPerBlock + k] == p) sum += sh[k];
                uint64_t base = atomicAdd(reinterpret_cast<unsigned long long*>(p),
                                          (unsigned long long)sum);
                for (int k = i; k < kWavesPerBlock; ++k)
                    if (sh[k] && sh[kWavesPerBlock + k] == p) {
                        sh[2 * kWavesPerBlock + k] = base;
                        base += sh[k];
                        done |= 1u << k;
                    }
            }
        }
        __syncthreads();
        return sh[2 * kWavesPerBlock + w];
    }
};

// A multi-row tile over more than LX rows stages every sr-th row's candidate
// start, up to sr = kLxStride (beyond, its rows are searched in global memory).
constexpr uint32_t kLxStride = 16;

// LDS words per wave of a walk with IPT candidates per lane that stages the
// candidate starts of up to LX rows of a multi-row tile (a tile over more rows
// searches its rows in global memory): stage placed over lx and win (see
// walk_tile).
template <int IPT, uint32_t LX = 64u * IPT>
constexpr uint32_t wave_lds_words() {
    return LX + 1u + 2u * kWin > 64u * GV_MAXK ? LX + 1u + 2u * kWin : 64u * GV_MAXK;
}

// LDS of a wave: lx (LX + 1 words), the candidate starts of the tile's
// rows relative to t0, so that a lane finds its row without global reads; win
// (2 * kWin words), windows of the other back lists; stage (64 * (L + 1)
// words), survivor rows on their way out.  stage is used only after lx and win
// are done with, so a kernel may place it over them (wave_lds_words).
// reserve: the BlockReserve of the block (extension only; every wave of the
// block reaches it once per call of walk_tile).
template <bool kCount, int IPT, uint32_t LX = 64u * IPT>
__device__ __forceinline__ void walk_tile(const DevGraph& g, const gv_step_t& st,
                                          const gv_step_t& tail, bool has_tail, const TileIn& in,
                                          uint64_t t0, uint64_t t1, const TileOut& o,
                                          uint32_t* stage, uint32_t* lx, uint32_t* win,
                                          uint32_t* gbits, TileAcc& acc,
                                          const BlockReserve& reserve) {
    const unsigned lane = __lane_id();
    const uint32_t L = in.L, W = L + 1;
    const uint64_t r_lo = wave_find_row(in.pf, in.row_lo, in.row_hi, t0);
    // the tile's last row is mostly within 64 rows of its first: one round of
    // 64 consecutive rows, then the wide search past them (when none of them
    // ends after t1 - 1, r_lo + 64 < row_hi, since csum[row_hi - 1] does)
    uint64_t r_hi;
    {
        const uint64_t x = r_lo + lane;
        const uint64_t m = __ballot(x < in.row_hi && in.pf.end(x) > t1 - 1);
        r_hi = m ? r_lo + (uint64_t)__builtin_ctzll(m) : wave_find_row(in.pf, r_lo + 64, in.row_hi, t1 - 1);
    }
    auto rowp = [&](uint64_t r) { return in.rows + (r - in.row_base) * L; };

    // anchors (back-edge indices) whose segment is also the fused last
    // vertex's segment
    uint32_t smask = 0;
    if (kCount && has_tail && tail.label == st.label && tail.back_col[0] != L && in.anc == g.nbr) {
        for (uint32_t i = 0; i < st.nback; ++i)
            if (st.back_col[i] == tail.back_col[0] && st.back_part[i] == tail.back_part[0])
                smask |= 1u << i;
    }

    // ---- checked candidates of the tile: ok[it], v[it], r[it], and for the
    // fused last step the index of v in the segment when it is shared (else ~0u)
    bool okk[IPT];
    uint32_t vv[IPT];
    uint64_t rr[IPT];
    uint32_t six[IPT];
    if (r_lo == r_hi && st.ncoll <= 3 && st.nback <= 3) {
        // one row: its data is loaded once and kept in scalars; ids are below
        // 0xffffffff, so an unused same-label slot never matches
        const uint64_t r = r_lo;
        const uint64_t first = in.pf.begin(r);
        const uint32_t abeg = in.anc_beg[r];
        const uint32_t skip = in.anc_i[r];
        const uint32_t* row = rowp(r);
        const uint32_t cv0 = st.ncoll > 0 ? row[st.coll_col[0]] : 0xffffffffu;
        const uint32_t cv1 = st.ncoll > 1 ? row[st.coll_col[1]] : 0xffffffffu;
        const uint32_t cv2 = st.ncoll > 2 ? row[st.coll_col[2]] : 0xffffffffu;
        const bool shared = (smask >> skip) & 1u;
        uint32_t ob0 = 0, oe0 = 0, ob1 = 0, oe1 = 0, nob = 0;
        for (uint32_t i = 0; i < st.nback; ++i) {
            if (i == skip) continue;
            uint32_t b, e;
            seg(g, st.back_part[i], row[st.back_col[i]], st.label, b, e);
            if (nob == 0) {
                ob0 = b;
                oe0 = e;
            } else {
                ob1 = b;
                oe1 = e;
            }
            ++nob;
        }
        if (lane == 0) acc.rq += 4ull * L + 17ull + 8ull * nob;
        // the tile's candidates are consecutive entries of the anchor segment,
        // so they span [vmin, vmax]: the part of each other list in that range
        // is loaded into LDS once and searched there
        uint32_t n0 = 0, n1 = 0;
        bool w0 = false, w1 = false;
        if (nob > 0) {
            const uint32_t vmin = in.anc[abeg + (uint32_t)(t0 - first)];
            const uint32_t vmax = in.anc[abeg + (uint32_t)(t1 - 1 - first)];
            const uint32_t l0 = wave_lower_bound(g.nbr, ob0, oe0, vmin);
            const uint32_t h0 = wave_lower_bound(g.nbr, l0, oe0, vmax + 1);
            n0 = h0 - l0;
            w0 = n0 <= kWin;
            if (lane == 0) acc.rq += 8ull * probes(oe0 - ob0);
            if (w0) {
                for (uint32_t i = lane; i < n0; i += 64) win[i] = g.nbr[l0 + i];
                if (lane == 0) acc.rq += 4ull * n0;
            }
            if (nob > 1) {
                const uint32_t l1 = wave_lower_bound(g.nbr, ob1, oe1, vmin);
                const uint32_t h1 = wave_lower_bound(g.nbr, l1, oe1, vmax + 1);
                n1 = h1 - l1;
                w1 = n1 <= kWin;
                if (lane == 0) acc.rq += 8ull * probes(oe1 - ob1);
                if (w1) {
                    for (uint32_t i = lane; i < n1; i += 64) win[kWin + i] = g.nbr[l1 + i];
                    if (lane == 0) acc.rq += 4ull * n1;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // 32-bit offsets inside the tile (t1 - t0 <= 64 IPT) and the row; a
        // lane takes IPT consecutive candidates, so that its run can be
        // merged with the windows
        const uint32_t tn = (uint32_t)(t1 - t0);
        const uint32_t o0 = (uint32_t)(t0 - first);
        const uint32_t* anc = in.anc + abeg + o0;
        const uint32_t k0 = lane * (uint32_t)IPT;
        const uint32_t nv = tn > k0 ? (tn - k0 < (uint32_t)IPT ? tn - k0 : (uint32_t)IPT) : 0u;
#pragma unroll
        for (int it = 0; it < IPT; ++it) {
            const uint32_t k = k0 + (uint32_t)it;
            uint32_t v = 0;
            bool ok = false;
            if (k < tn) {
                v = anc[k];
                acc.rq += 4;
                ok = v != cv0 && v != cv1 && v != cv2;
            }
            okk[it] = ok;
            vv[it] = v;
            rr[it] = r;
            six[it] = shared ? o0 + k : ~0u;
        }
        if (nob > 0) {
            if (w0) {
                window_member<IPT>(win, n0, vv, nv, okk);
            } else {
#pragma unroll
                for (int it = 0; it < IPT; ++it)
                    if (okk[it]) {
                        acc.rq += 4ull * probes(oe0 - ob0);
                        okk[it] = contains_u32(g.nbr, ob0, oe0, vv[it]);
                    }
            }
        }
        if (nob > 1) {
            if (w1) {
                window_member<IPT>(win + kWin, n1, vv, nv, okk);
            } else {
#pragma unroll
                for (int it = 0; it < IPT; ++it)
                    if (okk[it]) {
                        acc.rq += 4ull * probes(oe1 - ob1);
                        okk[it] = contains_u32(g.nbr, ob1, oe1, vv[it]);
                    }
            }
        }
        // the windows are rewritten by the next tile
        if (nob > 0) __builtin_amdgcn_wave_barrier();
    } else {
        // rows r_lo..r_hi hold the tile: stage the candidate starts of every
        // sr-th of them in LDS relative to t0 (lx[i] = start of row
        // r_lo + i * sr, lx[ns] = end of row r_hi; sr = 1 when the rows fit
        // LX).  Rows after r_lo start inside the tile, so only lx[0] (clamped
        // to 0) and lx[ns] (clamped to 2^32 - 1) are not exact, which the
        // search over c - t0 < 64 * IPT does not see; row r_lo's own start is
        // first0.  A lane finds its run of sr rows in LDS and, for sr > 1,
        // its row among them in global memory.
        const uint64_t nr = r_hi - r_lo + 1;
        const uint32_t sr = nr <= LX ? 1u : (uint32_t)((nr + LX - 1) / LX);
        const bool staged = sr <= kLxStride;
        const uint32_t ns = staged ? (uint32_t)((nr + sr - 1) / sr) : 0u;
        const uint64_t first0 = in.pf.begin(r_lo);
        // up to kRowStage rows: each row's data is also staged once (in the
        // window space, kRowWords words a row): anchor start, the other
        // back lists' segments, the same-label values and whether the row's
        // anchor segment is the fused last vertex's
        const bool rstaged = staged && nr <= kRowStage && st.nback <= 3 && st.ncoll <= 3;
        const uint32_t nob = st.nback - 1;
        if (staged) {
#pragma unroll 4
            for (uint32_t i = lane; i <= ns; i += 64) {
                const uint64_t x = i < ns ? in.pf.begin(r_lo + (uint64_t)i * sr) : in.pf.end(r_hi);
                lx[i] = x <= t0 ? 0u : (x - t0 > 0xffffffffull ? 0xffffffffu : (uint32_t)(x - t0));
            }
            if (rstaged && lane < nr) {
                const uint64_t r = r_lo + lane;
                const uint32_t* row = rowp(r);
                const uint32_t skip = in.anc_i[r];
                uint32_t* q = win + lane * kRowWords;
                q[0] = in.anc_beg[r];
                q[1] = q[2] = q[3] = q[4] = 0;
                uint32_t k = 0;
                for (uint32_t i = 0; i < st.nback; ++i) {
                    if (i == skip) continue;
                    uint32_t b, e;
                    seg(g, st.back_part[i], row[st.back_col[i]], st.label, b, e);
                    q[1 + 2 * k] = b;
                    q[2 + 2 * k] = e;
                    ++k;
                }
                q[5] = st.ncoll > 0 ? row[st.coll_col[0]] : 0xffffffffu;
                q[6] = st.ncoll > 1 ? row[st.coll_col[1]] : 0xffffffffu;
                q[7] = st.ncoll > 2 ? row[st.coll_col[2]] : 0xffffffffu;
                q[8] = (smask >> skip) & 1u;
                acc.rq += 4ull * L + 17ull + 8ull * nob;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll
        for (int it = 0; it < IPT; ++it) {
            if (rstaged) {
                const uint64_t c = t0 + (uint64_t)it * 64 + lane;
                bool ok = false;
                uint32_t v = 0, sx = ~0u, i = 0;
                if (c < t1) {
                    // first i with lx[i + 1] > c - t0
                    const uint32_t cr = (uint32_t)(c - t0);
                    uint32_t n = (uint32_t)nr;
                    while (n > 0) {
                        const uint32_t h = n >> 1;
                        if (lx[i + h + 1] <= cr) {
                            i += h + 1;
                            n -= h + 1;
                        } else {
                            n = h;
                        }
                    }
                    const uint32_t* q = win + i * kRowWords;
                    const uint32_t off = i == 0 ? (uint32_t)(c - first0) : cr - lx[i];
                    v = in.anc[q[0] + off];
                    acc.rq += 4;
                    ok = v != q[5] && v != q[6] && v != q[7];
                    if (ok && nob > 0) {
                        acc.rq += 4ull * probes(q[2] - q[1]);
                        ok = contains_u32(g.nbr, q[1], q[2], v);
                    }
                    if (ok && nob > 1) {
                        acc.rq += 4ull * probes(q[4] - q[3]);
                        ok = contains_u32(g.nbr, q[3], q[4], v);
                    }
                    if (q[8]) sx = off;
                }
                okk[it] = ok;
                vv[it] = v;
                rr[it] = r_lo + i;
                six[it] = sx;
                continue;
            }
            const uint64_t c = t0 + (uint64_t)it * 64 + lane;
            bool ok = false;
            uint32_t v = 0, sx = ~0u;
            uint64_t r = r_lo;
            if (c < t1) {
                uint64_t first;
                if (staged) {
                    // first i with lx[i + 1] > c - t0
                    const uint32_t cr = (uint32_t)(c - t0);
                    uint32_t lo = 0, n = ns;
                    while (n > 0) {
                        const uint32_t h = n >> 1;
                        if (lx[lo + h + 1] <= cr) {
                            lo += h + 1;
                            n -= h + 1;
                        } else {
                            n = h;
                        }
                    }
                    if (sr == 1) {
                        r = r_lo + lo;
                        first = lo == 0 ? first0 : t0 + lx[lo];
                    } else {
                        const uint64_t b0 = r_lo + (uint64_t)lo * sr;
                        const uint64_t b1 = b0 + sr < r_hi + 1 ? b0 + sr : r_hi + 1;
                        r = find_row(in.pf, b0, b1, c);
                        first = in.pf.begin(r);
                    }
                } else {
                    r = find_row(in.pf, r_lo, r_hi + 1, c);
                    first = in.pf.begin(r);
                }
                v = in.anc[in.anc_beg[r] + (uint32_t)(c - first)];
                acc.rq += 4;
                const uint32_t* row = rowp(r);
                ok = true;
                for (uint32_t j = 0; j < st.ncoll && ok; ++j) ok = row[st.coll_col[j]] != v;
                const uint32_t skip = in.anc_i[r];
                if ((smask >> skip) & 1u) sx = (uint32_t)(c - first);
                for (uint32_t i = 0; i < st.nback && ok; ++i) {
                    if (i == skip) continue;
                    uint32_t b, e;
                    seg(g, st.back_part[i], row[st.back_col[i]], st.label, b, e);
                    acc.rq += 8ull + 4ull * probes(e - b);
                    ok = contains_u32(g.nbr, b, e, v);
                }
            }
            // the row and its sizing entries count once per run of lanes sharing it
            const uint64_t prev_r = __shfl_up(r, 1, 64);
            if (c < t1 && (lane == 0 || prev_r != r)) acc.rq += 4ull * L + 17ull;
            okk[it] = ok;
            vv[it] = v;
            rr[it] = r;
            six[it] = sx;
        }
        if (staged) __builtin_amdgcn_wave_barrier();
    }

    if (kCount) {
#pragma unroll
        for (int it = 0; it < IPT; ++it) {
            uint64_t n = okk[it] ? 1 : 0;
            const uint32_t* row = rowp(rr[it]);
            if (okk[it] && has_tail) {
                n = tail_count(g, tail, L, row, vv[it], rr[it], six[it] != ~0u, six[it], acc.tc,
                               acc.rq);
                acc.bytes += 8ull * W + 8ull + 4ull * (acc.tc.e - acc.tc.b);
            }
            acc.count += n;
            if (n && gbits) {
                const uint32_t gid = g.gof[row[0]];
                if (gid != acc.last_gid) {
                    atomicOr(&gbits[gid >> 5], 1u << (gid & 31));
                    acc.last_gid = gid;
                }
            }
        }
        return;
    }
    // extension: one reservation for the whole tile, then each group of 64
    // candidates stages its survivors in LDS and writes them contiguously
    uint32_t tot = 0;
#pragma unroll
    for (int it = 0; it < IPT; ++it) tot += (uint32_t)__popcll(__ballot(okk[it]));
    uint64_t base = o.origin != ~0ull ? t0 - o.origin : reserve(o.out_n, tot);
    if (tot == 0) return;
    if (base + tot > o.cap) {
        if (lane == 0) atomicOr(o.overflow, 1ull);
        return;
    }
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
        const uint64_t mask = __ballot(okk[it]);
        const uint32_t n = (uint32_t)__popcll(mask);
        if (n == 0) continue;
        if (okk[it]) {
            acc.bytes += 4ull * W;
            acc.rq += 4ull * W;
            const uint32_t* row = rowp(rr[it]);
            const uint32_t p = lanes_below(mask) * W;
            for (uint32_t j = 0; j < L; ++j) stage[p + j] = row[j];
            stage[p + L] = vv[it];
            if (o.par) o.par[base + lanes_below(mask)] = (uint32_t)(rr[it] - in.row_base);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        uint32_t* dst = o.out + base * W;
        for (uint32_t j = lane; j < n * W; j += 64) dst[j] = stage[j];
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        base += n;
    }
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
    // a query's folded acc[0..3], written by k_fold_acc straight into pinned
    // coherent host memory (host and device addresses), so reading a result
    // takes no copy
    unsigned long long* hres = nullptr;
    unsigned long long* dres = nullptr;
    // acc and lvl are all zero (gv_mine_levels leaves them so; the other
    // entry points that write them clear the flag)
    bool clean = true;
    uint32_t* gbits = nullptr;
    uint32_t gbits_words = 0;
    // scan scratch
    uint64_t* tiles = nullptr;
    uint64_t tiles_cap = 0;
    int num_cu = 256;
    // profiling: HIP events around every launch, per kernel kind
    bool prof = false;
    bool prof_bytes = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    unsigned long long* preq = nullptr;   // requested bytes per kind (device)
    unsigned long long* punits = nullptr; // rows or candidates per kind (device)
    // rows of each level's table when a whole query runs with extents in
    // device memory (gv_mine_levels)
    unsigned long long* lvl = nullptr;
    // per-vertex weights of a pendant chain (gv_mine_levels), nv1 entries
    unsigned long long* chainw = nullptr;
    uint32_t* chainneed = nullptr;        // chainw[y] is needed when == chain_epoch
    uint64_t chainw_cap = 0;
    // prefix-shared counted step (gv_mine_levels): parent row of every row of
    // the counted table, the parents' intersections and their lengths
    uint32_t* par = nullptr;
    uint64_t par_cap = 0;
    uint32_t* isect = nullptr;
    uint64_t isect_cap = 0;
    uint32_t* ilen = nullptr;
    uint64_t ilen_cap = 0;
    uint32_t* pmark = nullptr;            // parent p has rows when pmark[p] == pmark_epoch
    uint64_t pmark_cap = 0;
    uint32_t pmark_epoch = 0;
    uint32_t* plist = nullptr;            // parents with rows: small ones, then large ones from plist_cap / 2
    uint64_t plist_cap = 0;
    uint32_t chain_epoch = 0;
    // workspace and resume state of the reference-layout launchers
    // (launchers.hip), created on their first call
    struct RefState* ref = nullptr;
};

// Frees ctx->ref (launchers.hip); called by gv_ctx_destroy.
void ref_state_free(gv_ctx* c);

struct gv_graph {
    gv_ctx* ctx = nullptr;
    uint64_t nv1 = 0;
    uint32_t nlabels = 0;
    uint32_t nparts = 0;
    uint64_t nnbr = 0;
    uint32_t max_deg = 0;
    uint64_t dev_bytes = 0;
    std::vector<uint32_t> label_values;       // dense -> original
    std::vector<uint32_t> part_edge_label;    // partition -> edge label
    std::vector<uint64_t> lvoff;              // host copy of start-table offsets
    std::vector<uint64_t> pair_edges;         // [part][NL][NL] directed edges by label pair
    uint32_t* d_loff = nullptr;
    uint32_t* d_nbr = nullptr;
    uint32_t* d_sp = nullptr;
    uint64_t sp_n = 0;
    uint32_t* d_lverts = nullptr;
    uint32_t* d_gof = nullptr;
    uint32_t ngraphs = 0;
    // label-pair edge tables (rows (v, u) of width 2): pe for partition p,
    // labels a, b at rows [pe_off[(p*NL + a)*NL + b], pe_off[... + 1]); pel
    // (u < v, labels a = a) at rows [pel_off[p*NL + a], pel_off[p*NL + a + 1])
    std::vector<uint64_t> pe_off, pel_off;
    uint32_t* d_pe = nullptr;
    uint32_t* d_pel = nullptr;
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
    d.sp = g->d_sp;
    d.sp_n = g->sp_n;
    d.gof = g->d_gof;
    d.nl1 = g->nlabels + 1;
    d.part_stride = g->nv1 * (uint64_t)(g->nlabels + 1);
    return d;
}

enum { kKindSize = 0, kKindScan = 1, kKindExt = 2, kKindCount = 3, kKindBatch = 4, kKindRows = 5,
       kKindIsect = 6, kKindChain = 7, kKindRef = 8, kKinds = 9 };

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
    return (c->prof_bytes && c->preq) ? c->preq + kind : nullptr;
}

unsigned long long* units_ptr(gv_ctx* c, int kind) {
    return (c->prof_bytes && c->punits) ? c->punits + kind : nullptr;
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

// leaf groups over a core of d->width columns; every group's C(s, m) must be
// exact in 64 bits: max_deg^m < 2^64
bool valid_leaves(const gv_leaves_t* d, const gv_graph* g) {
    if (!d || d->width < 1 || d->k > GV_MAXK || d->width >= d->k) return false;
    if (d->ngroups < 1 || d->ngroups > GV_MAX_LEAF_GROUPS) return false;
    uint32_t total = 0;
    for (uint32_t i = 0; i < d->ngroups; ++i) {
        const gv_leaf_group_t& gr = d->group[i];
        if (gr.anchor_col >= d->width || gr.part >= std::max<uint32_t>(g->nparts, 1) ||
            gr.label >= g->nlabels || gr.m < 1 || gr.ncoll >= GV_MAXK)
            return false;
        if (gr.restrict_col != GV_NONE && gr.restrict_col >= d->width) return false;
        for (uint32_t j = 0; j < gr.ncoll; ++j)
            if (gr.coll_col[j] >= d->width) return false;
        double bits = gr.m * std::log2((double)g->max_deg + 1.0);
        if (bits >= 63.0) return false;
        total += gr.m;
    }
    if (total != d->k - d->width) return false;
    uint32_t seen[GV_MAX_LEAF_GROUPS] = {};
    for (uint32_t j = 0; j < total; ++j) {
        if (d->group_of[j] >= d->ngroups) return false;
        seen[d->group_of[j]]++;
    }
    for (uint32_t i = 0; i < d->ngroups; ++i)
        if (seen[i] != d->group[i].m) return false;
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
