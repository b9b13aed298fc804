// libgeneva_hip.so: gfx950 kernels and the C ABI declared in include/geneva_hip.h.
//
// Execution of one extension step over a table of partial embeddings:
//   1. k_size   : per row, pick the shortest back-neighbour list (label
//                 segment of the CSR, O(1) lookup), cut it at the restriction
//                 bound by binary search, record its length.
//   2. k_scan_* : grid-level inclusive scan of the lengths (uint64) = the
//                 candidate index space of the step.
//   3. k_extend : load-balanced expansion over candidate indices.  Each
//                 wave64 owns a contiguous tile of candidates, so a hub row's
//                 candidates are spread over as many waves as it needs;
//                 consecutive lanes read consecutive neighbour entries
//                 (coalesced).  Each candidate is checked against the
//                 other back-neighbour lists by binary search (sorted-set
//                 intersection), the same-label columns (injectivity); the
//                 wave compacts survivors with ballot + mbcnt, reserves output
//                 with one atomic, stages the rows in LDS and writes them as
//                 one contiguous coalesced block.
//   4. k_count  : the same walk for the last step, counting instead of
//                 writing; optionally fused with a last step that has a
//                 single back edge, whose count is a binary search per row.
#include "common.h"

namespace gvi {
thread_local std::string g_last_error;
}

namespace {
// ---------------------------------------------------------------------------
// sizing
struct SizeArgs {
    DevGraph g;
    const uint32_t* rows;
    uint64_t nrows;
    uint32_t L;
    uint32_t* anc_beg;
    uint32_t* cnt;
    uint8_t* anc_i;
    unsigned long long* bytes;
    unsigned long long* req;   // bytes this launch reads and writes, each access once
    unsigned long long* units; // rows sized (profiling)
    // extent in device memory (nullable): rows = min(*nrows_dev, nrows), none
    // when *stop is set
    const unsigned long long* nrows_dev;
    const unsigned long long* stop;
    int64_t start_pos;         // >= 0: rows are the start table at this lverts position
    uint64_t* tile_sums;       // count sum per sizing tile (the scan's first pass)
    gv_step_t st;
};

// rows of a launch whose extent may live in device memory
__device__ __forceinline__ uint64_t dev_rows(uint64_t n, const unsigned long long* n_dev,
                                             const unsigned long long* stop) {
    if (stop && *stop) return 0;
    if (!n_dev) return n;
    const uint64_t d = *n_dev;
    return d < n ? d : n;
}

// kProf: requested bytes and units are accumulated (profiling builds of the
// launch); otherwise that arithmetic is dead and compiled out.
template <bool kProf>
__global__ __launch_bounds__(kBlock) void k_size(SizeArgs a) {
    uint64_t local = 0, rq = 0;
    const uint64_t nrows = dev_rows(a.nrows, a.nrows_dev, a.stop);
    if (kProf && a.units && blockIdx.x == 0 && threadIdx.x == 0 && nrows) atomicAdd(a.units, (unsigned long long)nrows);
    const uint64_t ntiles = (nrows + kSizeTile - 1) / kSizeTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint64_t ri[kSizeRows];
        bool ok[kSizeRows];
#pragma unroll
        for (int k = 0; k < kSizeRows; ++k) {
            ri[k] = t * kSizeTile + (uint64_t)k * kBlock + threadIdx.x;
            ok[k] = ri[k] < nrows;
        }
        uint32_t beg[kSizeRows], len[kSizeRows], cnt[kSizeRows], ai[kSizeRows];
        size_rows<kSizeRows>(a.g, a.st, a.rows, a.L, ri, ok, beg, len, cnt, ai, local, rq, a.start_pos);
        uint64_t ts = 0;
#pragma unroll
        for (int k = 0; k < kSizeRows; ++k) {
            if (!ok[k]) continue;
            a.anc_beg[ri[k]] = beg[k];
            a.cnt[ri[k]] = cnt[k];
            a.anc_i[ri[k]] = (uint8_t)ai[k];
            ts += cnt[k];
        }
        // the tile's count sum: the scan's first pass, done here so that the
        // counts are not read back for it
        __shared__ uint64_t tsum[kWavesPerBlock];
        ts = wave_sum_u64(ts);
        if (lane_id() == 0) tsum[threadIdx.x >> 6] = ts;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t x = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) x += tsum[i];
            a.tile_sums[t] = x;
        }
        __syncthreads();
    }
    if (kProf && a.units) {
        const uint64_t w = wave_sum_u64(local);
        if (lane_id() == 0 && w) atomicAdd(a.units + kKinds, (unsigned long long)w);
    }
    block_acc_add2(nullptr, 0, a.bytes, local);
    if (kProf && a.req) {
        rq = wave_sum_u64(rq);
        if (lane_id() == 0 && rq) atomicAdd(a.req, (unsigned long long)rq);
    }
}

// ---------------------------------------------------------------------------
// count by rows (row_count in common.h): one thread per row, no candidate walk
struct RowsArgs {
    DevGraph g;
    const uint32_t* rows;
    uint64_t nrows;
    uint32_t L;
    int mode;
    gv_step_t st;
    gv_step_t tail;
    unsigned long long* count;
    unsigned long long* bytes;
    unsigned long long* req;
    unsigned long long* units;
    const unsigned long long* nrows_dev;   // as in SizeArgs
    const unsigned long long* stop;
    uint32_t* gbits;
};

template <bool kProf>
__global__ __launch_bounds__(kBlock) void k_count_rows(RowsArgs a) {
    uint64_t cnt = 0, by = 0, rq = 0;
    uint32_t last_gid = 0xffffffffu;
    const uint64_t nrows = dev_rows(a.nrows, a.nrows_dev, a.stop);
    if (kProf && a.units && blockIdx.x == 0 && threadIdx.x == 0 && nrows) atomicAdd(a.units, (unsigned long long)nrows);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        const uint32_t* row = a.rows + r * a.L;
        uint32_t b, e;
        seg(a.g, a.st.back_part[0], row[a.st.back_col[0]], a.st.label, b, e);
        by += 4ull * a.L + 8ull + 4ull * (e - b);
        rq += 4ull * a.L + 8ull;
        const uint64_t n = row_count(a.g, a.st, a.tail, a.mode, a.L, row, b, e, by, rq);
        cnt += n;
        if (n && a.gbits) {
            const uint32_t gid = a.g.gof[row[0]];
            if (gid != last_gid) {
                atomicOr(&a.gbits[gid >> 5], 1u << (gid & 31));
                last_gid = gid;
            }
        }
    }
    if (kProf && a.units) {
        const uint64_t w = wave_sum_u64(by);
        if (lane_id() == 0 && w) atomicAdd(a.units + kKinds, (unsigned long long)w);
    }
    block_acc_add2(a.count, cnt, a.bytes, by);
    if (kProf && a.req) {
        rq = wave_sum_u64(rq);
        if (lane_id() == 0 && rq) atomicAdd(a.req, (unsigned long long)rq);
    }
}

// ---------------------------------------------------------------------------
// count by leaf groups: one thread per core row; the row's embeddings are the
// product over groups of C(s_g, m_g) (gv_leaves_t), and its algorithmic bytes
// follow the rows the enumeration would have held at every width: with j_g
// leaves of group g placed, there are prod_g C(s_g, j_g) of them.
struct LeavesArgs {
    DevGraph g;
    const uint32_t* rows;
    uint64_t nrows;
    const unsigned long long* nrows_dev;   // as in SizeArgs
    const unsigned long long* stop;
    gv_leaves_t d;
    unsigned long long* count;
    unsigned long long* bytes;
    unsigned long long* req;
    unsigned long long* units;
    uint32_t* gbits;
};

template <bool kProf>
__global__ __launch_bounds__(kBlock) void k_count_leaves(LeavesArgs a) {
    uint64_t cnt = 0, by = 0, rq = 0;
    uint32_t last_gid = 0xffffffffu;
    const uint64_t nrows = dev_rows(a.nrows, a.nrows_dev, a.stop);
    if (kProf && a.units && blockIdx.x == 0 && threadIdx.x == 0 && nrows) atomicAdd(a.units, (unsigned long long)nrows);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        const uint32_t* row = a.rows + r * a.d.width;
        rq += 4ull * a.d.width;
        const uint64_t n = leaf_row_count(a.g, a.d, row, by, rq);
        cnt += n;
        if (n && a.gbits) {
            const uint32_t gid = a.g.gof[row[0]];
            if (gid != last_gid) {
                atomicOr(&a.gbits[gid >> 5], 1u << (gid & 31));
                last_gid = gid;
            }
        }
    }
    if (kProf && a.units) {
        const uint64_t w = wave_sum_u64(by);
        if (lane_id() == 0 && w) atomicAdd(a.units + kKinds, (unsigned long long)w);
    }
    block_acc_add2(a.count, cnt, a.bytes, by);
    if (kProf && a.req) {
        rq = wave_sum_u64(rq);
        if (lane_id() == 0 && rq) atomicAdd(a.req, (unsigned long long)rq);
    }
}

// ---------------------------------------------------------------------------
// pendant chain: the last two query vertices u1, u2 hang as a path y - u1 - u2
// off core column a, carry labels no other query vertex has and are not
// restricted, so no constraint ties them to the rest of the row: a core row
// has S(y) = sum over x in N_l1(y) of |N_l2(x)| embeddings.  S is computed once
// per distinct anchor vertex y (marked by the core rows, then one wave per
// marked vertex, lanes over its segment) instead of once per core row.
struct ChainArgs {
    DevGraph g;
    const uint32_t* verts;      // vertices of the anchor's label
    uint64_t nverts;
    uint32_t l1, p1, l2, p2;
    unsigned long long* w;      // S by vertex id
    const uint32_t* need;       // need[y] == epoch: S(y) is read by some core row
    uint32_t epoch;
};

__global__ __launch_bounds__(kBlock) void k_chain_mark(const uint32_t* rows, uint64_t nrows_max,
                                                       const unsigned long long* nrows_dev,
                                                       const unsigned long long* stop, uint32_t L,
                                                       uint32_t acol, uint32_t* need, uint32_t epoch) {
    const uint64_t nrows = dev_rows(nrows_max, nrows_dev, stop);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride)
        need[rows[r * L + acol]] = epoch;
}

constexpr uint32_t kChainVertsPerWave = 4;

__global__ __launch_bounds__(kBlock) void k_chain_weights(ChainArgs a) {
    const unsigned lane = lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    // kChainVertsPerWave vertices per wave: lanes test their marks, then the
    // wave sums each marked vertex's segment together (few vertices per wave,
    // so that many waves hide the dependent loads)
    for (uint64_t base = wave * kChainVertsPerWave; base < a.nverts; base += nwaves * kChainVertsPerWave) {
        const uint64_t i = base + lane;
        const bool in = lane < kChainVertsPerWave && i < a.nverts;
        const uint32_t yl = in ? a.verts[i] : 0u;
        uint64_t m = __ballot(in && a.need[yl] == a.epoch);
        while (m) {
            const unsigned j = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t y = __shfl(yl, j, 64);
            uint32_t b, e;
            seg(a.g, a.p1, y, a.l1, b, e);
            uint64_t s = 0;
            for (uint32_t k = b + lane; k < e; k += 64) {
                uint32_t xb, xe;
                seg(a.g, a.p2, a.g.nbr[k], a.l2, xb, xe);
                s += xe - xb;
            }
            s = wave_sum_u64(s);
            if (lane == 0) a.w[y] = s;
        }
    }
}

struct ChainCountArgs {
    DevGraph g;
    const uint32_t* rows;
    uint64_t nrows;
    const unsigned long long* nrows_dev;   // as in SizeArgs
    const unsigned long long* stop;
    uint32_t L, a, l1, p1;
    const unsigned long long* w;
    unsigned long long* count;
    unsigned long long* bytes;
    unsigned long long* req;
    unsigned long long* units;
    uint32_t* gbits;
};

// Per core row: S(row[a]); algorithmic bytes of the rows the enumeration would
// hold: the core row (4L + 8 + 4 P1), its P1 written extensions and their
// reads (P1 (8(L+1) + 8)) and the last vertex's lists (4 S), P1 = |N_l1(y)|.
template <bool kProf>
__global__ __launch_bounds__(kBlock) void k_count_chain(ChainCountArgs a) {
    uint64_t cnt = 0, by = 0, rq = 0;
    uint32_t last_gid = 0xffffffffu;
    const uint64_t nrows = dev_rows(a.nrows, a.nrows_dev, a.stop);
    if (kProf && a.units && blockIdx.x == 0 && threadIdx.x == 0 && nrows)
        atomicAdd(a.units, (unsigned long long)nrows);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        const uint32_t* row = a.rows + r * a.L;
        const uint32_t y = row[a.a];
        uint32_t b, e;
        seg(a.g, a.p1, y, a.l1, b, e);
        const uint64_t p1 = e - b, s = a.w[y];
        by += 4ull * a.L + 8ull + 4ull * p1 + p1 * (8ull * (a.L + 1) + 8ull) + 4ull * s;
        rq += 4ull * a.L + 16ull;
        cnt += s;
        if (s && a.gbits) {
            const uint32_t gid = a.g.gof[row[0]];
            if (gid != last_gid) {
                atomicOr(&a.gbits[gid >> 5], 1u << (gid & 31));
                last_gid = gid;
            }
        }
    }
    if (kProf && a.units) {
        const uint64_t w = wave_sum_u64(by);
        if (lane_id() == 0 && w) atomicAdd(a.units + kKinds, (unsigned long long)w);
    }
    block_acc_add2(a.count, cnt, a.bytes, by);
    if (kProf && a.req) {
        rq = wave_sum_u64(rq);
        if (lane_id() == 0 && rq) atomicAdd(a.req, (unsigned long long)rq);
    }
}

// ---------------------------------------------------------------------------
// prefix-shared counted step.  The counted vertex u has back edges into the
// parent prefix (columns < L - 1, the same for every row extended from one
// parent row) and possibly to column L - 1.  The prefix lists are
// intersected once per parent (k_isect, after a sizing of the parent table by
// the prefix step), and each row then only meets its parent's intersection
// with the list of column L - 1 (k_child_size sizes it, the counting walk
// reads its candidates from the intersections).  The algorithmic bytes stay
// those of u's step over the full lists, added by k_child_size.
constexpr uint32_t kIsectWave = 256;    // larger parents: k_isect_big
constexpr uint32_t kIsectLds = 4096;    // LDS entries of k_isect_big's lists (16 KB: 8 blocks a CU)

struct IsectArgs {
    DevGraph g;
    const uint32_t* rows;       // parent table, width L
    uint64_t nrows;
    const unsigned long long* nrows_dev;
    const unsigned long long* stop;
    uint32_t L;
    const uint32_t* anc_beg;    // sizing of the parents by the prefix step
    const uint32_t* cnt;
    const uint8_t* anc_i;
    const uint64_t* csum;
    uint32_t* out;              // intersections, parent p's at csum[p] - cnt[p]
    uint64_t out_cap;
    uint32_t* len;              // entries of each parent's intersection
    unsigned long long* overflow;
    const uint32_t* list;       // the parents to intersect (k_mark_parents)
    const unsigned long long* nlist;
    gv_step_t st;               // the prefix step: prefix lists, same-label prefix columns
};

// The parents of the rows, listed for intersection: the small ones (at most
// kIsectWave anchor entries) at list, the others at list + half; one atomic
// per block round and list.  A row whose row before has another parent claims
// its parent with an exchange on mark[p] (rows of one parent are mostly
// adjacent, so few rows try); only the row that finds another epoch there
// lists it.  Each parent is therefore listed at most once, whichever waves or
// chiplets see its rows, and each list holds at most the parent count
// (half >= the parent table's rows).
__global__ __launch_bounds__(kBlock) void k_mark_parents(const uint32_t* par, uint64_t nrows_max,
                                                         const unsigned long long* nrows_dev,
                                                         const unsigned long long* stop,
                                                         const uint32_t* pcnt, uint32_t* mark,
                                                         uint32_t epoch, uint32_t* list, uint64_t half,
                                                         unsigned long long* nsmall,
                                                         unsigned long long* nbig) {
    __shared__ uint32_t wn[2][kWavesPerBlock];
    __shared__ uint64_t wb[2][kWavesPerBlock];
    const unsigned lane = lane_id(), wib = threadIdx.x >> 6;
    const uint64_t n = dev_rows(nrows_max, nrows_dev, stop);
    // block rounds of kBlock rows: the block's waves share one atomic per list
    for (uint64_t b0 = (uint64_t)blockIdx.x * kBlock; b0 < n; b0 += (uint64_t)gridDim.x * kBlock) {
        const uint64_t r = b0 + threadIdx.x;
        const uint32_t p = r < n ? par[r] : 0u;
        const bool fresh = r < n && (r == 0 || par[r - 1] != p) && mark[p] != epoch &&
                           atomicExch(&mark[p], epoch) != epoch;
        const bool big = fresh && pcnt[p] > kIsectWave;
        const uint64_t ms = __ballot(fresh && !big), mb = __ballot(big);
        if (lane == 0) {
            wn[0][wib] = (uint32_t)__popcll(ms);
            wn[1][wib] = (uint32_t)__popcll(mb);
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            const int l = threadIdx.x;
            uint32_t tot = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) tot += wn[l][i];
            uint64_t base = tot ? atomicAdd(l == 0 ? nsmall : nbig, (unsigned long long)tot) : 0ull;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                wb[l][i] = base;
                base += wn[l][i];
            }
        }
        __syncthreads();
        if (fresh && !big) list[wb[0][wib] + lanes_below(ms)] = p;
        if (big) list[half + wb[1][wib] + lanes_below(mb)] = p;
        __syncthreads();
    }
}

// one wave per parent: the anchor's (cut) entries in order, those in every
// other prefix list and not equal to a same-label prefix value written
// contiguously, so each intersection is sorted
__global__ __launch_bounds__(kBlock) void k_isect(IsectArgs a) {
    const unsigned lane = lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t n = dev_rows(a.nrows, a.nrows_dev, a.stop);
    if (n == 0) return;
    if (a.csum[n - 1] > a.out_cap) {
        if (wave == 0 && lane == 0) atomicOr(a.overflow, 1ull);
        return;
    }
    const uint64_t nl = *a.nlist;
    for (uint64_t li = wave; li < nl; li += nwaves) {
        const uint64_t p = a.list[li];
        const uint32_t* row = a.rows + p * a.L;
        const uint32_t c = a.cnt[p], ab = a.anc_beg[p], skip = a.anc_i[p];
        const uint64_t o = a.csum[p] - c;
        uint32_t ob[3] = {0, 0, 0}, oe[3] = {0, 0, 0}, no = 0;
        for (uint32_t i = 0; i < a.st.nback; ++i) {
            if (i == skip) continue;
            uint32_t b, e;
            seg(a.g, a.st.back_part[i], row[a.st.back_col[i]], a.st.label, b, e);
            if (no == 0) {
                ob[0] = b;
                oe[0] = e;
            } else if (no == 1) {
                ob[1] = b;
                oe[1] = e;
            } else {
                ob[2] = b;
                oe[2] = e;
            }
            ++no;
        }
        const uint32_t cv0 = a.st.ncoll > 0 ? row[a.st.coll_col[0]] : 0xffffffffu;
        const uint32_t cv1 = a.st.ncoll > 1 ? row[a.st.coll_col[1]] : 0xffffffffu;
        const uint32_t cv2 = a.st.ncoll > 2 ? row[a.st.coll_col[2]] : 0xffffffffu;
        uint32_t written = 0;
        for (uint32_t k0 = 0; k0 < c; k0 += 64) {
            const uint32_t k = k0 + lane;
            bool ok = k < c;
            const uint32_t v = ok ? a.g.nbr[ab + k] : 0u;
            ok = ok && v != cv0 && v != cv1 && v != cv2;
            if (ok && no > 0) ok = contains_u32(a.g.nbr, ob[0], oe[0], v);
            if (ok && no > 1) ok = contains_u32(a.g.nbr, ob[1], oe[1], v);
            if (ok && no > 2) ok = contains_u32(a.g.nbr, ob[2], oe[2], v);
            const uint64_t m = __ballot(ok);
            if (ok) a.out[o + written + lanes_below(m)] = v;
            written += (uint32_t)__popcll(m);
        }
        if (lane == 0) a.len[p] = written;
    }
}

// Parents with more than kIsectWave anchor entries: one workgroup each, the
// other prefix lists staged in LDS when they fit (kIsectLds entries in all),
// the 4 waves taking 64-entry chunks in rounds and writing their survivors in
// chunk order (a prefix of the waves' counts per round), so the intersection
// stays sorted.
__global__ __launch_bounds__(kBlock) void k_isect_big(IsectArgs a) {
    __shared__ uint32_t lst[kIsectLds];
    __shared__ uint32_t wcnt[kWavesPerBlock];
    const unsigned lane = lane_id(), wib = threadIdx.x >> 6;
    const uint64_t n = dev_rows(a.nrows, a.nrows_dev, a.stop);
    if (n == 0 || a.csum[n - 1] > a.out_cap) return;   // k_isect flags the overflow
    const uint64_t nl = *a.nlist;
    for (uint64_t li = blockIdx.x; li < nl; li += gridDim.x) {
        const uint64_t p = a.list[li];
        const uint32_t c = a.cnt[p];
        const uint32_t* row = a.rows + p * a.L;
        const uint32_t ab = a.anc_beg[p], skip = a.anc_i[p];
        const uint64_t o = a.csum[p] - c;
        uint32_t ob[3] = {0, 0, 0}, oe[3] = {0, 0, 0}, no = 0;
        for (uint32_t i = 0; i < a.st.nback; ++i) {
            if (i == skip) continue;
            uint32_t b, e;
            seg(a.g, a.st.back_part[i], row[a.st.back_col[i]], a.st.label, b, e);
            if (no == 0) {
                ob[0] = b;
                oe[0] = e;
            } else if (no == 1) {
                ob[1] = b;
                oe[1] = e;
            } else {
                ob[2] = b;
                oe[2] = e;
            }
            ++no;
        }
        // stage the other lists in LDS (one after another) when they fit
        const uint32_t l0 = oe[0] - ob[0], l1 = oe[1] - ob[1], l2 = oe[2] - ob[2];
        const bool in_lds = (uint64_t)l0 + l1 + l2 <= kIsectLds;
        if (in_lds) {
            for (uint32_t i = threadIdx.x; i < l0; i += kBlock) lst[i] = a.g.nbr[ob[0] + i];
            for (uint32_t i = threadIdx.x; i < l1; i += kBlock) lst[l0 + i] = a.g.nbr[ob[1] + i];
            for (uint32_t i = threadIdx.x; i < l2; i += kBlock) lst[l0 + l1 + i] = a.g.nbr[ob[2] + i];
        }
        __syncthreads();
        const uint32_t cv0 = a.st.ncoll > 0 ? row[a.st.coll_col[0]] : 0xffffffffu;
        const uint32_t cv1 = a.st.ncoll > 1 ? row[a.st.coll_col[1]] : 0xffffffffu;
        const uint32_t cv2 = a.st.ncoll > 2 ? row[a.st.coll_col[2]] : 0xffffffffu;
        uint32_t written = 0;
        for (uint32_t r0 = 0; r0 < c; r0 += 64 * kWavesPerBlock) {
            const uint32_t k = r0 + wib * 64 + lane;
            bool ok = k < c;
            const uint32_t v = ok ? a.g.nbr[ab + k] : 0u;
            ok = ok && v != cv0 && v != cv1 && v != cv2;
            if (in_lds) {
                if (ok && no > 0) {
                    const uint32_t q = lds_lower_bound(lst, 0, l0, v);
                    ok = q < l0 && lst[q] == v;
                }
                if (ok && no > 1) {
                    const uint32_t q = lds_lower_bound(lst, l0, l0 + l1, v);
                    ok = q < l0 + l1 && lst[q] == v;
                }
                if (ok && no > 2) {
                    const uint32_t q = lds_lower_bound(lst, l0 + l1, l0 + l1 + l2, v);
                    ok = q < l0 + l1 + l2 && lst[q] == v;
                }
            } else {
                if (ok && no > 0) ok = contains_u32(a.g.nbr, ob[0], oe[0], v);
                if (ok && no > 1) ok = contains_u32(a.g.nbr, ob[1], oe[1], v);
                if (ok && no > 2) ok = contains_u32(a.g.nbr, ob[2], oe[2], v);
            }
            const uint64_t m = __ballot(ok);
            if (lane == 0) wcnt[wib] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t before = 0, all = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                before += i < (int)wib ? wcnt[i] : 0u;
                all += wcnt[i];
            }
            if (ok) a.out[o + written + before + lanes_below(m)] = v;
            written += all;
            __syncthreads();
        }
        if (threadIdx.x == 0) a.len[p] = written;
    }
}

struct ChildSizeArgs {
    DevGraph g;
    const uint32_t* rows;       // counted table, width L
    uint64_t nrows;
    const unsigned long long* nrows_dev;
    const unsigned long long* stop;
    uint32_t L;
    const uint32_t* par;        // parent of each row
    const uint32_t* pcnt;       // parents' sizing (offsets of their intersections)
    const uint64_t* pcsum;
    const uint32_t* ilen;
    const uint32_t* isect;
    uint32_t* anc_beg;          // the counted step's sizing, anchor = the intersection
    uint32_t* cnt;
    uint8_t* anc_i;
    uint64_t* tile_sums;        // as in k_size
    unsigned long long* bytes;
    gv_step_t st;               // u's step (its full lists give the algorithmic bytes)
    uint32_t cut_last;          // u is restricted below column L - 1
};

__global__ __launch_bounds__(kBlock) void k_child_size(ChildSizeArgs a) {
    uint64_t by = 0;
    const uint64_t nrows = dev_rows(a.nrows, a.nrows_dev, a.stop);
    const uint64_t ntiles = (nrows + kSizeTile - 1) / kSizeTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint64_t ts = 0;
        for (int k = 0; k < kSizeRows; ++k) {
            const uint64_t r = t * kSizeTile + (uint64_t)k * kBlock + threadIdx.x;
            if (r >= nrows) continue;
            const uint32_t* row = a.rows + r * a.L;
            const uint32_t p = a.par[r];
            const uint32_t o = (uint32_t)(a.pcsum[p] - a.pcnt[p]);
            uint32_t c = a.ilen[p];
            if (a.cut_last) c = lower_bound_u32(a.isect, o, o + c, row[a.L - 1]) - o;
            a.anc_beg[r] = o;
            a.cnt[r] = c;
            a.anc_i[r] = 0;
            ts += c;
            by += 4ull * a.L;
            for (uint32_t i = 0; i < a.st.nback; ++i) {
                uint32_t b, e;
                seg(a.g, a.st.back_part[i], row[a.st.back_col[i]], a.st.label, b, e);
                by += 8ull + 4ull * (e - b);
            }
        }
        __shared__ uint64_t tsum[kWavesPerBlock];
        ts = wave_sum_u64(ts);
        if (lane_id() == 0) tsum[threadIdx.x >> 6] = ts;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t x = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) x += tsum[i];
            a.tile_sums[t] = x;
        }
        __syncthreads();
    }
    block_acc_add2(nullptr, 0, a.bytes, by);
}

constexpr uint32_t kChainMinDegree = 256;

// The pendant chain of the last two steps of q (counted step w = k - 2),
// or false.
bool chain_tail(const gv_query_t* q, uint32_t w) {
    if (q->k < 3 || w + 2 != q->k) return false;
    const gv_step_t& u1 = q->step[w];
    const gv_step_t& u2 = q->step[w + 1];
    // single back edges, u1 into the core, u2 onto u1; no restriction; no
    // earlier vertex with either label (u2's list would include u1 if they
    // shared one)
    return u1.nback == 1 && u1.back_col[0] < w && u1.ncoll == 0 && u1.restrict_col == GV_NONE &&
           u2.nback == 1 && u2.back_col[0] == w && u2.ncoll == 0 && u2.restrict_col == GV_NONE;
}

// Counted step w of q split into its prefix part (back edges and same-label
// columns below w - 1, the restriction if it lies there) and the part left to
// each row (the edge to column w - 1, its same-label check and restriction),
// when at least two back edges lie in the prefix (k_isect).  child's list 0
// stands for the parent's intersection (its column is never read).
bool prefix_split(const gv_query_t* q, uint32_t w, gv_step_t& pre, gv_step_t& child,
                  uint32_t& cut_last) {
    if (w < 3) return false;
    const gv_step_t& st = q->step[w];
    pre = st;
    pre.nback = 0;
    bool has_last = false;
    uint8_t part_last = 0;
    for (uint32_t i = 0; i < st.nback; ++i) {
        if (st.back_col[i] < w - 1) {
            pre.back_col[pre.nback] = st.back_col[i];
            pre.back_part[pre.nback] = st.back_part[i];
            ++pre.nback;
        } else {
            has_last = true;
            part_last = st.back_part[i];
        }
    }
    if (pre.nback < 2 || pre.nback > 4) return false;
    pre.restrict_col = st.restrict_col != GV_NONE && st.restrict_col < w - 1 ? st.restrict_col : GV_NONE;
    pre.ncoll = 0;
    child = st;
    child.ncoll = 0;
    for (uint32_t j = 0; j < st.ncoll; ++j) {
        if (st.coll_col[j] < w - 1) pre.coll_col[pre.ncoll++] = st.coll_col[j];
        else child.coll_col[child.ncoll++] = st.coll_col[j];
    }
    if (pre.ncoll > 3) return false;
    child.nback = has_last ? 2 : 1;
    child.back_col[0] = 0;
    child.back_part[0] = pre.back_part[0];
    child.back_col[1] = (uint8_t)(w - 1);
    child.back_part[1] = part_last;
    child.restrict_col = GV_NONE;
    cut_last = st.restrict_col == w - 1 ? 1u : 0u;
    return true;
}

// grow-only device buffer of n uint32
int ensure_u32(gv_ctx* c, uint32_t*& p, uint64_t& cap, uint64_t n) {
    if (n <= cap) return GV_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, std::max<uint64_t>(n, 1) * sizeof(uint32_t)) != hipSuccess) return GV_ERR_NOMEM;
    cap = n;
    return GV_OK;
}

// ---------------------------------------------------------------------------
// grid-level inclusive scan of uint32 counts into uint64 sums (reduce, scan of
// tile sums, scan + add)

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* wsum,
                                                         uint64_t& total) {
    const unsigned lane = lane_id(), w = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(x, o, 64);
        if (lane >= (unsigned)o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint64_t off = 0;
    total = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
        if (i < (int)w) off += wsum[i];
        total += wsum[i];
    }
    __syncthreads();
    return off + x - v;
}

// The scan of a sizing launch's counts into inclusive uint64 sums has its
// first pass in k_size (one count sum per kSizeTile rows); n may come from
// device memory (n_dev, clamped to n_max; nothing when *stop is set).
constexpr int kFinalItems = kSizeRows;                    // counts per thread
constexpr int kFinalTile = kBlock * kFinalItems;          // rows per tile
constexpr int kFinalPadded = kFinalTile + kFinalTile / 16;
constexpr int kTileItems = 16;                            // tile sums per thread per pass
static_assert((uint64_t)kFinalTile == kSizeTile, "scan tiles are sizing tiles");

// exclusive scan of the tile sums in place, one workgroup (kTileItems
// consecutive sums per thread per pass)
__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint64_t* tile_sums, uint64_t n_max,
                                                       const unsigned long long* n_dev,
                                                       const unsigned long long* stop,
                                                       unsigned long long* units,
                                                       unsigned long long* req) {
    __shared__ uint64_t wsum[kWavesPerBlock];
    const uint64_t n = dev_rows(n_max, n_dev, stop);
    const uint64_t ntiles = (n + kFinalTile - 1) / kFinalTile;
    if (threadIdx.x == 0 && n) {
        if (units) atomicAdd(units, (unsigned long long)n);
        // counts read, sums written; a tile sum written (k_size), read and
        // written here, read by k_scan_final
        if (req) atomicAdd(req, (unsigned long long)(12ull * n + 32ull * ntiles));
    }
    uint64_t carry = 0;
    for (uint64_t b = 0; b < ntiles; b += (uint64_t)kBlock * kTileItems) {
        const uint64_t first = b + (uint64_t)threadIdx.x * kTileItems;
        uint64_t loc[kTileItems];
        uint64_t s = 0;
#pragma unroll
        for (int i = 0; i < kTileItems; ++i) {
            const uint64_t idx = first + i;
            loc[i] = s;
            s += idx < ntiles ? tile_sums[idx] : 0;
        }
        uint64_t tot;
        const uint64_t ex = carry + block_exclusive_scan(s, wsum, tot);
#pragma unroll
        for (int i = 0; i < kTileItems; ++i) {
            const uint64_t idx = first + i;
            if (idx < ntiles) tile_sums[idx] = ex + loc[i];
        }
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_final(const uint32_t* in, uint64_t n_max,
                                                       const unsigned long long* n_dev,
                                                       const unsigned long long* stop,
                                                       const uint64_t* tile_off, uint64_t* out) {
    // the counts and the sums share one padded tile: the counts are all read
    // (and the block scan has synchronised) before the first sum is written
    __shared__ uint64_t outv[kFinalPadded / 2];
    uint32_t* vals = reinterpret_cast<uint32_t*>(outv);
    __shared__ uint64_t wsum[kWavesPerBlock];
    const uint64_t n = dev_rows(n_max, n_dev, stop);
    const uint64_t ntiles = (n + kFinalTile - 1) / kFinalTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t base = t * kFinalTile;
        for (int i = 0; i < kFinalItems; ++i) {
            const uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
            vals[lds_pad16(i * kBlock + threadIdx.x)] = idx < n ? in[idx] : 0u;
        }
        __syncthreads();
        uint64_t loc[kFinalItems];
        uint64_t s = 0;
#pragma unroll
        for (int i = 0; i < kFinalItems; ++i) {
            s += vals[lds_pad16(threadIdx.x * kFinalItems + i)];
            loc[i] = s;
        }
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan(s, wsum, tot) + tile_off[t];
        // write back through LDS in two halves to keep coalesced stores
        for (int half = 0; half < 2; ++half) {
            __syncthreads();
            const int lo = half * (kFinalTile / 2);
#pragma unroll
            for (int i = 0; i < kFinalItems; ++i) {
                const int p = threadIdx.x * kFinalItems + i - lo;
                if (p >= 0 && p < kFinalTile / 2) outv[lds_pad16(p)] = ex + loc[i];
            }
            __syncthreads();
            for (int i = threadIdx.x; i < kFinalTile / 2; i += kBlock) {
                const uint64_t idx = base + lo + i;
                if (idx < n) out[idx] = outv[lds_pad16(i)];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// extension and counting over candidate tiles

struct WalkArgs {
    DevGraph g;
    const uint32_t* rows;
    uint64_t nrows;
    uint32_t L;
    const uint32_t* anc_beg;
    const uint32_t* cnt;
    const uint64_t* csum;     // inclusive
    const uint8_t* anc_i;
    uint64_t c0, c1;
    // extension output
    uint32_t* out;
    uint64_t out_cap;
    unsigned long long* out_n;
    unsigned long long* overflow;
    // counting
    unsigned long long* count;
    uint32_t* gbits;
    unsigned long long* bytes;
    unsigned long long* req;   // bytes this launch reads and writes, each access once
    unsigned long long* units; // candidates walked (profiling)
    // extent in device memory (nullable): rows = min(*nrows_dev, nrows) and
    // the candidates are all of theirs, [0, csum[rows - 1]); none when *stop
    const unsigned long long* nrows_dev;
    const unsigned long long* stop;
    // prefix-shared step (nullable): candidates come from anc_base instead of
    // the graph's neighbour array; extension: the input row of every output
    // row is written to par
    const uint32_t* anc_base;
    uint32_t* par;
    gv_step_t st;
    gv_step_t tail;
    int has_tail;
};



// Eight waves per SIMD (at most 64 VGPRs): the walk waits on dependent loads,
// and more waves in flight hide more of that latency than the few registers
// the limit spills cost.
template <bool kCount, bool kProf, int IPT = kIPT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_walk(WalkArgs a) {
    __shared__ uint32_t wl[kWavesPerBlock][wave_lds_words<IPT>()];
    constexpr uint64_t kT = 64 * IPT;
    const unsigned lane = lane_id();
    const unsigned wib = threadIdx.x >> 6;
    uint64_t nrows = a.nrows, c0 = a.c0, c1 = a.c1;
    if (a.nrows_dev || a.stop) {
        nrows = dev_rows(a.nrows, a.nrows_dev, a.stop);
        c0 = 0;
        c1 = nrows ? a.csum[nrows - 1] : 0;
    }
    if (kProf && a.units && blockIdx.x == 0 && threadIdx.x == 0 && c1 > c0)
        atomicAdd(a.units, (unsigned long long)(c1 - c0));
    TileIn in;
    in.rows = a.rows;
    in.anc = a.anc_base ? a.anc_base : a.g.nbr;
    in.row_base = 0;
    in.L = a.L;
    in.pf.csum = a.csum;
    in.pf.cnt = a.cnt;
    in.anc_beg = a.anc_beg;
    in.anc_i = a.anc_i;
    in.row_lo = 0;
    in.row_hi = nrows;
    TileOut o{a.out, a.out_cap, a.out_n, a.overflow, a.par};
    // a step with one back list and no same-label column keeps every
    // candidate: rows go to their candidate index, the count is the total
    const bool direct = !kCount && a.st.nback == 1 && a.st.ncoll == 0;
    if (direct) {
        o.origin = c0;
        if (blockIdx.x == 0 && threadIdx.x == 0 && c1 > c0) atomicAdd(a.out_n, (unsigned long long)(c1 - c0));
    }
    TileAcc acc;
    __shared__ uint64_t resv[3 * kWavesPerBlock];
    const BlockReserve reserve{resv};
    // block rounds: wave wib of block b takes tile (round * blocks + b) * 4 + wib,
    // and every wave of the block runs every round (the reservation is shared)
    const uint64_t nb = gridDim.x;
    for (uint64_t T0 = (uint64_t)blockIdx.x * kWavesPerBlock; c0 + T0 * kT < c1;
         T0 += nb * kWavesPerBlock) {
        const uint64_t t0 = c0 + (T0 + wib) * kT;
        if (t0 < c1) {
            const uint64_t t1 = t0 + kT < c1 ? t0 + kT : c1;
            walk_tile<kCount, IPT>(a.g, a.st, a.tail, a.has_tail != 0, in, t0, t1, o,
                                    kCount ? nullptr : wl[wib], wl[wib], wl[wib] + kT + 1, a.gbits, acc,
                                    reserve);
        } else if (!kCount && !direct) {
            reserve(nullptr, 0);
        }
    }
    if (kProf && a.units) {
        const uint64_t w = wave_sum_u64(acc.bytes);
        if (lane == 0 && w) atomicAdd(a.units + kKinds, (unsigned long long)w);
    }
    block_acc_add2(kCount ? a.count : nullptr, acc.count, a.bytes, acc.bytes);
    if (kProf && a.req) {
        acc.rq = wave_sum_u64(acc.rq);
        if (lane == 0 && acc.rq) atomicAdd(a.req, (unsigned long long)acc.rq);
    }
}

// zero a query's accumulators, level counters and graph bitset in one launch
__global__ void k_zero_query(unsigned long long* acc, uint32_t nacc, unsigned long long* lvl,
                             uint32_t nlvl, uint32_t* bits, uint32_t nbits) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x < nlvl) lvl[threadIdx.x] = 0;
    for (uint32_t i = t; i < nacc; i += stride) acc[i] = 0;
    for (uint32_t i = t; i < nbits; i += stride) bits[i] = 0;
}

// acc[0] and acc[1] = the sums over the spread lines of the count and the
// algorithmic bytes (one wave of 64); with out, acc[0..3] go there instead
// and acc (nacc words) and lvl (nlvl words) are cleared for the next query
__global__ void k_fold_acc(unsigned long long* acc, unsigned long long* out, uint32_t nacc,
                           unsigned long long* lvl, uint32_t nlvl) {
    uint64_t cnt = 0, by = 0;
    for (uint32_t i = threadIdx.x; i < kAccSlots; i += 64) {
        cnt += acc[kAccLine0 + i * kAccStride];
        by += acc[kAccLine0 + i * kAccStride + 1];
    }
    cnt = wave_sum_u64(cnt);
    by = wave_sum_u64(by);
    if (!out) {
        if (threadIdx.x == 0) {
            acc[0] = cnt;
            acc[1] = by;
        }
        return;
    }
    const uint64_t rows = acc[2], over = acc[3];
    __syncthreads();   // every read of acc is done before it is cleared
    if (threadIdx.x == 0) {
        out[0] = cnt;
        out[1] = by;
        out[2] = rows;
        out[3] = over;
    }
    for (uint32_t i = threadIdx.x; i < nacc; i += 64) acc[i] = 0;
    for (uint32_t i = threadIdx.x; i < nlvl; i += 64) lvl[i] = 0;
}

__global__ void k_copy_back(const uint32_t* src, uint32_t* dst, uint64_t n) {
    // overlap-safe when dst >= src: each element is moved by the thread that
    // owns it, highest indices first within the grid-stride sweep
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0) return;
    for (uint64_t k = tid; k < n; k += stride) {
        uint64_t i = n - 1 - k;
        dst[i] = src[i];
    }
}

}  // namespace

namespace {

// lverts position of a width-1 table that lies inside lverts (a start table
// or a contiguous shard of one), else -1
int64_t start_pos_of(const gv_graph* g, const uint32_t* rows, uint32_t L) {
    if (L != 1 || !g->d_lverts || rows < g->d_lverts || rows >= g->d_lverts + g->sp_n) return -1;
    return (int64_t)(rows - g->d_lverts);
}

// the context's graph bitset holds at least the graph's owners
int ensure_gbits(gv_ctx* c, const gv_graph* g) {
    if (!g || !g->ngraphs) return GV_OK;
    const uint32_t words = (g->ngraphs + 31) / 32;
    if (words <= c->gbits_words) return GV_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->gbits) HIPCHK(hipFree(c->gbits));
    c->gbits = nullptr;
    c->gbits_words = 0;
    if (hipMalloc(&c->gbits, words * sizeof(uint32_t)) != hipSuccess) return GV_ERR_NOMEM;
    c->gbits_words = words;
    return GV_OK;
}

// tile sums of a sizing launch over n rows
int ensure_tiles(gv_ctx* c, uint64_t n) {
    const uint64_t ntiles = (n + kSizeTile - 1) / kSizeTile;
    if (ntiles > c->tiles_cap) {
        if (c->tiles) HIPCHK(hipFree(c->tiles));
        c->tiles = nullptr;
        uint64_t cap = std::max<uint64_t>(ntiles, 1024);
        if (hipMalloc(&c->tiles, cap * sizeof(uint64_t)) != hipSuccess) {
            c->tiles_cap = 0;
            return GV_ERR_NOMEM;
        }
        c->tiles_cap = cap;
    }
    return GV_OK;
}

// Inclusive scan of n counts (n_max when n_dev is null, else min(*n_dev,
// n_max) read on the device; nothing when *stop is set) that the sizing launch
// just before wrote, with their tile sums in c->tiles.
int scan_counts(gv_ctx* c, const uint32_t* cnt, uint64_t n_max, const unsigned long long* n_dev,
                const unsigned long long* stop, uint64_t* csum) {
    const uint64_t ntiles = (n_max + kSizeTile - 1) / kSizeTile;
    const unsigned grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(ntiles, 1), 16ull * c->num_cu);
    ProfScope ps(c, kKindScan);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, c->stream, c->tiles, n_max, n_dev,
                       stop, units_ptr(c, kKindScan), req_ptr(c, kKindScan));
    hipLaunchKernelGGL(k_scan_final, dim3(grid), dim3(kBlock), 0, c->stream, cnt, n_max, n_dev, stop,
                       c->tiles, csum);
    HIPCHK(hipGetLastError());
    return GV_OK;
}

}  // namespace

extern "C" {

const char* gv_status_string(int s) {
    switch (s) {
        case GV_OK: return "ok";
        case GV_ERR_ARG: return "invalid argument";
        case GV_ERR_HIP: return "HIP runtime error";
        case GV_ERR_NOMEM: return "device allocation failed";
        case GV_ERR_CAPACITY: return "output capacity exceeded";
        case GV_ERR_FORMAT: return "malformed partition block";
        case GV_ERR_NODEV: return "no HIP device";
        default: return "unknown status";
    }
}

const char* gv_last_error(void) { return g_last_error.c_str(); }

int gv_device_count(int* n) {
    if (!n) return GV_ERR_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return GV_OK;
}

int gv_ctx_create(int device, gv_ctx_t* out) {
    if (!out) return GV_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GV_ERR_NODEV;
    if (device < 0 || device >= n) return GV_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    gv_ctx* c = new gv_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GV_ERR_HIP;
    }
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    if (hipMalloc(&c->acc, kAccWords * sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return GV_ERR_NOMEM;
    }
    HIPCHK(hipMemsetAsync(c->acc, 0, kAccWords * sizeof(unsigned long long), c->stream));
    if (hipMalloc(&c->lvl, (GV_MAXK + 1) * sizeof(unsigned long long)) != hipSuccess) {
        hipFree(c->acc);
        delete c;
        return GV_ERR_NOMEM;
    }
    HIPCHK(hipMemsetAsync(c->lvl, 0, (GV_MAXK + 1) * sizeof(unsigned long long), c->stream));
    if (hipHostMalloc((void**)&c->hres, 4 * sizeof(unsigned long long), hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&c->dres, c->hres, 0) != hipSuccess) {
        gv_ctx_destroy(c);
        return GV_ERR_NOMEM;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    *out = c;
    return GV_OK;
}

int gv_ctx_destroy(gv_ctx_t c) {
    if (!c) return GV_ERR_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (c->acc) hipFree(c->acc);
    if (c->hres) hipHostFree(c->hres);
    if (c->gbits) hipFree(c->gbits);
    if (c->tiles) hipFree(c->tiles);
    if (c->preq) hipFree(c->preq);
    if (c->punits) hipFree(c->punits);
    if (c->lvl) hipFree(c->lvl);
    if (c->chainw) hipFree(c->chainw);
    if (c->chainneed) hipFree(c->chainneed);
    if (c->par) hipFree(c->par);
    if (c->isect) hipFree(c->isect);
    if (c->ilen) hipFree(c->ilen);
    if (c->pmark) hipFree(c->pmark);
    if (c->plist) hipFree(c->plist);
    ref_state_free(c);
    for (auto& r : c->recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (auto& e : c->pool) hipEventDestroy(e);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    hipStreamDestroy(c->stream);
    delete c;
    return GV_OK;
}

int gv_ctx_synchronize(gv_ctx_t c) {
    if (!c) return GV_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    return GV_OK;
}

int gv_mem_get_info(gv_ctx_t c, uint64_t* f, uint64_t* t) {
    if (!c || !f || !t) return GV_ERR_ARG;
    size_t a = 0, b = 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemGetInfo(&a, &b));
    *f = a;
    *t = b;
    return GV_OK;
}

int gv_malloc(gv_ctx_t c, uint64_t bytes, void** p) {
    if (!c || !p) return GV_ERR_ARG;
    *p = nullptr;
    if (bytes == 0) return GV_OK;
    if (hipSetDevice(c->device) != hipSuccess) return GV_ERR_HIP;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        g_last_error = "hipMalloc failed";
        return GV_ERR_NOMEM;
    }
    return GV_OK;
}

int gv_free(gv_ctx_t c, void* p) {
    if (!c) return GV_ERR_ARG;
    if (p) HIPCHK(hipFree(p));
    return GV_OK;
}

int gv_memcpy_h2d(gv_ctx_t c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return GV_ERR_ARG;
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return GV_OK;
}

int gv_memcpy_d2h(gv_ctx_t c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return GV_ERR_ARG;
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GV_OK;
}

// Decode the reference partition blocks into the label-segmented CSR.
int gv_graph_create(gv_ctx_t c, const uint32_t* vLabel, uint32_t nblocks,
                    const uint32_t* const* blocks, const uint32_t* block_edge_labels,
                    gv_graph_t* out) {
    if (!c || !vLabel || !out || (nblocks && (!blocks || !block_edge_labels)) || nblocks > 255)
        return GV_ERR_ARG;
    *out = nullptr;
    const uint32_t maxid = vLabel[0];
    const uint64_t nv1 = (uint64_t)maxid + 1;
    // dense labels, ascending original value
    std::vector<uint32_t> vals;
    vals.reserve(nv1);
    for (uint64_t v = 1; v < nv1; ++v)
        if (vLabel[v] != 0xffffffffu) vals.push_back(vLabel[v]);
    std::sort(vals.begin(), vals.end());
    vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
    const uint32_t NL = (uint32_t)vals.size();
    if (NL == 0 || NL > 65535) return GV_ERR_FORMAT;
    std::vector<uint32_t> dl(nv1, 0xffffffffu);
    for (uint64_t v = 1; v < nv1; ++v)
        if (vLabel[v] != 0xffffffffu)
            dl[v] = (uint32_t)(std::lower_bound(vals.begin(), vals.end(), vLabel[v]) - vals.begin());
    const uint64_t nl1 = NL + 1;
    std::vector<uint32_t> loff(nv1 * nl1 * (uint64_t)std::max<uint32_t>(nblocks, 1), 0u);
    std::vector<uint32_t> nbr;
    std::vector<uint64_t> pairs((uint64_t)std::max<uint32_t>(nblocks, 1) * NL * NL, 0);
    for (uint32_t p = 0; p < nblocks; ++p) {
        const uint32_t* blk = blocks[p];
        const uint32_t tot = blk[0], S = blk[1], n = blk[2], u4 = blk[4];
        if (tot < 5 || n == 0 || n > 256 || 5ull + 2 * n + 2 + S + 1 > tot || 5ull + u4 > tot)
            return GV_ERR_FORMAT;
        const uint32_t* a0 = blk + 5;
        const uint32_t* a1 = blk + 5 + 2 * n + 2;
        const uint32_t* a3 = blk + 5 + u4;
        const uint64_t nn = tot - 5 - u4;
        uint32_t* lo = loff.data() + (uint64_t)p * nv1 * nl1;
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t start = a0[j], plo = a0[n + j], phi = a0[n + 1 + j];
            if (phi < plo || phi > S) return GV_ERR_FORMAT;
            for (uint32_t k = plo; k < phi; ++k) {
                const uint64_t src = (uint64_t)start + (k - plo);
                const uint32_t b = a1[k], e = a1[k + 1];
                if (src == 0 || src >= nv1 || e < b || e > nn) return GV_ERR_FORMAT;
                uint32_t* row = lo + src * nl1;
                uint32_t next_label = 0;
                uint32_t q = b;
                while (q < e) {
                    const uint32_t x = a3[q];
                    if (x == 0 || x >= nv1 || dl[x] == 0xffffffffu) return GV_ERR_FORMAT;
                    const uint32_t lab = dl[x];
                    uint32_t r = q;
                    while (r < e && a3[r] < nv1 && dl[a3[r]] == lab) ++r;
                    if (lab < next_label) return GV_ERR_FORMAT;  // groups must ascend
                    const uint32_t L = r - q;
                    uint32_t pre = 0;
                    if (L > 130 && L <= 4128) pre = (L + 128) / 129;
                    else if (L > 4128) pre = 32;
                    for (uint32_t l = next_label; l <= lab; ++l) row[l] = (uint32_t)nbr.size();
                    if (dl[src] != 0xffffffffu)
                        pairs[((uint64_t)p * NL + dl[src]) * NL + lab] += r - q - pre;
                    for (uint32_t s = q + pre; s < r; ++s) {
                        if (a3[s] == 0 || a3[s] >= nv1) return GV_ERR_FORMAT;
                        nbr.push_back(a3[s]);
                    }
                    next_label = lab + 1;
                    q = r;
                }
                for (uint32_t l = next_label; l <= NL; ++l) row[l] = (uint32_t)nbr.size();
                if (nbr.size() >= 0xffffffffull) return GV_ERR_FORMAT;
            }
        }
    }
    // largest neighbour list (any partition): bounds every label segment
    uint32_t max_deg = 0;
    for (uint32_t p = 0; p < std::max<uint32_t>(nblocks, 1); ++p) {
        const uint32_t* lo = loff.data() + (uint64_t)p * nv1 * nl1;
        for (uint64_t v = 1; v < nv1; ++v)
            max_deg = std::max(max_deg, lo[v * nl1 + NL] - lo[v * nl1]);
    }
    // start tables per label
    std::vector<uint64_t> lvoff(nl1, 0);
    for (uint64_t v = 1; v < nv1; ++v)
        if (dl[v] != 0xffffffffu) lvoff[dl[v] + 1]++;
    for (uint32_t l = 1; l <= NL; ++l) lvoff[l] += lvoff[l - 1];
    std::vector<uint32_t> lverts(lvoff[NL]);
    {
        std::vector<uint64_t> pos(lvoff.begin(), lvoff.end() - 1);
        for (uint64_t v = 1; v < nv1; ++v)
            if (dl[v] != 0xffffffffu) lverts[pos[dl[v]]++] = (uint32_t)v;
    }
    // start planes (label-major offsets in lverts order)
    const uint64_t spn = lverts.size();
    std::vector<uint32_t> sp((uint64_t)std::max<uint32_t>(nblocks, 1) * nl1 * spn + 1, 0u);
    for (uint32_t p = 0; p < std::max<uint32_t>(nblocks, 1); ++p) {
        const uint32_t* lo = loff.data() + (uint64_t)p * nv1 * nl1;
        for (uint32_t l = 0; l < nl1; ++l) {
            uint32_t* plane = sp.data() + ((uint64_t)p * nl1 + l) * spn;
            for (uint64_t i = 0; i < spn; ++i) plane[i] = lo[(uint64_t)lverts[i] * nl1 + l];
        }
    }
    // label-pair edge tables: for partition p and labels a, b, the rows
    // (v, u) with label(v) = a, u a label-b neighbour of v, v then u
    // ascending (pe), and for a = b the rows with u < v (pel): the width-2
    // table of a first step from label a, unrestricted or restricted below
    // column 0
    const uint32_t np1 = std::max<uint32_t>(nblocks, 1);
    std::vector<uint64_t> pe_off((uint64_t)np1 * NL * NL + 1, 0), pel_off((uint64_t)np1 * NL + 1, 0);
    for (uint32_t p = 0; p < np1; ++p) {
        const uint32_t* lo = loff.data() + (uint64_t)p * nv1 * nl1;
        for (uint32_t a = 0; a < NL; ++a) {
            uint64_t low = 0;
            for (uint64_t i = lvoff[a]; i < lvoff[a + 1]; ++i) {
                const uint64_t v = lverts[i];
                const uint32_t b0 = lo[v * nl1 + a], b1 = lo[v * nl1 + a + 1];
                low += std::lower_bound(nbr.begin() + b0, nbr.begin() + b1, (uint32_t)v) - (nbr.begin() + b0);
            }
            pel_off[(uint64_t)p * NL + a + 1] = low;
            for (uint32_t b = 0; b < NL; ++b)
                pe_off[((uint64_t)p * NL + a) * NL + b + 1] = pairs[((uint64_t)p * NL + a) * NL + b];
        }
    }
    for (uint64_t i = 1; i < pe_off.size(); ++i) pe_off[i] += pe_off[i - 1];
    for (uint64_t i = 1; i < pel_off.size(); ++i) pel_off[i] += pel_off[i - 1];
    std::vector<uint32_t> pe(2 * pe_off.back() + 2, 0u), pel(2 * pel_off.back() + 2, 0u);
    for (uint32_t p = 0; p < np1; ++p) {
        const uint32_t* lo = loff.data() + (uint64_t)p * nv1 * nl1;
        for (uint32_t a = 0; a < NL; ++a) {
            uint64_t* cur = &pe_off[((uint64_t)p * NL + a) * NL];
            std::vector<uint64_t> at(cur, cur + NL);
            uint64_t lat = pel_off[(uint64_t)p * NL + a];
            for (uint64_t i = lvoff[a]; i < lvoff[a + 1]; ++i) {
                const uint32_t v = lverts[i];
                for (uint32_t b = 0; b < NL; ++b) {
                    const uint32_t s0 = lo[(uint64_t)v * nl1 + b], s1 = lo[(uint64_t)v * nl1 + b + 1];
                    for (uint32_t s = s0; s < s1; ++s) {
                        pe[2 * at[b]] = v;
                        pe[2 * at[b] + 1] = nbr[s];
                        ++at[b];
                        if (b == a && nbr[s] < v) {
                            pel[2 * lat] = v;
                            pel[2 * lat + 1] = nbr[s];
                            ++lat;
                        }
                    }
                }
            }
        }
    }
    gv_graph* g = new gv_graph();
    g->ctx = c;
    g->nv1 = nv1;
    g->sp_n = spn;
    g->pe_off = pe_off;
    g->pel_off = pel_off;
    g->nlabels = NL;
    g->nparts = nblocks;
    g->nnbr = nbr.size();
    g->max_deg = max_deg;
    g->label_values = vals;
    g->part_edge_label.assign(block_edge_labels, block_edge_labels + nblocks);
    g->lvoff = lvoff;
    g->pair_edges = pairs;
    if (nbr.empty()) nbr.push_back(0);
    if (lverts.empty()) lverts.push_back(0);
    int rc = GV_OK;
    auto up = [&](uint32_t** d, const void* h, uint64_t bytes) {
        if (rc != GV_OK) return;
        if (hipMalloc(d, bytes) != hipSuccess) { rc = GV_ERR_NOMEM; return; }
        if (hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = GV_ERR_HIP;
        g->dev_bytes += bytes;
    };
    HIPCHK(hipSetDevice(c->device));
    up(&g->d_loff, loff.data(), loff.size() * sizeof(uint32_t));
    up(&g->d_nbr, nbr.data(), nbr.size() * sizeof(uint32_t));
    up(&g->d_sp, sp.data(), sp.size() * sizeof(uint32_t));
    up(&g->d_lverts, lverts.data(), lverts.size() * sizeof(uint32_t));
    up(&g->d_pe, pe.data(), pe.size() * sizeof(uint32_t));
    up(&g->d_pel, pel.data(), pel.size() * sizeof(uint32_t));
    if (rc != GV_OK) {
        gv_graph_destroy(g);
        return rc;
    }
    *out = g;
    return GV_OK;
}

int gv_graph_destroy(gv_graph_t g) {
    if (!g) return GV_ERR_ARG;
    if (g->ctx) hipSetDevice(g->ctx->device);
    if (g->d_loff) hipFree(g->d_loff);
    if (g->d_nbr) hipFree(g->d_nbr);
    if (g->d_sp) hipFree(g->d_sp);
    if (g->d_lverts) hipFree(g->d_lverts);
    if (g->d_pe) hipFree(g->d_pe);
    if (g->d_pel) hipFree(g->d_pel);
    if (g->d_gof) hipFree(g->d_gof);
    delete g;
    return GV_OK;
}

int gv_graph_set_owner(gv_graph_t g, const uint32_t* graph_of, uint32_t ngraphs) {
    if (!g || !graph_of || ngraphs == 0) return GV_ERR_ARG;
    for (uint64_t v = 0; v < g->nv1; ++v)
        if (graph_of[v] >= ngraphs) return GV_ERR_ARG;
    if (g->d_gof) HIPCHK(hipFree(g->d_gof));
    g->d_gof = nullptr;
    if (hipMalloc(&g->d_gof, g->nv1 * sizeof(uint32_t)) != hipSuccess) return GV_ERR_NOMEM;
    HIPCHK(hipMemcpy(g->d_gof, graph_of, g->nv1 * sizeof(uint32_t), hipMemcpyHostToDevice));
    g->ngraphs = ngraphs;
    return GV_OK;
}

int gv_graph_owners(gv_graph_t g, uint32_t* ngraphs) {
    if (!g || !ngraphs) return GV_ERR_ARG;
    *ngraphs = g->d_gof ? g->ngraphs : 0;
    return GV_OK;
}

int gv_graph_label_index(gv_graph_t g, uint32_t label, uint32_t* dense) {
    if (!g || !dense) return GV_ERR_ARG;
    auto it = std::lower_bound(g->label_values.begin(), g->label_values.end(), label);
    if (it == g->label_values.end() || *it != label) return GV_ERR_ARG;
    *dense = (uint32_t)(it - g->label_values.begin());
    return GV_OK;
}

int gv_graph_part_index(gv_graph_t g, uint32_t edge_label, uint32_t* part) {
    if (!g || !part) return GV_ERR_ARG;
    for (uint32_t p = 0; p < g->nparts; ++p)
        if (g->part_edge_label[p] == edge_label) {
            *part = p;
            return GV_OK;
        }
    return GV_ERR_ARG;
}

int gv_graph_info(gv_graph_t g, uint64_t* nv1, uint32_t* nl, uint32_t* np, uint64_t* nnbr,
                  uint64_t* bytes) {
    if (!g) return GV_ERR_ARG;
    if (nv1) *nv1 = g->nv1;
    if (nl) *nl = g->nlabels;
    if (np) *np = g->nparts;
    if (nnbr) *nnbr = g->nnbr;
    if (bytes) *bytes = g->dev_bytes;
    return GV_OK;
}

int gv_graph_max_degree(gv_graph_t g, uint32_t* d) {
    if (!g || !d) return GV_ERR_ARG;
    *d = g->max_deg;
    return GV_OK;
}

int gv_graph_label_stats(gv_graph_t g, uint32_t part, uint64_t* vb, uint64_t* eb) {
    if (!g || part >= std::max<uint32_t>(g->nparts, 1)) return GV_ERR_ARG;
    const uint32_t NL = g->nlabels;
    if (vb)
        for (uint32_t l = 0; l < NL; ++l) vb[l] = g->lvoff[l + 1] - g->lvoff[l];
    if (eb)
        for (uint64_t i = 0; i < (uint64_t)NL * NL; ++i) eb[i] = g->pair_edges[(uint64_t)part * NL * NL + i];
    return GV_OK;
}

int gv_graph_start_rows(gv_graph_t g, uint32_t l, const uint32_t** rows, uint64_t* n) {
    if (!g || !rows || !n || l >= g->nlabels) return GV_ERR_ARG;
    *rows = g->d_lverts + g->lvoff[l];
    *n = g->lvoff[l + 1] - g->lvoff[l];
    return GV_OK;
}

int gv_frontier_create(gv_ctx_t c, uint64_t cap, gv_frontier_t* out) {
    if (!c || !out || cap == 0) return GV_ERR_ARG;
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    gv_frontier* f = new gv_frontier();
    f->ctx = c;
    f->cap = cap;
    bool ok = hipMalloc(&f->anc_beg, cap * 4) == hipSuccess &&
              hipMalloc(&f->cnt, cap * 4) == hipSuccess &&
              hipMalloc(&f->anc_i, cap) == hipSuccess &&
              hipMalloc(&f->csum, cap * 8) == hipSuccess;
    if (!ok) {
        gv_frontier_destroy(f);
        return GV_ERR_NOMEM;
    }
    *out = f;
    return GV_OK;
}

int gv_frontier_destroy(gv_frontier_t f) {
    if (!f) return GV_ERR_ARG;
    if (f->anc_beg) hipFree(f->anc_beg);
    if (f->cnt) hipFree(f->cnt);
    if (f->anc_i) hipFree(f->anc_i);
    if (f->csum) hipFree(f->csum);
    delete f;
    return GV_OK;
}

int gv_size_rows(gv_ctx_t c, gv_graph_t g, const gv_step_t* st, const uint32_t* rows,
                 uint32_t L, uint64_t nrows, gv_frontier_t f, uint64_t* total) {
    if (!c || !g || !f || !total || !rows || L == 0 || !valid_step(st, L, g)) return GV_ERR_ARG;
    c->clean = false;
    if (nrows > f->cap) return GV_ERR_CAPACITY;
    *total = 0;
    f->nrows = nrows;
    f->total = 0;
    if (nrows == 0) return GV_OK;
    SizeArgs a;
    memset(&a, 0, sizeof(a));
    a.g = dev_view(g);
    a.rows = rows;
    a.nrows = nrows;
    a.L = L;
    a.anc_beg = f->anc_beg;
    a.cnt = f->cnt;
    a.anc_i = f->anc_i;
    a.bytes = c->acc + kAccLine0 + 1;
    a.req = req_ptr(c, kKindSize);
    a.units = units_ptr(c, kKindSize);
    a.st = *st;
    a.start_pos = start_pos_of(g, rows, L);
    {
        int rc0 = ensure_tiles(c, nrows);
        if (rc0 != GV_OK) return rc0;
        a.tile_sums = c->tiles;
        ProfScope ps(c, kKindSize);
        if (c->prof_bytes)
            hipLaunchKernelGGL(k_size<true>, dim3(grid_for(nrows, kSizeTile, c->num_cu * 16)),
                               dim3(kBlock), 0, c->stream, a);
        else
            hipLaunchKernelGGL(k_size<false>, dim3(grid_for(nrows, kSizeTile, c->num_cu * 16)),
                               dim3(kBlock), 0, c->stream, a);
    }
    HIPCHK(hipGetLastError());
    int rc = scan_counts(c, f->cnt, nrows, nullptr, nullptr, f->csum);
    if (rc != GV_OK) return rc;
    HIPCHK(hipMemcpyAsync(total, f->csum + (nrows - 1), sizeof(uint64_t), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    f->total = *total;
    return GV_OK;
}

static int launch_walk(gv_ctx* c, WalkArgs& a, bool count) {
    const uint64_t span = a.c1 - a.c0;
    const uint64_t tiles = (span + kTile - 1) / kTile;
    const int blocks = grid_for(tiles, kWavesPerBlock, c->num_cu * 32);
    a.req = req_ptr(c, count ? kKindCount : kKindExt);
    a.units = units_ptr(c, count ? kKindCount : kKindExt);
    ProfScope ps(c, count ? kKindCount : kKindExt);
    if (count) {
        if (c->prof_bytes) hipLaunchKernelGGL((k_walk<true, true>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
        else hipLaunchKernelGGL((k_walk<true, false>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
    } else {
        if (c->prof_bytes) hipLaunchKernelGGL((k_walk<false, true>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
        else hipLaunchKernelGGL((k_walk<false, false>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
    }
    HIPCHK(hipGetLastError());
    return GV_OK;
}

static void walk_common(WalkArgs& a, gv_ctx* c, gv_graph* g, const gv_step_t* st,
                        const uint32_t* rows, uint32_t L, gv_frontier* f, uint64_t c0,
                        uint64_t c1) {
    memset(&a, 0, sizeof(a));
    a.g = dev_view(g);
    a.rows = rows;
    a.nrows = f->nrows;
    a.L = L;
    a.anc_beg = f->anc_beg;
    a.cnt = f->cnt;
    a.csum = f->csum;
    a.anc_i = f->anc_i;
    a.c0 = c0;
    a.c1 = c1;
    a.count = c->acc + kAccLine0;
    a.bytes = c->acc + kAccLine0 + 1;
    a.out_n = c->acc + 2;
    a.overflow = c->acc + 3;
    a.st = *st;
}

int gv_ext_emb(gv_ctx_t c, gv_graph_t g, const gv_step_t* st, const uint32_t* rows, uint32_t L,
               uint64_t nrows, gv_frontier_t f, uint64_t c0, uint64_t c1, uint32_t* out,
               uint64_t max_rows, uint64_t* written) {
    if (!c || !g || !f || !written || !valid_step(st, L, g) || c0 > c1) return GV_ERR_ARG;
    c->clean = false;
    if (nrows != f->nrows || c1 > f->total) return GV_ERR_ARG;
    *written = 0;
    if (c0 == c1 || nrows == 0) return GV_OK;
    if (!out || !rows) return GV_ERR_ARG;
    WalkArgs a;
    walk_common(a, c, g, st, rows, L, f, c0, c1);
    a.out = out;
    a.out_cap = max_rows;
    HIPCHK(hipMemsetAsync(c->acc + 2, 0, 2 * sizeof(unsigned long long), c->stream));
    int rc = launch_walk(c, a, false);
    if (rc != GV_OK) return rc;
    unsigned long long h[2];
    HIPCHK(hipMemcpyAsync(h, c->acc + 2, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h[1]) return GV_ERR_CAPACITY;
    *written = h[0];
    return GV_OK;
}

int gv_count_emb(gv_ctx_t c, gv_graph_t g, const gv_step_t* st, const gv_step_t* tail,
                 const uint32_t* rows, uint32_t L, uint64_t nrows, gv_frontier_t f, uint64_t c0,
                 uint64_t c1) {
    if (!c || !g || !f || !valid_step(st, L, g) || c0 > c1) return GV_ERR_ARG;
    c->clean = false;
    if (tail && !valid_tail(tail, L, g)) return GV_ERR_ARG;
    if (nrows != f->nrows || c1 > f->total) return GV_ERR_ARG;
    if (c0 == c1 || nrows == 0) return GV_OK;
    if (!rows) return GV_ERR_ARG;
    WalkArgs a;
    walk_common(a, c, g, st, rows, L, f, c0, c1);
    if (tail) {
        a.tail = *tail;
        a.has_tail = 1;
    }
    a.gbits = (g->d_gof && c->gbits && c->gbits_words * 32ull >= g->ngraphs) ? c->gbits : nullptr;
    return launch_walk(c, a, true);
}

int gv_count_rows_mode(const gv_step_t* st, const gv_step_t* tail, uint32_t L) {
    return count_rows_mode(st, tail, L);
}

static void leaves_args(LeavesArgs& a, gv_ctx* c, gv_graph* g, const gv_leaves_t* d,
                        const uint32_t* rows, uint64_t nrows) {
    memset(&a, 0, sizeof(a));
    a.g = dev_view(g);
    a.rows = rows;
    a.nrows = nrows;
    a.d = *d;
    a.count = c->acc + kAccLine0;
    a.bytes = c->acc + kAccLine0 + 1;
    a.req = req_ptr(c, kKindRows);
    a.units = units_ptr(c, kKindRows);
    a.gbits = (g->d_gof && c->gbits && c->gbits_words * 32ull >= g->ngraphs) ? c->gbits : nullptr;
}

int gv_count_leaves(gv_ctx_t c, gv_graph_t g, const gv_leaves_t* d, const uint32_t* rows,
                    uint64_t nrows) {
    if (!c || !g || !valid_leaves(d, g)) return GV_ERR_ARG;
    c->clean = false;
    if (nrows == 0) return GV_OK;
    if (!rows) return GV_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    LeavesArgs a;
    leaves_args(a, c, g, d, rows, nrows);
    {
        ProfScope ps(c, kKindRows);
        if (c->prof_bytes)
            hipLaunchKernelGGL(k_count_leaves<true>, dim3(grid_for(nrows, kBlock, c->num_cu * 16)),
                               dim3(kBlock), 0, c->stream, a);
        else
            hipLaunchKernelGGL(k_count_leaves<false>, dim3(grid_for(nrows, kBlock, c->num_cu * 16)),
                               dim3(kBlock), 0, c->stream, a);
    }
    HIPCHK(hipGetLastError());
    return GV_OK;
}

int gv_count_rows(gv_ctx_t c, gv_graph_t g, const gv_step_t* st, const gv_step_t* tail,
                  const uint32_t* rows, uint32_t L, uint64_t nrows) {
    if (!c || !g || !valid_step(st, L, g)) return GV_ERR_ARG;
    c->clean = false;
    if (tail && !valid_tail(tail, L, g)) return GV_ERR_ARG;
    const int mode = count_rows_mode(st, tail, L);
    if (mode < 0) return GV_ERR_ARG;
    if (nrows == 0) return GV_OK;
    if (!rows) return GV_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    RowsArgs a;
    memset(&a, 0, sizeof(a));
    a.g = dev_view(g);
    a.rows = rows;
    a.nrows = nrows;
    a.L = L;
    a.mode = mode;
    a.st = *st;
    if (tail) a.tail = *tail;
    a.count = c->acc + kAccLine0;
    a.bytes = c->acc + kAccLine0 + 1;
    a.req = req_ptr(c, kKindRows);
    a.units = units_ptr(c, kKindRows);
    a.gbits = (g->d_gof && c->gbits && c->gbits_words * 32ull >= g->ngraphs) ? c->gbits : nullptr;
    {
        ProfScope ps(c, kKindRows);
        if (c->prof_bytes)
            hipLaunchKernelGGL(k_count_rows<true>, dim3(grid_for(nrows, kBlock, c->num_cu * 16)),
                               dim3(kBlock), 0, c->stream, a);
        else
            hipLaunchKernelGGL(k_count_rows<false>, dim3(grid_for(nrows, kBlock, c->num_cu * 16)),
                               dim3(kBlock), 0, c->stream, a);
    }
    HIPCHK(hipGetLastError());
    return GV_OK;
}

int gv_mine_levels(gv_ctx_t c, gv_graph_t g, const gv_query_t* q, const uint32_t* start_rows,
                   uint64_t nstart, uint32_t* const* tables, const uint64_t* table_rows,
                   const gv_frontier_t* frontiers, int count_by_rows, const gv_leaves_t* leaves,
                   uint64_t* count, uint64_t* alg_bytes, uint32_t* bits, uint32_t nwords) {
    if (!c || !g || !q || !tables || !table_rows || !frontiers || q->k < 2 || q->k > GV_MAXK)
        return GV_ERR_ARG;
    if (leaves && (!valid_leaves(leaves, g) || leaves->k != q->k)) return GV_ERR_ARG;
    if (nstart && !start_rows) return GV_ERR_ARG;
    if (bits && nwords > (g->ngraphs + 31) / 32) return GV_ERR_ARG;
    const uint32_t k = q->k;
    // the widths that are materialised or sized, and their capacities
    uint32_t wlast = 1;   // width of the counted step's input
    for (uint32_t w = 1; w < k; ++w) {
        const gv_step_t& st = q->step[w];
        if (!valid_step(&st, w, g)) return GV_ERR_ARG;
        const bool last = w + 1 == k;
        const bool fused = w + 2 == k && q->step[w + 1].nback == 1;
        if (fused && !valid_tail(&q->step[w + 1], w, g)) return GV_ERR_ARG;
        const uint64_t cap = w == 1 ? nstart : table_rows[w];
        if (w >= 2 && (!tables[w] || cap == 0)) return GV_ERR_ARG;
        wlast = w;
        if (leaves && w == leaves->width) break;
        const bool by_rows = (last || fused) && count_by_rows &&
                             count_rows_mode(&st, fused ? &q->step[w + 1] : nullptr, w) >= 0;
        if (!by_rows && (!frontiers[w] || frontiers[w]->cap < cap)) return GV_ERR_ARG;
        if (last || fused) break;
        if (!tables[w + 1]) return GV_ERR_ARG;
    }
    HIPCHK(hipSetDevice(c->device));
    // the counted step in prefix-shared form (with the counting modes only,
    // as the leaf, chain and row counts)
    gv_step_t pre, child;
    uint32_t cut_last = 0;
    bool split = false;
    if (count_by_rows && !(leaves && wlast == leaves->width) && wlast >= 3) {
        const bool lastw = wlast + 1 == k;
        const bool fusedw = wlast + 2 == k && q->step[wlast + 1].nback == 1;
        const gv_step_t* tl = fusedw ? &q->step[wlast + 1] : nullptr;
        const bool chain = fusedw && g->max_deg > kChainMinDegree && chain_tail(q, wlast);
        const bool rowsm = count_rows_mode(&q->step[wlast], tl, wlast) >= 0;
        if ((lastw || fusedw) && !chain && !rowsm && frontiers[wlast - 1] && table_rows[wlast - 1] &&
            table_rows[wlast])
            split = prefix_split(q, wlast, pre, child, cut_last);
    }
    if (split) {
        const uint64_t icap = std::min<uint64_t>(table_rows[wlast] * wlast, 0xffffffffull);
        int rc = ensure_u32(c, c->par, c->par_cap, table_rows[wlast]);
        if (rc == GV_OK) rc = ensure_u32(c, c->ilen, c->ilen_cap, table_rows[wlast - 1]);
        if (rc == GV_OK) rc = ensure_u32(c, c->isect, c->isect_cap, icap);
        if (rc == GV_OK) rc = ensure_u32(c, c->plist, c->plist_cap, 2 * table_rows[wlast - 1]);
        if (rc == GV_OK && c->pmark_cap < table_rows[wlast - 1]) {
            rc = ensure_u32(c, c->pmark, c->pmark_cap, table_rows[wlast - 1]);
            if (rc == GV_OK) {
                HIPCHK(hipMemsetAsync(c->pmark, 0, c->pmark_cap * sizeof(uint32_t), c->stream));
                c->pmark_epoch = 0;
            }
        }
        if (rc != GV_OK) return rc;
        if (++c->pmark_epoch == 0) {   // marks of 2^32 queries ago could match
            HIPCHK(hipMemsetAsync(c->pmark, 0, c->pmark_cap * sizeof(uint32_t), c->stream));
            c->pmark_epoch = 1;
        }
    }
    int rc0 = ensure_gbits(c, g);
    if (rc0 != GV_OK) return rc0;
    const DevGraph dg = dev_view(g);
    unsigned long long* stop = c->acc + 3;
    uint32_t* gb = (g->d_gof && c->gbits && c->gbits_words * 32ull >= g->ngraphs) ? c->gbits : nullptr;
    if (!c->clean || gb) {
        const uint32_t words = g->ngraphs ? (g->ngraphs + 31) / 32 : 0;
        hipLaunchKernelGGL(k_zero_query, dim3(std::max<uint32_t>(1, std::min<uint32_t>((words + 255) / 256, 1024))),
                           dim3(256), 0, c->stream, c->acc, kAccWords, c->lvl, (uint32_t)(GV_MAXK + 1),
                           gb, gb ? words : 0u);
    }
    c->clean = false;
    const uint32_t* rows = start_rows;
    for (uint32_t w = 1; w <= wlast && nstart; ++w) {
        const gv_step_t& st = q->step[w];
        const uint64_t cap = w == 1 ? nstart : table_rows[w];
        const unsigned long long* ndev = w == 1 ? nullptr : c->lvl + w;
        const bool last = w + 1 == k;
        const bool fused = w + 2 == k && q->step[w + 1].nback == 1;
        const gv_step_t* tail = fused ? &q->step[w + 1] : nullptr;
        if (leaves && w == leaves->width) {
            LeavesArgs a;
            leaves_args(a, c, g, leaves, rows, cap);
            a.nrows_dev = ndev;
            a.stop = stop;
            ProfScope ps(c, kKindRows);
            if (c->prof_bytes)
                hipLaunchKernelGGL(k_count_leaves<true>, dim3(grid_for(cap, kBlock, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            else
                hipLaunchKernelGGL(k_count_leaves<false>, dim3(grid_for(cap, kBlock, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            HIPCHK(hipGetLastError());
            break;
        }
        // memoised chain weights pay when anchors repeat over many rows, i.e.
        // on graphs with hubs; on small-degree graphs the walk is cheaper
        if (fused && count_by_rows && g->max_deg > kChainMinDegree && chain_tail(q, w)) {
            // weights of every vertex of the anchor's label, then one pass
            // over the core rows
            const uint32_t acol = st.back_col[0];
            const uint32_t alab = acol == 0 ? q->label0 : q->step[acol].label;
            if (c->chainw_cap < g->nv1) {
                HIPCHK(hipStreamSynchronize(c->stream));
                if (c->chainw) HIPCHK(hipFree(c->chainw));
                if (c->chainneed) HIPCHK(hipFree(c->chainneed));
                c->chainw = nullptr;
                c->chainneed = nullptr;
                c->chainw_cap = 0;
                if (hipMalloc(&c->chainw, g->nv1 * sizeof(unsigned long long)) != hipSuccess ||
                    hipMalloc(&c->chainneed, g->nv1 * sizeof(uint32_t)) != hipSuccess)
                    return GV_ERR_NOMEM;
                HIPCHK(hipMemsetAsync(c->chainneed, 0, g->nv1 * sizeof(uint32_t), c->stream));
                c->chainw_cap = g->nv1;
                c->chain_epoch = 0;
            }
            if (++c->chain_epoch == 0) {   // marks of 2^32 queries ago could match
                HIPCHK(hipStreamSynchronize(c->stream));
                HIPCHK(hipMemsetAsync(c->chainneed, 0, c->chainw_cap * sizeof(uint32_t), c->stream));
                c->chain_epoch = 1;
            }
            ChainArgs ca;
            ca.g = dg;
            ca.verts = g->d_lverts + g->lvoff[alab];
            ca.nverts = g->lvoff[alab + 1] - g->lvoff[alab];
            ca.l1 = st.label;
            ca.p1 = st.back_part[0];
            ca.l2 = q->step[w + 1].label;
            ca.p2 = q->step[w + 1].back_part[0];
            ca.w = c->chainw;
            ca.need = c->chainneed;
            ca.epoch = c->chain_epoch;
            ChainCountArgs a;
            memset(&a, 0, sizeof(a));
            a.g = dg;
            a.rows = rows;
            a.nrows = cap;
            a.nrows_dev = ndev;
            a.stop = stop;
            a.L = w;
            a.a = acol;
            a.l1 = st.label;
            a.p1 = st.back_part[0];
            a.w = c->chainw;
            a.count = c->acc + kAccLine0;
            a.bytes = c->acc + kAccLine0 + 1;
            a.req = req_ptr(c, kKindChain);
            a.units = units_ptr(c, kKindChain);
            a.gbits = gb;
            ProfScope ps(c, kKindChain);
            hipLaunchKernelGGL(k_chain_mark, dim3(grid_for(cap, kBlock, c->num_cu * 16)), dim3(kBlock), 0,
                               c->stream, rows, cap, ndev, stop, w, acol, c->chainneed, c->chain_epoch);
            if (ca.nverts)
                hipLaunchKernelGGL(k_chain_weights,
                                   dim3(grid_for(ca.nverts * (64 / kChainVertsPerWave), kBlock,
                                                 c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, ca);
            if (c->prof_bytes)
                hipLaunchKernelGGL(k_count_chain<true>, dim3(grid_for(cap, kBlock, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            else
                hipLaunchKernelGGL(k_count_chain<false>, dim3(grid_for(cap, kBlock, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            HIPCHK(hipGetLastError());
            break;
        }
        if ((last || fused) && count_by_rows && count_rows_mode(&st, tail, w) >= 0) {
            RowsArgs a;
            memset(&a, 0, sizeof(a));
            a.g = dg;
            a.rows = rows;
            a.nrows = cap;
            a.nrows_dev = ndev;
            a.stop = stop;
            a.L = w;
            a.mode = count_rows_mode(&st, tail, w);
            a.st = st;
            if (tail) a.tail = *tail;
            a.count = c->acc + kAccLine0;
            a.bytes = c->acc + kAccLine0 + 1;
            a.req = req_ptr(c, kKindRows);
            a.units = units_ptr(c, kKindRows);
            a.gbits = gb;
            ProfScope ps(c, kKindRows);
            if (c->prof_bytes)
                hipLaunchKernelGGL(k_count_rows<true>, dim3(grid_for(cap, kBlock, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            else
                hipLaunchKernelGGL(k_count_rows<false>, dim3(grid_for(cap, kBlock, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            HIPCHK(hipGetLastError());
            break;
        }
        gv_frontier* f = frontiers[w];
        const bool split_here = split && w == wlast;
        if (split_here) {
            // parents: sized by the prefix step (no algorithmic bytes), then
            // intersected; the rows then sized by their parents' intersections
            gv_frontier* fp = frontiers[w - 1];
            const uint64_t pcap = table_rows[w - 1];
            const unsigned long long* pndev = c->lvl + (w - 1);
            const uint32_t* prows = tables[w - 1];
            SizeArgs sa;
            memset(&sa, 0, sizeof(sa));
            sa.g = dg;
            sa.rows = prows;
            sa.nrows = pcap;
            sa.nrows_dev = pndev;
            sa.stop = stop;
            sa.L = w - 1;
            sa.anc_beg = fp->anc_beg;
            sa.cnt = fp->cnt;
            sa.anc_i = fp->anc_i;
            sa.st = pre;
            sa.start_pos = -1;
            int rc1 = ensure_tiles(c, std::max<uint64_t>(pcap, cap));
            if (rc1 != GV_OK) return rc1;
            sa.tile_sums = c->tiles;
            {
                ProfScope ps(c, kKindIsect);
                hipLaunchKernelGGL(k_size<false>, dim3(grid_for(pcap, kSizeTile, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, sa);
                HIPCHK(hipGetLastError());
            }
            rc1 = scan_counts(c, fp->cnt, pcap, pndev, stop, fp->csum);
            if (rc1 != GV_OK) return rc1;
            IsectArgs ia;
            memset(&ia, 0, sizeof(ia));
            ia.g = dg;
            ia.rows = prows;
            ia.nrows = pcap;
            ia.nrows_dev = pndev;
            ia.stop = stop;
            ia.L = w - 1;
            ia.anc_beg = fp->anc_beg;
            ia.cnt = fp->cnt;
            ia.anc_i = fp->anc_i;
            ia.csum = fp->csum;
            ia.out = c->isect;
            ia.out_cap = c->isect_cap;
            ia.len = c->ilen;
            ia.overflow = stop;
            ia.st = pre;
            // only the parents of some row are intersected
            ProfScope ps(c, kKindIsect);
            const uint64_t half = c->plist_cap / 2;
            hipLaunchKernelGGL(k_mark_parents, dim3(grid_for(cap, kBlock, c->num_cu * 4)), dim3(kBlock), 0,
                               c->stream, c->par, cap, ndev, stop, fp->cnt, c->pmark, c->pmark_epoch,
                               c->plist, half, c->acc + kAccParSmall, c->acc + kAccParBig);
            ia.list = c->plist;
            ia.nlist = c->acc + kAccParSmall;
            hipLaunchKernelGGL(k_isect, dim3(c->num_cu * 16), dim3(kBlock), 0, c->stream, ia);
            ia.list = c->plist + half;
            ia.nlist = c->acc + kAccParBig;
            hipLaunchKernelGGL(k_isect_big, dim3(c->num_cu * 8), dim3(kBlock), 0, c->stream, ia);
            ChildSizeArgs ca;
            memset(&ca, 0, sizeof(ca));
            ca.g = dg;
            ca.rows = rows;
            ca.nrows = cap;
            ca.nrows_dev = ndev;
            ca.stop = stop;
            ca.L = w;
            ca.par = c->par;
            ca.pcnt = fp->cnt;
            ca.pcsum = fp->csum;
            ca.ilen = c->ilen;
            ca.isect = c->isect;
            ca.anc_beg = f->anc_beg;
            ca.cnt = f->cnt;
            ca.anc_i = f->anc_i;
            ca.tile_sums = c->tiles;
            ca.bytes = c->acc + kAccLine0 + 1;
            ca.st = st;
            ca.cut_last = cut_last;
            hipLaunchKernelGGL(k_child_size, dim3(grid_for(cap, kSizeTile, c->num_cu * 16)), dim3(kBlock),
                               0, c->stream, ca);
            HIPCHK(hipGetLastError());
        } else {
            SizeArgs a;
            memset(&a, 0, sizeof(a));
            a.g = dg;
            a.rows = rows;
            a.nrows = cap;
            a.nrows_dev = ndev;
            a.stop = stop;
            a.L = w;
            a.anc_beg = f->anc_beg;
            a.cnt = f->cnt;
            a.anc_i = f->anc_i;
            a.bytes = c->acc + kAccLine0 + 1;
            a.req = req_ptr(c, kKindSize);
            a.units = units_ptr(c, kKindSize);
            a.st = st;
            a.start_pos = start_pos_of(g, rows, w);
            int rc0 = ensure_tiles(c, cap);
            if (rc0 != GV_OK) return rc0;
            a.tile_sums = c->tiles;
            ProfScope ps(c, kKindSize);
            if (c->prof_bytes)
                hipLaunchKernelGGL(k_size<true>, dim3(grid_for(cap, kSizeTile, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            else
                hipLaunchKernelGGL(k_size<false>, dim3(grid_for(cap, kSizeTile, c->num_cu * 16)),
                                   dim3(kBlock), 0, c->stream, a);
            HIPCHK(hipGetLastError());
        }
        int rc = scan_counts(c, f->cnt, cap, ndev, stop, f->csum);
        if (rc != GV_OK) return rc;
        // the frontier's host-side extents are unknown here
        f->nrows = 0;
        f->total = 0;
        WalkArgs a;
        memset(&a, 0, sizeof(a));
        a.g = dg;
        a.rows = rows;
        a.nrows = cap;
        a.nrows_dev = ndev;   // null at level 1: rows = nstart, candidates from the scan
        a.stop = stop;
        a.L = w;
        a.anc_beg = f->anc_beg;
        a.cnt = f->cnt;
        a.csum = f->csum;
        a.anc_i = f->anc_i;
        a.count = c->acc + kAccLine0;
        a.bytes = c->acc + kAccLine0 + 1;
        a.overflow = stop;
        a.st = split_here ? child : st;
        a.anc_base = split_here ? c->isect : nullptr;
        if (split && w + 1 == wlast) a.par = c->par;
        const bool counting = last || fused;
        if (counting) {
            if (tail) {
                a.tail = *tail;
                a.has_tail = 1;
            }
            a.gbits = gb;
        } else {
            a.out = tables[w + 1];
            a.out_cap = table_rows[w + 1];
            a.out_n = c->lvl + w + 1;
        }
        a.req = req_ptr(c, counting ? kKindCount : kKindExt);
        a.units = units_ptr(c, counting ? kKindCount : kKindExt);
        {
            ProfScope ps(c, counting ? kKindCount : kKindExt);
            const dim3 grid(c->num_cu * 32);
            if (counting && c->prof_bytes)
                hipLaunchKernelGGL((k_walk<true, true, kCountIPT>), grid, dim3(kBlock), 0, c->stream, a);
            else if (counting)
                hipLaunchKernelGGL((k_walk<true, false, kCountIPT>), grid, dim3(kBlock), 0, c->stream, a);
            else if (c->prof_bytes && a.st.nback == 1)
                hipLaunchKernelGGL((k_walk<false, true, kSingleIPT>), grid, dim3(kBlock), 0, c->stream, a);
            else if (c->prof_bytes)
                hipLaunchKernelGGL((k_walk<false, true>), grid, dim3(kBlock), 0, c->stream, a);
            else if (a.st.nback == 1)
                hipLaunchKernelGGL((k_walk<false, false, kSingleIPT>), grid, dim3(kBlock), 0, c->stream, a);
            else
                hipLaunchKernelGGL((k_walk<false, false>), grid, dim3(kBlock), 0, c->stream, a);
            HIPCHK(hipGetLastError());
        }
        if (counting) break;
        rows = tables[w + 1];
    }
    hipLaunchKernelGGL(k_fold_acc, dim3(1), dim3(64), 0, c->stream, c->acc, c->dres, kAccWords,
                       c->lvl, (uint32_t)(GV_MAXK + 1));
    HIPCHK(hipGetLastError());
    c->clean = true;
    if (bits && nwords)
        HIPCHK(hipMemcpyAsync(bits, c->gbits, nwords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const volatile unsigned long long* h = c->hres;
    if (h[3]) return GV_ERR_CAPACITY;
    if (count) *count = h[0];
    if (alg_bytes) *alg_bytes = h[1];
    return GV_OK;
}

int gv_counters_reset(gv_ctx_t c, gv_graph_t g) {
    if (!c) return GV_ERR_ARG;
    c->clean = false;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemsetAsync(c->acc, 0, kAccWords * sizeof(unsigned long long), c->stream));
    if (g && g->ngraphs) {
        int rc = ensure_gbits(c, g);
        if (rc != GV_OK) return rc;
        HIPCHK(hipMemsetAsync(c->gbits, 0, c->gbits_words * sizeof(uint32_t), c->stream));
    }
    return GV_OK;
}

int gv_counters_read(gv_ctx_t c, uint64_t* count, uint64_t* bytes, uint32_t* bits,
                     uint32_t nwords) {
    if (!c) return GV_ERR_ARG;
    c->clean = false;
    unsigned long long h[2];
    hipLaunchKernelGGL(k_fold_acc, dim3(1), dim3(64), 0, c->stream, c->acc, nullptr, 0u, nullptr, 0u);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, c->acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    if (bits && nwords) {
        if (nwords > c->gbits_words) return GV_ERR_ARG;
        HIPCHK(hipMemcpyAsync(bits, c->gbits, nwords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (count) *count = h[0];
    if (bytes) *bytes = h[1];
    return GV_OK;
}

int gv_copy_partial_emb(gv_ctx_t c, const uint32_t* src, uint32_t* dst, uint64_t n) {
    if (!c || (n && (!src || !dst))) return GV_ERR_ARG;
    if (n == 0 || src == dst) return GV_OK;
    if (dst + n <= src || dst >= src + n) {   // disjoint
        HIPCHK(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToDevice, c->stream));
        return GV_OK;
    }
    if (dst < src) {
        // dst overlaps the head of src: copy in chunks from the front, each
        // no longer than the shift, so no chunk overwrites what a later one
        // still has to read
        const uint64_t shift = (uint64_t)(src - dst);
        for (uint64_t beg = 0; beg < n; beg += shift) {
            const uint64_t len = std::min(shift, n - beg);
            HIPCHK(hipMemcpyAsync(dst + beg, src + beg, len * 4, hipMemcpyDeviceToDevice, c->stream));
        }
        return GV_OK;
    }
    // dst overlaps the tail of src: copy in chunks from the end, each chunk
    // no longer than the shift so no chunk reads what an earlier one wrote
    const uint64_t shift = (uint64_t)(dst - src);
    uint64_t end = n;
    while (end > 0) {
        uint64_t len = std::min(shift, end);
        uint64_t beg = end - len;
        hipLaunchKernelGGL(k_copy_back, dim3(grid_for(len, kBlock, 4096)), dim3(kBlock), 0,
                           c->stream, src + beg, dst + beg, len);
        HIPCHK(hipGetLastError());
        end = beg;
    }
    return GV_OK;
}

int gv_profile_enable(gv_ctx_t c, int on) {
    if (!c) return GV_ERR_ARG;
    if (on && !c->preq) {
        if (hipMalloc(&c->preq, kKinds * sizeof(unsigned long long)) != hipSuccess) return GV_ERR_NOMEM;
        HIPCHK(hipMemsetAsync(c->preq, 0, kKinds * sizeof(unsigned long long), c->stream));
    }
    if (on && !c->punits) {
        // [kind]: work units, [kKinds + kind]: algorithmic bytes B
        if (hipMalloc(&c->punits, 2 * kKinds * sizeof(unsigned long long)) != hipSuccess)
            return GV_ERR_NOMEM;
        HIPCHK(hipMemsetAsync(c->punits, 0, 2 * kKinds * sizeof(unsigned long long), c->stream));
    }
    // 1: HIP events around every launch; 2: also requested bytes and units,
    // accumulated by the kernels' counting variants
    c->prof = on != 0;
    c->prof_bytes = on >= 2;
    return GV_OK;
}

int gv_profile_reset(gv_ctx_t c) {
    if (!c) return GV_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& r : c->recs) {
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    if (c->preq) HIPCHK(hipMemsetAsync(c->preq, 0, kKinds * sizeof(unsigned long long), c->stream));
    if (c->punits) HIPCHK(hipMemsetAsync(c->punits, 0, 2 * kKinds * sizeof(unsigned long long), c->stream));
    return GV_OK;
}

int gv_profile_units(gv_ctx_t c, int kind, uint64_t* units) {
    if (!c || !units || kind < 0 || kind >= kKinds) return GV_ERR_ARG;
    *units = 0;
    if (c->punits) {
        unsigned long long h[2 * kKinds];
        HIPCHK(hipMemcpyAsync(h, c->punits, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *units = h[kind];
    }
    return GV_OK;
}

int gv_profile_alg_bytes(gv_ctx_t c, int kind, uint64_t* bytes) {
    if (!c || !bytes || kind < 0 || kind >= kKinds) return GV_ERR_ARG;
    *bytes = 0;
    if (c->punits) {
        unsigned long long h[2 * kKinds];
        HIPCHK(hipMemcpyAsync(h, c->punits, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *bytes = h[kKinds + kind];
    }
    return GV_OK;
}

int gv_profile_read(gv_ctx_t c, int kind, uint64_t* launches, double* total_ms,
                    uint64_t* req_bytes) {
    if (!c || kind < 0 || kind >= kKinds) return GV_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    uint64_t n = 0;
    double ms = 0;
    for (auto& r : c->recs) {
        if (r.kind != kind) continue;
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        ms += t;
        ++n;
    }
    uint64_t rq = 0;
    if (c->preq) {
        unsigned long long h[kKinds];
        HIPCHK(hipMemcpy(h, c->preq, sizeof(h), hipMemcpyDeviceToHost));
        rq = h[kind];
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (req_bytes) *req_bytes = rq;
    return GV_OK;
}

int gv_profile_records(gv_ctx_t c, int* kinds, float* ms, uint64_t cap, uint64_t* n) {
    if (!c || !n) return GV_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    *n = c->recs.size();
    for (uint64_t i = 0; i < c->recs.size() && i < cap; ++i) {
        if (kinds) kinds[i] = c->recs[i].kind;
        if (ms) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, c->recs[i].a, c->recs[i].b));
            ms[i] = t;
        }
    }
    return GV_OK;
}

int gv_mark(gv_ctx_t c, int slot) {
    if (!c || slot < 0 || slot >= 8) return GV_ERR_ARG;
    HIPCHK(hipEventRecord(c->ev[slot], c->stream));
    return GV_OK;
}

int gv_elapsed_ms(gv_ctx_t c, int a, int b, float* ms) {
    if (!c || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return GV_ERR_ARG;
    HIPCHK(hipEventSynchronize(c->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return GV_OK;
}

}  // extern "C"
