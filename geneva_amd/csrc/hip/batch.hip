// Batched execution: many queries in flight, one captured graph per iteration.
//
// Device state per slot is split in a descriptor (the compiled query, written
// at admission) and a compact dynamic record (polled by the host).  One
// iteration is the kernel sequence
//   k_b_prep   one workgroup: which slots run, their mode (extend / count /
//              count with a single-edge tail), their offset in the iteration's
//              row space, deferral when the row space is full;
//   k_b_size   per row of every running slot: shortest back-neighbour list,
//              restriction cut, candidate count;
//              each sizing tile's candidate prefix (within the tile) and total;
//   k_b_tscan  one workgroup: the tiles' prefix, each slot's candidate range
//              and walk tiles;
//   k_b_walk   load-balanced walk over the candidates of all slots at once
//              (a wave's tile may straddle slots; per-slot groups are formed
//              with ballots), writing each slot's next table or counting;
//   k_b_post   one workgroup: advance, finish or flag overflow.
// Every kernel reads its extent from device memory, so the sequence is
// captured once and replayed without host synchronisation.
#include "common.h"

#include <optional>

namespace {

// Rows per thread of the engine's sizing passes.  Its launches size many
// slots at once and are large, so four rows per thread with their loads
// interleaved beat two here (the plain path's single-query levels use two).
constexpr int kBSizeRows = 4;
constexpr int kBSizeThreads = 256;
constexpr uint64_t kBSizeTile = (uint64_t)kBSizeThreads * kBSizeRows;
static_assert(kBSizeTile == (1u << kPTileShift), "a sizing tile is a tile of the candidate prefix");
constexpr uint32_t kMaxSlots = 1024;

struct BDesc {               // written by the host at admission
    gv_query_t q;
    uint32_t start_off;      // offset of label0's start rows in lverts
    uint32_t start_n;
    uint32_t has_leaves;     // leaves: counted per row at width leaves.width
    gv_leaves_t leaves;
    // a first step that only lists the label-pair edges (one back edge to
    // column 0, at most the restriction below it) starts the slot at width 2
    // on the graph's label-pair edge table: its rows, never written
    const uint32_t* init_rows;
    // candidates per walk tile of this slot (64 x its candidates per lane, at
    // most the kernel's); 0 = the batch's
    uint32_t tile;
    // prefix sharing: a follower's first steps equal those of the slot
    // `leader` (admitted with serial leader_serial) up to the table of width
    // fork_width.  It waits until the leader's input table has that width and
    // then extends the leader's table in the same iteration as the leader,
    // taking over the leader's counters; kNoSlot = none
    uint32_t leader;
    uint32_t fork_width;
    uint64_t leader_serial;
    // where the slot starts (width, rows, bytes of a pair start): a follower
    // that cannot share restarts from here
    uint32_t init_width;
    uint32_t piece;          // one of several pieces of the query's first table: never shares
    uint64_t init_n;
    uint64_t init_bytes;
};

constexpr uint32_t kNoSlot = 0xffffffffu;

struct BState {              // dynamic, polled by the host (mirrors gv_slot_info_t)
    uint64_t tag;
    uint32_t state;
    uint32_t width;          // columns of the current input table
    uint64_t rows;           // rows of the current input table
    unsigned long long count;
    unsigned long long bytes;
    // internal
    uint32_t sel;            // half holding the input table
    uint32_t mode;           // this iteration: 0 idle, 1 extend, 2 count, 3 count + tail,
                             // 4 leaf groups (counted while sizing)
    int32_t rmode;           // count modes: count_rows_mode of the step, or -1 (walk)
    unsigned long long out_rows;
    unsigned long long overflow;
    uint64_t row_end;        // end of this slot's rows in the iteration row space
    uint64_t cand_base;      // first candidate of this slot in the iteration's space
    uint64_t cand_end;
    uint64_t tile_end;       // end of this slot's tiles (tiles never straddle slots)
    uint64_t rtile_end;      // end of this slot's sizing tiles (kBSizeTile rows each)
    uint32_t defer;          // iterations in a row this slot waited for row space
    uint32_t wait;           // follower still waiting for its leader's table
    uint64_t serial;         // admission number (never 0 for an admitted slot)
    uint32_t src;            // slot and half holding this iteration's input table
    uint32_t src_sel;
    // slots sizing the same input table this iteration are sized together:
    // the group's head (szl) runs each of its tiles for the others too, so
    // their row and neighbour-list lines are read while still in the cache;
    // sznx = the next slot of the group, kNoSlot after the last
    uint32_t szl;
    uint32_t sznx;
};

// A slot that waited this many iterations in a row for row space is placed
// before the others, so a large request in a high slot is not passed over
// forever while lower slots keep refilling.
constexpr uint32_t kAgeLimit = 4;

struct BArgs {
    DevGraph g;
    const uint32_t* lverts;
    const BDesc* desc;
    BState* st;
    uint32_t nslots;
    uint32_t* half0;
    uint32_t* half1;
    uint64_t slot_uints;
    // per row of the iteration's row space (slots start at sizing tiles, so
    // rows between a slot's last row and the next tile are unused)
    uint32_t* anc_beg;
    uint32_t* cnt;              // candidate prefix within the row's sizing tile (inclusive)
    uint8_t* anc_i;
    uint64_t row_cap;
    uint64_t* tiles;            // per sizing tile: its total, then the tiles' exclusive prefix
    // walk tiles: the extension slots' first, then the counted slots'; slot
    // wslot[i] owns the tiles up to wend[i] (ascending)
    uint64_t* wend;
    uint16_t* wslot;
    unsigned long long* glob;   // [0] rows, [1] candidates, [2] tiles, [3] sizing tiles,
                                // [4] shared tables (forks), [5] followers restarted,
                                // [6] extension tiles, [7] rows sized with another slot's
                                // tiles (cumulative)
    uint64_t tile;              // candidates per walk tile (64 * IPT)
    uint32_t fuse;              // size slots with the same input table together
    uint32_t* gbits;            // nslots * gwords
    uint32_t gwords;
    // spread count / bytes accumulators: kBLines lines (kAccStride words) per
    // slot; blocks add to line blockIdx % kBLines of their slot and k_b_post
    // folds the lines into the slot's state
    unsigned long long* sacc;
    // algorithmic bytes B charged per kernel (nullable; profiling), spread
    // like sacc over kBLines lines: line[0] k_b_size, line[1] k_b_walk
    // (rows read, back lists read, rows counted there), [1] k_b_walk
    unsigned long long* kb;
};

constexpr uint32_t kBLines = 16;
constexpr uint64_t kKbBytes = 8ull * kBLines * kAccStride;

__device__ __forceinline__ unsigned long long* slot_line(const BArgs& a, uint32_t s) {
    return a.sacc + ((uint64_t)s * kBLines + blockIdx.x % kBLines) * kAccStride;
}

// the per-kernel B line of this block (one address for all blocks made the
// profiled kernels wait on that one address)
__device__ __forceinline__ unsigned long long* kb_line(const BArgs& a) {
    return a.kb + (uint64_t)(blockIdx.x % kBLines) * kAccStride;
}

__device__ __forceinline__ const uint32_t* slot_rows(const BArgs& a, uint32_t s, const BState& S,
                                                     const BDesc& D) {
    if (S.width == 1) return a.lverts + D.start_off;
    if (S.width == 2 && D.init_rows) return D.init_rows;
    return (S.src_sel ? a.half1 : a.half0) + (uint64_t)S.src * a.slot_uints;
}

__device__ __forceinline__ uint32_t* slot_out(const BArgs& a, uint32_t s, const BState& S) {
    return (S.sel ? a.half0 : a.half1) + (uint64_t)s * a.slot_uints;
}

// slot owning global row r: first s with row_end[s] > r (row_end in LDS)
__device__ __forceinline__ uint32_t slot_of(const uint64_t* row_end, uint32_t n, uint64_t r) {
    uint32_t lo = 0;
    while (n > 0) {
        uint32_t h = n >> 1;
        if (row_end[lo + h] <= r) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

// Block-wide exclusive scan over up to 1024 values held one per thread.
__device__ uint64_t block_scan_1024(uint64_t v, uint64_t* wsum, uint64_t& total) {
    const unsigned lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
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
    for (unsigned i = 0; i < nw; ++i) {
        if (i < w) off += wsum[i];
        total += wsum[i];
    }
    __syncthreads();
    return off + x - v;
}

// Sizing tiles of a slot with `rows` rows, rounded up to a multiple of the
// GPU's XCD count.  Workgroup b runs on XCD b mod kXcds and the sizing grid
// is a multiple of kXcds, so tile k of every slot lands on the same XCD.
constexpr uint64_t kXcds = 8;
__device__ __forceinline__ uint64_t slot_tiles(uint64_t rows) {
    return ((rows + kBSizeTile - 1) / kBSizeTile + kXcds - 1) / kXcds * kXcds;
}

// Mode of a slot whose input table has width w: 1 extend, 2 count the last
// step, 3 count the last two (single-edge tail), 4 leaf groups; rmode = the
// per-row counting mode of a counted step, or -1 (walk).
__device__ uint32_t slot_mode(const BDesc& D, uint32_t w, int32_t& rmode) {
    const bool last = w + 1 == D.q.k;
    const bool tail = !last && w + 2 == D.q.k && D.q.step[w + 1].nback == 1;
    uint32_t mode = last ? 2 : (tail ? 3 : 1);
    if (D.has_leaves && w == D.leaves.width) mode = 4;
    rmode = mode == 1 || mode == 4
                ? -1
                : count_rows_mode(&D.q.step[w], mode == 3 ? &D.q.step[w + 1] : nullptr, w);
    return mode;
}

// A running slot's mode and rows this iteration; finishes a slot whose table
// is empty or complete and flags one whose table can never fit an iteration.
__device__ uint32_t plan_running(const BArgs& a, BState& S, const BDesc& D, uint64_t& rows,
                                 int32_t& rmode) {
    rows = 0;
    rmode = -1;
    const uint64_t n = S.width == 1 ? D.start_n : S.rows;
    if (n == 0 || S.width >= D.q.k) {
        S.state = GV_SLOT_DONE;
        return 0;
    }
    if (n > a.row_cap) {
        S.state = GV_SLOT_OVERFLOW;
        return 0;
    }
    rows = n;
    return slot_mode(D, S.width, rmode);
}

__global__ __launch_bounds__(1024) void k_b_prep(BArgs a) {
    __shared__ uint64_t wsum[16];
    __shared__ uint32_t nfollow[kMaxSlots];   // followers sharing each leader's table now
    __shared__ uint8_t ran[kMaxSlots];
    __shared__ uint64_t gkey[kMaxSlots];      // input table address (0: not running)
    __shared__ uint64_t grows[kMaxSlots];
    __shared__ uint32_t gwid[kMaxSlots];
    __shared__ uint32_t hslot[2 * kMaxSlots];  // sizing groups: bucket -> head slot
    __shared__ uint32_t lnext[kMaxSlots];      // a head's last linked member
    const uint32_t s = threadIdx.x;
    const bool in = s < a.nslots;
    nfollow[s] = 0;
    uint64_t rows = 0;
    uint32_t mode = 0;
    int32_t rmode = -1;
    bool waiting = false;
    if (in) {
        BState& S = a.st[s];
        S.out_rows = 0;
        if (S.state == GV_SLOT_RUNNING) {
            if (S.wait) waiting = true;
            else mode = plan_running(a, S, a.desc[s], rows, rmode);
        }
    }
    __syncthreads();   // every leader's state of this iteration is settled
    // a waiting follower: share its leader's table if the leader holds it now;
    // finish as the leader did if its table ran empty or overflowed before
    // that width (the follower's identical table would have too); restart on
    // its own if the leader went past that width or its slot changed hands
    uint32_t leader = kNoSlot;
    if (waiting) {
        BState& S = a.st[s];
        const BDesc& D = a.desc[s];
        const BState& P = a.st[D.leader];
        const uint32_t d = D.fork_width;
        int act = 2;   // 0 wait, 1 share, 2 restart, 3 empty, 4 overflow
        if (P.serial == D.leader_serial) {
            if (P.state == GV_SLOT_RUNNING) act = P.wait || P.width < d ? 0 : (P.width == d ? 1 : 2);
            else if (P.state == GV_SLOT_DONE) act = P.width <= d ? 3 : 2;
            else if (P.state == GV_SLOT_OVERFLOW) act = P.width <= d ? 4 : 2;
        }
        if (act == 1) {
            leader = D.leader;
            atomicAdd(&nfollow[leader], 1u);
        } else if (act == 3) {
            S.wait = 0;
            S.bytes = P.bytes;
            S.state = GV_SLOT_DONE;
        } else if (act == 4) {
            S.wait = 0;
            S.state = GV_SLOT_OVERFLOW;
        } else if (act == 2) {
            S.wait = 0;
            S.width = D.init_width;
            S.rows = D.init_n;
            S.bytes = D.init_bytes;
            S.sel = 0;
            mode = plan_running(a, S, D, rows, rmode);
            atomicAdd(&a.glob[5], 1ull);
        }
    }
    __syncthreads();
    // a leader and the followers sharing its table run as one unit of
    // (1 + followers) x its rows; a unit larger than row_cap never fits, so
    // its followers restart on their own
    bool share = false;
    if (leader != kNoSlot) {
        const uint64_t pn = a.st[leader].rows;
        if (pn * (1 + nfollow[leader]) > a.row_cap) {
            BState& S = a.st[s];
            const BDesc& D = a.desc[s];
            S.wait = 0;
            S.width = D.init_width;
            S.rows = D.init_n;
            S.bytes = D.init_bytes;
            S.sel = 0;
            mode = plan_running(a, S, D, rows, rmode);
            atomicAdd(&a.glob[5], 1ull);
            leader = kNoSlot;
        } else {
            share = true;
        }
    }
    uint64_t unit = share ? 0 : rows;
    if (in && !share && nfollow[s] && rows * (1 + nfollow[s]) <= a.row_cap) unit = rows * (1 + nfollow[s]);
    // which units run: those that waited kAgeLimit iterations first, then the
    // rest, each group in slot order while the rows fit row_cap; the others
    // wait (and age)
    const bool aged = in && !share && a.st[s].defer >= kAgeLimit;
    uint64_t tot, tot_a;
    const uint64_t ra = aged ? unit : 0;
    const uint64_t ex_a = block_scan_1024(ra, wsum, tot);
    const bool run_a = aged && ex_a + ra <= a.row_cap;
    block_scan_1024(run_a ? unit : 0, wsum, tot_a);
    const uint64_t rb = aged ? 0 : unit;
    const uint64_t ex_b = block_scan_1024(rb, wsum, tot);
    bool run = run_a || (!aged && tot_a + ex_b + rb <= a.row_cap);
    ran[s] = run;
    __syncthreads();
    uint32_t src = s, src_sel = in ? a.st[s].sel : 0;
    if (share) {
        // runs with its leader or keeps waiting
        run = ran[leader];
        if (run) {
            BState& S = a.st[s];
            const BState& P = a.st[leader];
            const uint32_t d = a.desc[s].fork_width;
            S.wait = 0;
            S.width = d;
            S.rows = P.rows;
            S.bytes = P.bytes;
            S.sel = 0;
            rows = P.rows;
            mode = slot_mode(a.desc[s], d, rmode);
            src = leader;
            src_sel = P.sel;
            atomicAdd(&a.glob[4], 1ull);
        }
    } else if (in && mode != 0) {
        a.st[s].defer = run ? 0 : a.st[s].defer + 1;
    }
    if (!run) {
        rows = 0;
        mode = 0;
    }
    // sizing groups: running slots whose input tables are one table (same
    // rows at the same address: slots starting from one start list or
    // label-pair range, a leader and the followers reading its table)
    uint64_t key = 0;
    uint32_t kw = 0;
    if (in && mode != 0 && a.fuse) {
        const BState& S = a.st[s];
        const BDesc& D = a.desc[s];
        kw = S.width;
        const uint32_t* p = kw == 1                    ? a.lverts + D.start_off
                            : kw == 2 && D.init_rows ? D.init_rows
                                                     : (src_sel ? a.half1 : a.half0) + (uint64_t)src * a.slot_uints;
        key = (uint64_t)p;
    }
    gkey[s] = key;
    grows[s] = rows;
    gwid[s] = kw;
    hslot[s] = kNoSlot;
    hslot[s + kMaxSlots] = kNoSlot;
    lnext[s] = kNoSlot;
    __syncthreads();
    // the group's head is the slot that claims the table's bucket of an
    // open-addressing set in LDS (twice as many buckets as slots, so a probe
    // always ends); the others link themselves after it
    uint32_t szl = s, linked = kNoSlot;
    if (key) {
        uint32_t h = (uint32_t)((key >> 2) * 0x9E3779B97F4A7C15ull >> 40 ^ rows * 0x85EBCA6Bu ^ kw) &
                     (2 * kMaxSlots - 1);
        for (;;) {
            const uint32_t old = atomicCAS(&hslot[h], kNoSlot, s);
            if (old == kNoSlot) break;
            if (gkey[old] == key && grows[old] == rows && gwid[old] == kw) {
                szl = old;
                break;
            }
            h = (h + 1) & (2 * kMaxSlots - 1);
        }
        if (szl != s) {
            linked = atomicExch(&lnext[szl], s);
            atomicAdd(&a.glob[7], (unsigned long long)rows);
        }
    }
    __syncthreads();
    const uint32_t sznx = szl == s ? lnext[s] : linked;
    // the running slots' row ranges, in slot order, each from the start of a
    // sizing tile (kBSizeTile rows of one slot), so that row r lies in tile
    // r / kBSizeTile of the iteration's candidate prefix; a slot's tiles are
    // a multiple of kXcds (see slot_tiles)
    uint64_t total_rows;
    block_scan_1024(rows, wsum, total_rows);
    const uint64_t nrt = slot_tiles(rows);
    uint64_t rtot;
    const uint64_t rex = block_scan_1024(nrt, wsum, rtot);
    if (in) {
        BState& S = a.st[s];
        S.mode = mode;
        S.rmode = rmode;
        S.src = src;
        S.src_sel = src_sel;
        S.szl = szl;
        S.sznx = sznx;
        S.row_end = rex * kBSizeTile + rows;
        S.rtile_end = rex + nrt;
    }
    if (s == 0) {
        a.glob[0] = total_rows;
        a.glob[1] = 0;
        a.glob[3] = rtot;
    }
}

// A thread's R consecutive rows of a per-row array in one store (R words
// from an R-aligned row).
template <int R>
__device__ __forceinline__ void store_rows(uint32_t* p, const uint32_t (&v)[R]) {
    if constexpr (R == 4) {
        *reinterpret_cast<uint4*>(p) = make_uint4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 2) {
        *reinterpret_cast<uint2*>(p) = make_uint2(v[0], v[1]);
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) p[k] = v[k];
    }
}
template <int R>
__device__ __forceinline__ void store_bytes(uint8_t* p, const uint32_t (&v)[R]) {
    if constexpr (R == 4) {
        *reinterpret_cast<uint32_t*>(p) =
            (v[0] & 0xffu) | (v[1] & 0xffu) << 8 | (v[2] & 0xffu) << 16 | (v[3] & 0xffu) << 24;
    } else if constexpr (R == 2) {
        *reinterpret_cast<uint16_t*>(p) = (uint16_t)((v[0] & 0xffu) | (v[1] & 0xffu) << 8);
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) p[k] = (uint8_t)v[k];
    }
}

__device__ __forceinline__ Prefix two_level(const BArgs& a) {
    Prefix p;
    p.tpre = a.tiles;
    p.loc = a.cnt;
    return p;
}

// Sizing: one block per sizing tile (kBSizeTile rows of one slot, kBSizeRows
// consecutive rows per thread).  Besides each row's anchor segment it writes
// the candidate prefix of the tile's rows (two-level form: loc = inclusive
// within the tile, in `cnt`) and the tile's total in tiles[tile].
__global__ __launch_bounds__(kBSizeThreads) void k_b_size(BArgs a) {
    __shared__ uint64_t rtile_end[kMaxSlots];
    __shared__ uint16_t szl[kMaxSlots], sznx[kMaxSlots];
    __shared__ uint64_t wsum[kBSizeThreads / 64];
    for (uint32_t i = threadIdx.x; i < a.nslots; i += blockDim.x) {
        rtile_end[i] = a.st[i].rtile_end;
        szl[i] = (uint16_t)a.st[i].szl;
        sznx[i] = (uint16_t)a.st[i].sznx;
    }
    __syncthreads();
    const uint64_t ntiles = a.glob[3];
    const unsigned lane = __lane_id();
    uint32_t cur = 0xffffffffu, last_gid = 0xffffffffu;
    uint64_t acc = 0, cacc = 0, rq = 0;
    auto flush = [&]() {
        if (cur == 0xffffffffu) return;
        const uint64_t w = wave_sum_u64(acc), c = wave_sum_u64(cacc);
        unsigned long long* line = slot_line(a, cur);
        if (lane == 0 && c) atomicAdd(line, (unsigned long long)c);
        if (lane == 0 && w) atomicAdd(line + 1, (unsigned long long)w);
        if (a.kb && lane == 0 && w) atomicAdd(kb_line(a), (unsigned long long)w);
        acc = cacc = 0;
    };
    for (uint64_t RT0 = blockIdx.x; RT0 < ntiles; RT0 += gridDim.x) {
        // every row of this tile belongs to one slot: its step is read once
        const uint32_t s0 = __builtin_amdgcn_readfirstlane(slot_of(rtile_end, a.nslots, RT0));
        // a slot of a sizing group is sized with the group's head slot:
        // the same tile of each slot of the group, one after the other
        if (szl[s0] != s0) continue;
        const uint64_t back = rtile_end[s0] - RT0;
        for (uint32_t s = s0; s != (kNoSlot & 0xffffu); s = __builtin_amdgcn_readfirstlane(sznx[s])) {
        const uint64_t RT = rtile_end[s] - back;
        if (s != cur) {
            flush();
            cur = s;
            last_gid = 0xffffffffu;
        }
        const BState& S = a.st[s];
        const BDesc& D = a.desc[s];
        const uint32_t L = S.width;
        const uint64_t rows = L == 1 ? D.start_n : S.rows;
        const uint64_t tile0 = (RT - (rtile_end[s] - slot_tiles(rows))) * kBSizeTile;
        const gv_step_t& st = D.q.step[L];
        const uint32_t* trows = slot_rows(a, s, S, D);
        uint64_t li[kBSizeRows];
        bool ok[kBSizeRows];
#pragma unroll
        for (int k = 0; k < kBSizeRows; ++k) {
            li[k] = tile0 + (uint64_t)threadIdx.x * kBSizeRows + k;
            ok[k] = li[k] < rows;
        }
        uint32_t beg[kBSizeRows], len[kBSizeRows], cnt[kBSizeRows], ai[kBSizeRows];
#pragma unroll
        for (int k = 0; k < kBSizeRows; ++k) beg[k] = cnt[k] = ai[k] = 0;
        if (S.mode == 4) {
            // leaf groups on these core rows: counted here, no candidates (a
            // row's own read is part of the leaf bytes)
#pragma unroll
            for (int k = 0; k < kBSizeRows; ++k) {
                if (!ok[k]) continue;
                const uint32_t* row = trows + li[k] * L;
                uint64_t by = 0;
                const uint64_t n = leaf_row_count(a.g, D.leaves, row, by, rq);
                cacc += n;
                acc += by;
                if (n && a.g.gof) {
                    const uint32_t gid = a.g.gof[row[0]];
                    if (gid != last_gid) {
                        atomicOr(&a.gbits[(uint64_t)s * a.gwords + (gid >> 5)], 1u << (gid & 31));
                        last_gid = gid;
                    }
                }
            }
        } else {
            uint64_t by = 0;
            size_rows<kBSizeRows>(a.g, st, trows, L, li, ok, beg, len, cnt, ai, by, rq,
                                 L == 1 ? (int64_t)D.start_off : (int64_t)-1);
            acc += by;
            if (S.rmode >= 0) {
                // counted here, per row: no candidates are left for the walk
                const gv_step_t& tl = D.q.step[L + 1 < GV_MAXK ? L + 1 : GV_MAXK - 1];
#pragma unroll
                for (int k = 0; k < kBSizeRows; ++k) {
                    if (!ok[k]) continue;
                    uint64_t tby = 0;
                    const uint64_t n = row_count(a.g, st, tl, S.rmode, L, trows + li[k] * L, beg[k],
                                                 beg[k] + len[k], tby, rq);
                    acc += tby;
                    cnt[k] = 0;
                    cacc += n;
                    if (n && a.g.gof) {
                        const uint32_t gid = a.g.gof[trows[li[k] * L]];
                        if (gid != last_gid) {
                            atomicOr(&a.gbits[(uint64_t)s * a.gwords + (gid >> 5)], 1u << (gid & 31));
                            last_gid = gid;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kBSizeRows; ++k)
                if (!ok[k]) cnt[k] = 0;
        }
        // the tile's candidate prefix in row order (below 2^32: a row has at
        // most max_deg < 2^32 / kBSizeTile candidates)
        uint32_t incl[kBSizeRows];
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < kBSizeRows; ++k) {
            run += cnt[k];
            incl[k] = run;
        }
        uint64_t btot;
        const uint32_t bex = (uint32_t)block_scan_1024(run, wsum, btot);
        const uint64_t r0 = RT * kBSizeTile + (uint64_t)threadIdx.x * kBSizeRows;
#pragma unroll
        for (int k = 0; k < kBSizeRows; ++k) incl[k] += bex;
        store_rows(a.cnt + r0, incl);
        store_rows(a.anc_beg + r0, beg);
        store_bytes(a.anc_i + r0, ai);
        if (threadIdx.x == 0) a.tiles[RT] = btot;
        }
    }
    flush();
}

constexpr int kTScanPer = 16;

// The candidate prefix of the tiles (exclusive, in place in `tiles`), then
// each running slot's candidate range and its walk tiles (tiles never
// straddle slots).  One workgroup.
__global__ __launch_bounds__(1024) void k_b_tscan(BArgs a) {
    __shared__ uint64_t wsum[16];
    const uint64_t nt = a.glob[3];
    // rounds of kTScanPer consecutive tiles per thread (16-byte loads and
    // stores, all of a round's loads in flight at once), carried over rounds
    uint64_t cand = 0;
    for (uint64_t r0 = 0; r0 < nt; r0 += (uint64_t)kTScanPer * blockDim.x) {
        const uint64_t b0 = r0 + (uint64_t)threadIdx.x * kTScanPer;
        uint64_t v[kTScanPer];
        const bool full = b0 + kTScanPer <= nt;
        if (full) {
            const ulonglong2* p = reinterpret_cast<const ulonglong2*>(a.tiles + b0);
#pragma unroll
            for (int k = 0; k < kTScanPer / 2; ++k) {
                const ulonglong2 x = p[k];
                v[2 * k] = x.x;
                v[2 * k + 1] = x.y;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kTScanPer; ++k) v[k] = b0 + k < nt ? a.tiles[b0 + k] : 0;
        }
        uint64_t sum = 0;
#pragma unroll
        for (int k = 0; k < kTScanPer; ++k) sum += v[k];
        uint64_t tot;
        uint64_t ex = cand + block_scan_1024(sum, wsum, tot);
#pragma unroll
        for (int k = 0; k < kTScanPer; ++k) {
            const uint64_t x = v[k];
            v[k] = ex;
            ex += x;
        }
        if (full) {
            ulonglong2* p = reinterpret_cast<ulonglong2*>(a.tiles + b0);
#pragma unroll
            for (int k = 0; k < kTScanPer / 2; ++k) p[k] = make_ulonglong2(v[2 * k], v[2 * k + 1]);
        } else {
#pragma unroll
            for (int k = 0; k < kTScanPer; ++k)
                if (b0 + k < nt) a.tiles[b0 + k] = v[k];
        }
        cand += tot;
    }
    __syncthreads();
    const uint32_t s = threadIdx.x;
    uint64_t nwt = 0, cb = 0, ce = 0;
    if (s < a.nslots) {
        const BState& S = a.st[s];
        if (S.mode != 0) {
            const Prefix pf = two_level(a);
            const uint64_t rows = S.width == 1 ? a.desc[s].start_n : S.rows;
            const uint64_t re = S.row_end, rb = re - rows;
            cb = pf.begin(rb);
            ce = pf.end(re - 1);
            const uint64_t tl = a.desc[s].tile ? a.desc[s].tile : a.tile;
            nwt = (ce - cb + tl - 1) / tl;
        }
    }
    // the extension slots' tiles first, then the counted ones' (two kernels)
    const bool in = s < a.nslots;
    const bool ext = in && a.st[s].mode == 1;
    uint64_t t0, t1, n0, n1;
    const uint64_t wex0 = block_scan_1024(ext ? nwt : 0, wsum, t0);
    const uint64_t wex1 = block_scan_1024(in && !ext ? nwt : 0, wsum, t1);
    const uint64_t rk0 = block_scan_1024(ext ? 1 : 0, wsum, n0);
    const uint64_t rk1 = block_scan_1024(in && !ext ? 1 : 0, wsum, n1);
    if (in) {
        const uint64_t end = ext ? wex0 + nwt : t0 + wex1 + nwt;
        const uint64_t rank = ext ? rk0 : n0 + rk1;
        a.st[s].cand_base = cb;
        a.st[s].cand_end = ce;
        a.st[s].tile_end = end;
        a.wend[rank] = end;
        a.wslot[rank] = (uint16_t)s;
    }
    if (s == 0) {
        a.glob[1] = cand;
        a.glob[2] = t0 + t1;
        a.glob[6] = t0;
    }
}

// ---- walk -------------------------------------------------------------------

// Rows of a multi-row walk tile whose candidate starts a wave stages in LDS:
// twice the candidates of a tile, since a batch's tables hold many rows with
// no candidate (the rows of workload 3's first iteration average 0.43
// candidates), so its tiles span more rows than candidates.
template <int IPT>
constexpr uint32_t kBLxRows = 128u * IPT;

// The walk over the extension slots' tiles (kCount false) or the counted
// slots' (true): two kernels, so that each holds only its own path's
// registers (one more wave per SIMD each).
template <int IPT, bool kCount>
__global__ __launch_bounds__(kBlock) void k_b_walk(BArgs a) {
    __shared__ uint64_t wend[kMaxSlots];
    __shared__ uint16_t wslot[kMaxSlots];
    __shared__ uint32_t wl[kWavesPerBlock][wave_lds_words<IPT, kBLxRows<IPT>>()];
    for (uint32_t i = threadIdx.x; i < a.nslots; i += blockDim.x) {
        wend[i] = a.wend[i];
        wslot[i] = a.wslot[i];
    }
    __syncthreads();
    const uint64_t tlo = kCount ? a.glob[6] : 0, ntiles = kCount ? a.glob[2] : a.glob[6];
    const unsigned lane = __lane_id(), wib = threadIdx.x >> 6;
    uint32_t cur = 0xffffffffu;
    TileAcc acc;
    __shared__ uint64_t resv[3 * kWavesPerBlock];
    const BlockReserve reserve{resv};
    auto flush = [&]() {
        if (cur == 0xffffffffu) return;
        const uint64_t c = wave_sum_u64(acc.count), b = wave_sum_u64(acc.bytes);
        unsigned long long* line = slot_line(a, cur);
        if (lane == 0) {
            if (c) atomicAdd(line, (unsigned long long)c);
            if (b) atomicAdd(line + 1, (unsigned long long)b);
            if (a.kb && b) atomicAdd(kb_line(a) + 1, (unsigned long long)b);
        }
        acc = TileAcc();
    };
    // block rounds: wave wib of block b takes tile (round * blocks + b) * 4 + wib,
    // and every wave of the block runs every round (the reservation is shared)
    for (uint64_t T0 = tlo + (uint64_t)blockIdx.x * kWavesPerBlock; T0 < ntiles;
         T0 += (uint64_t)gridDim.x * kWavesPerBlock) {
        const uint64_t T = T0 + wib;
        if (T >= ntiles) {
            reserve(nullptr, 0);
            continue;
        }
        // the tile's slot, made wave-uniform so its descriptor is read once
        const uint32_t pos = __builtin_amdgcn_readfirstlane(slot_of(wend, a.nslots, T));
        const uint32_t s = wslot[pos];
        if (s != cur) {
            flush();
            cur = s;
        }
        const BState& S = a.st[s];
        const BDesc& D = a.desc[s];
        const uint32_t L = S.width;
        const uint64_t rows = L == 1 ? D.start_n : S.rows;
        const uint64_t tl = D.tile ? D.tile : a.tile;
        const uint64_t ntile_s = (S.cand_end - S.cand_base + tl - 1) / tl;
        const uint64_t k = T - (wend[pos] - ntile_s);
        const uint64_t t0 = S.cand_base + k * tl;
        const uint64_t t1 = t0 + tl < S.cand_end ? t0 + tl : S.cand_end;
        TileIn in;
        in.rows = slot_rows(a, s, S, D);
        in.anc = a.g.nbr;
        in.row_base = S.row_end - rows;
        in.L = L;
        in.pf = two_level(a);
        in.anc_beg = a.anc_beg;
        in.anc_i = a.anc_i;
        in.row_lo = S.row_end - rows;
        in.row_hi = S.row_end;
        uint32_t* gb = a.g.gof ? a.gbits + (uint64_t)s * a.gwords : nullptr;
        const gv_step_t& st = D.q.step[L];
        if constexpr (!kCount) {
            TileOut o{slot_out(a, s, S), a.slot_uints / (L + 1), &a.st[s].out_rows, &a.st[s].overflow};
            walk_tile<false, IPT, kBLxRows<IPT>>(a.g, st, st, false, in, t0, t1, o, wl[wib], wl[wib], wl[wib] + kBLxRows<IPT> + 1, gb,
                                  acc, reserve);
        } else {
            TileOut o{nullptr, 0, nullptr, nullptr};
            const gv_step_t& tl = D.q.step[L + 1 < GV_MAXK ? L + 1 : GV_MAXK - 1];
            walk_tile<true, IPT, kBLxRows<IPT>>(a.g, st, tl, S.mode == 3, in, t0, t1, o, nullptr, wl[wib], wl[wib] + kBLxRows<IPT> + 1, gb,
                                 acc, reserve);
            reserve(nullptr, 0);
        }
    }
    flush();
}

__global__ __launch_bounds__(1024) void k_b_post(BArgs a) {
    const uint32_t s = threadIdx.x;
    if (s >= a.nslots) return;
    BState& S = a.st[s];
    unsigned long long* lines = a.sacc + (uint64_t)s * kBLines * kAccStride;
    uint64_t cnt = 0, by = 0;
    for (uint32_t i = 0; i < kBLines; ++i) {
        cnt += lines[i * kAccStride];
        by += lines[i * kAccStride + 1];
        lines[i * kAccStride] = 0;
        lines[i * kAccStride + 1] = 0;
    }
    S.count += cnt;
    S.bytes += by;
    if (S.mode == 0) return;
    if (S.mode >= 2) {   // counted (count, count + tail, leaf groups)
        S.state = GV_SLOT_DONE;
        S.width = a.desc[s].q.k;
    } else if (S.overflow) {
        S.state = GV_SLOT_OVERFLOW;
    } else {
        S.rows = S.out_rows;
        S.width += 1;
        S.sel ^= 1u;
    }
    S.mode = 0;
}

// Slots named to a kernel by value (up to kMaxSlots), so no host staging
// buffer is read by a copy in flight.
struct SlotList {
    uint32_t n;
    uint16_t s[kMaxSlots];
};
constexpr uint32_t kStateWords = sizeof(BState) / 4;
static_assert(sizeof(BState) % 4 == 0, "BState is copied word by word");

unsigned list_blocks(uint64_t work) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((work + 255) / 256, 1024)); }

// Admitted slots: their staged states copied in and their graph bitsets
// cleared.
__global__ void k_b_admit(BState* st, const BState* stage, uint32_t* gbits, uint32_t gwords, SlotList L) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t0; i < (uint64_t)L.n * gwords; i += stride)
        gbits[(uint64_t)L.s[i / gwords] * gwords + i % gwords] = 0;
    for (uint64_t i = t0; i < (uint64_t)L.n * kStateWords; i += stride) {
        const uint32_t s = L.s[i / kStateWords];
        reinterpret_cast<uint32_t*>(st + s)[i % kStateWords] =
            reinterpret_cast<const uint32_t*>(stage + s)[i % kStateWords];
    }
}

// Released slots: their graph bitsets gathered into out[k * gwords ..] (k =
// position in the list; out nullable) and their states zeroed (free).
__global__ void k_b_release(BState* st, const uint32_t* gbits, uint32_t gwords, uint32_t* out, SlotList L) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (out)
        for (uint64_t i = t0; i < (uint64_t)L.n * gwords; i += stride)
            out[i] = gbits[(uint64_t)L.s[i / gwords] * gwords + i % gwords];
    for (uint64_t i = t0; i < (uint64_t)L.n * kStateWords; i += stride)
        reinterpret_cast<uint32_t*>(st + L.s[i / kStateWords])[i % kStateWords] = 0u;
}

}  // namespace

struct gv_batch {
    gv_ctx* ctx = nullptr;
    gv_graph* g = nullptr;
    uint32_t nslots = 0;
    uint64_t slot_bytes = 0;
    uint64_t row_cap = 0;
    uint64_t row_cap_alloc = 0;         // rows the sizing arrays were made for
    gv_tune_t tune{4, 16};
    BDesc* d_desc = nullptr;
    BState* d_st = nullptr;
    uint32_t* d_half[2] = {nullptr, nullptr};
    uint32_t* d_anc_beg = nullptr;
    uint32_t* d_cnt = nullptr;
    uint8_t* d_anc_i = nullptr;
    uint64_t* d_tiles = nullptr;
    unsigned long long* d_glob = nullptr;
    uint32_t* d_gbits = nullptr;
    unsigned long long* d_sacc = nullptr;
    uint32_t gwords = 0;
    std::vector<uint8_t> used;          // host view: slot occupied
    // Admission stages a slot's descriptor and state in pinned memory and
    // marks it dirty; the next run or poll copies each run of consecutive
    // dirty slots with one copy per array.  A staging entry is rewritten only
    // after a synchronisation that followed its copy (epochs below).
    std::vector<uint8_t> dirty;
    std::vector<uint64_t> flush_epoch;  // epoch of the slot's last staging copy
    uint64_t epoch = 1;                 // advanced by every synchronisation here
    BDesc* h_desc = nullptr;            // pinned staging, one entry per slot
    BState* h_st = nullptr;
    uint64_t* d_wend = nullptr;
    uint16_t* d_wslot = nullptr;
    BState* d_st_stage = nullptr;       // staged states on their way to d_st (flush)
    uint32_t* d_bits_out = nullptr;     // released slots' graph bitsets, gathered
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long* d_kb = nullptr;   // per-kernel B (profiling), kBLines lines
    bool kb_on = false;
    bool pair_start = true;               // first steps from the label-pair edge tables
    bool share = true;                    // followers share a leader's table (prefix sharing)
    bool fuse = true;                     // slots with one input table are sized together
    uint64_t next_serial = 1;
    std::vector<uint32_t> hwidth;         // host view: width at admission or last poll
    std::vector<uint32_t> hstate;
    // rows sized and candidates walked by each profiled iteration
    std::vector<uint64_t> hist;
};

namespace {

BArgs bargs(gv_batch* b) {
    BArgs a;
    memset(&a, 0, sizeof(a));
    a.g = dev_view(b->g);
    a.lverts = b->g->d_lverts;
    a.desc = b->d_desc;
    a.st = b->d_st;
    a.nslots = b->nslots;
    a.half0 = b->d_half[0];
    a.half1 = b->d_half[1];
    a.slot_uints = b->slot_bytes / 4;
    a.anc_beg = b->d_anc_beg;
    a.cnt = b->d_cnt;
    a.anc_i = b->d_anc_i;
    a.row_cap = b->row_cap;
    a.tiles = b->d_tiles;
    a.glob = b->d_glob;
    a.gbits = b->d_gbits;
    a.gwords = b->gwords;
    a.sacc = b->d_sacc;
    a.tile = 64ull * b->tune.ipt;
    a.fuse = b->fuse ? 1u : 0u;
    a.kb = b->kb_on ? b->d_kb : nullptr;
    a.wend = b->d_wend;
    a.wslot = b->d_wslot;
    return a;
}

template <int IPT>
void launch_walk(const BArgs& a, int blocks, hipStream_t st) {
    hipLaunchKernelGGL((k_b_walk<IPT, false>), dim3(blocks), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL((k_b_walk<IPT, true>), dim3(blocks), dim3(kBlock), 0, st, a);
}

// The kernels of one iteration on the context's stream; with `prof` each is
// bracketed by HIP events of its kind (prepare / tiles / retire: batch;
// sizing; scan; walk: extension).
void launch_iteration(gv_batch* b, bool prof) {
    gv_ctx* c = b->ctx;
    BArgs a = bargs(b);
    const int size_blocks = c->num_cu * (2048 / kBSizeThreads);
    const int walk_blocks = c->num_cu * (int)b->tune.walk_blocks_per_cu;
    // events only when profiling (never inside a capture)
    std::optional<ProfScope> ps;
    auto timed = [&](int kind) {
        ps.reset();
        if (prof) ps.emplace(c, kind);
    };
    timed(kKindBatch);
    hipLaunchKernelGGL(k_b_prep, dim3(1), dim3(1024), 0, c->stream, a);
    timed(kKindSize);
    hipLaunchKernelGGL(k_b_size, dim3(size_blocks), dim3(kBSizeThreads), 0, c->stream, a);
    timed(kKindScan);
    hipLaunchKernelGGL(k_b_tscan, dim3(1), dim3(1024), 0, c->stream, a);
    timed(kKindExt);
    switch (b->tune.ipt) {
        case 1: launch_walk<1>(a, walk_blocks, c->stream); break;
        case 2: launch_walk<2>(a, walk_blocks, c->stream); break;
        case 8: launch_walk<8>(a, walk_blocks, c->stream); break;
        default: launch_walk<4>(a, walk_blocks, c->stream); break;
    }
    timed(kKindBatch);
    hipLaunchKernelGGL(k_b_post, dim3(1), dim3(1024), 0, c->stream, a);
    ps.reset();
}

int capture(gv_batch* b) {
    gv_ctx* c = b->ctx;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    launch_iteration(b, false);
    hipError_t e1 = hipStreamEndCapture(c->stream, &b->graph);
    if (e1 != hipSuccess) {
        g_last_error = std::string("capture: ") + hipGetErrorString(e1);
        return GV_ERR_HIP;
    }
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return GV_OK;
}

int sync(gv_batch* b) {
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    ++b->epoch;
    return GV_OK;
}

// Copy the staged descriptors and states of the dirty slots and clear their
// graph bitsets: the staging entries of the slots lo..hi spanning them go over
// in two copies (descriptors straight to the device table: the device never
// writes it, and a slot's staged descriptor is what it last received; states
// to a device staging table), then one kernel applies the dirty slots' states.
int flush(gv_batch* b) {
    gv_ctx* c = b->ctx;
    SlotList L;
    L.n = 0;
    uint32_t lo = b->nslots, hi = 0;
    for (uint32_t s = 0; s < b->nslots; ++s) {
        if (!b->dirty[s]) continue;
        L.s[L.n++] = (uint16_t)s;
        lo = std::min(lo, s);
        hi = s;
    }
    if (L.n == 0) return GV_OK;
    const uint32_t n = hi - lo + 1;
    HIPCHK(hipMemcpyAsync(b->d_desc + lo, b->h_desc + lo, sizeof(BDesc) * n, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(b->d_st_stage + lo, b->h_st + lo, sizeof(BState) * n, hipMemcpyHostToDevice,
                          c->stream));
    hipLaunchKernelGGL(k_b_admit, dim3(list_blocks((uint64_t)L.n * std::max(b->gwords, kStateWords))),
                       dim3(256), 0, c->stream, b->d_st, b->d_st_stage, b->d_gbits, b->gwords, L);
    HIPCHK(hipGetLastError());
    for (uint32_t s = lo; s <= hi; ++s) {
        b->dirty[s] = 0;
        b->flush_epoch[s] = b->epoch;
    }
    return GV_OK;
}

// Free the listed slots on the device (states zeroed); with bits, their
// graph bitsets are copied to bits[k * gwords ..] (synchronises).
int release_list(gv_batch* b, const SlotList& L, uint32_t* bits) {
    if (L.n == 0) return GV_OK;
    hipLaunchKernelGGL(k_b_release, dim3(list_blocks((uint64_t)L.n * std::max(b->gwords, kStateWords))),
                       dim3(256), 0, b->ctx->stream, b->d_st, b->d_gbits, b->gwords,
                       bits ? b->d_bits_out : nullptr, L);
    HIPCHK(hipGetLastError());
    if (!bits) return GV_OK;
    HIPCHK(hipMemcpyAsync(bits, b->d_bits_out, 4ull * b->gwords * L.n, hipMemcpyDeviceToHost,
                          b->ctx->stream));
    return sync(b);
}

bool same_step(const gv_step_t& x, const gv_step_t& y) {
    if (x.label != y.label || x.col != y.col || x.nback != y.nback || x.restrict_col != y.restrict_col ||
        x.ncoll != y.ncoll)
        return false;
    for (int i = 0; i < x.nback && i < GV_MAXK; ++i)
        if (x.back_col[i] != y.back_col[i] || x.back_part[i] != y.back_part[i]) return false;
    for (int i = 0; i < x.ncoll && i < GV_MAXK; ++i)
        if (x.coll_col[i] != y.coll_col[i]) return false;
    return true;
}

// Widest table a slot extends from before it counts: its last step, a
// single-edge tail counted with it, or its leaf groups' core.
uint32_t counted_width(const BDesc& d) {
    uint32_t w = d.q.k - 1;
    if (d.q.k >= 3 && d.q.step[d.q.k - 1].nback == 1) w = d.q.k - 2;
    if (d.has_leaves) w = std::min(w, d.leaves.width);
    return w;
}

// Widest table a follower can take from a leader: as counted_width, except
// that a single-edge tail counted with the step before it can instead be
// counted on its own from the table of width k - 1 (the same count and
// algorithmic bytes).
uint32_t follow_width(const BDesc& d) {
    uint32_t w = d.q.k - 1;
    if (d.has_leaves) w = std::min(w, d.leaves.width);
    return w;
}

// The live slot whose table of the widest width d >= 3 equals this query's
// own table of width d (same start label and steps 1 .. d-1; the leader
// extends from width d); 0 = none.  A leader that itself waits to share a
// table at least as wide is replaced by its own leader.
uint32_t pick_leader(gv_batch* b, uint32_t self, const BDesc& d, uint32_t& leader) {
    const uint32_t cwf = follow_width(d);
    uint32_t best = 0;
    leader = kNoSlot;
    for (uint32_t p = 0; p < b->nslots; ++p) {
        if (p == self || !b->used[p] || b->hstate[p] != GV_SLOT_RUNNING) continue;
        const BDesc& e = b->h_desc[p];
        if (e.piece) continue;
        if (e.q.label0 != d.q.label0) continue;
        uint32_t m = 0;
        const uint32_t kk = std::min(e.q.k, d.q.k);
        while (m + 1 < kk && same_step(e.q.step[m + 1], d.q.step[m + 1])) ++m;
        uint32_t w = std::min({m + 1, cwf, counted_width(e)});
        if (w < 3 || w <= best || b->hwidth[p] > w) continue;
        best = w;
        leader = p;
    }
    while (leader != kNoSlot) {
        const BDesc& e = b->h_desc[leader];
        if (e.leader == kNoSlot || e.fork_width < best) break;
        if (!b->used[e.leader] || b->h_st[e.leader].serial != e.leader_serial) break;
        leader = e.leader;
    }
    return leader == kNoSlot ? 0 : best;
}

bool tune_ok(const gv_tune_t* t) {
    return t && (t->ipt == 1 || t->ipt == 2 || t->ipt == 4 || t->ipt == 8) &&
           t->walk_blocks_per_cu >= 1 && t->walk_blocks_per_cu <= 64;
}

}  // namespace

extern "C" {

int gv_batch_create(gv_ctx_t c, gv_graph_t g, uint32_t slots, uint64_t slot_bytes,
                    uint64_t row_cap, const gv_tune_t* tune, gv_batch_t* out) {
    if (!c || !g || !out || slots == 0 || slots > kMaxSlots || slot_bytes < 4096 || row_cap == 0)
        return GV_ERR_ARG;
    if (tune && !tune_ok(tune)) return GV_ERR_ARG;
    // a sizing tile's candidate prefix is kept in 32 bits
    if ((uint64_t)g->max_deg * kBSizeTile >= (1ull << 32)) return GV_ERR_ARG;
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    gv_batch* b = new gv_batch();
    b->ctx = c;
    b->g = g;
    b->nslots = slots;
    b->slot_bytes = slot_bytes & ~uint64_t(63);
    b->row_cap = row_cap;
    b->row_cap_alloc = row_cap;
    if (tune) b->tune = *tune;
    b->used.assign(slots, 0);
    b->dirty.assign(slots, 0);
    b->flush_epoch.assign(slots, 0);
    b->hwidth.assign(slots, 0);
    b->hstate.assign(slots, GV_SLOT_FREE);
    b->gwords = g->ngraphs ? (g->ngraphs + 31) / 32 : 1;
    // the row space: row_cap rows and fewer than kXcds tiles of unused rows
    // after each slot's
    const uint64_t space = row_cap + kXcds * kBSizeTile * (uint64_t)slots;
    const uint64_t ntiles = space / kBSizeTile;
    bool ok = hipMalloc(&b->d_desc, sizeof(BDesc) * slots) == hipSuccess &&
              hipMalloc(&b->d_st, sizeof(BState) * slots) == hipSuccess &&
              hipMalloc(&b->d_half[0], b->slot_bytes * slots) == hipSuccess &&
              hipMalloc(&b->d_half[1], b->slot_bytes * slots) == hipSuccess &&
              hipMalloc(&b->d_anc_beg, 4 * space) == hipSuccess &&
              hipMalloc(&b->d_cnt, 4 * space) == hipSuccess &&
              hipMalloc(&b->d_anc_i, space) == hipSuccess &&
              hipMalloc(&b->d_tiles, 8 * (ntiles + 1)) == hipSuccess &&
              hipMalloc(&b->d_glob, 8 * 8) == hipSuccess &&
              hipMalloc(&b->d_gbits, 4ull * b->gwords * slots) == hipSuccess &&
              hipMalloc(&b->d_sacc, 8ull * kBLines * kAccStride * slots) == hipSuccess &&
              hipHostMalloc((void**)&b->h_desc, sizeof(BDesc) * slots, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&b->h_st, sizeof(BState) * slots, hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&b->d_st_stage, sizeof(BState) * slots) == hipSuccess &&
              hipMalloc(&b->d_wend, 8ull * slots) == hipSuccess &&
              hipMalloc(&b->d_wslot, 2ull * slots) == hipSuccess &&
              hipMalloc(&b->d_bits_out, 4ull * b->gwords * slots) == hipSuccess;
    if (!ok) {
        gv_batch_destroy(b);
        return GV_ERR_NOMEM;
    }
    HIPCHK(hipMemsetAsync(b->d_desc, 0, sizeof(BDesc) * slots, c->stream));
    HIPCHK(hipMemsetAsync(b->d_st, 0, sizeof(BState) * slots, c->stream));
    HIPCHK(hipMemsetAsync(b->d_glob, 0, 8 * 8, c->stream));
    HIPCHK(hipMemsetAsync(b->d_gbits, 0, 4ull * b->gwords * slots, c->stream));
    HIPCHK(hipMemsetAsync(b->d_sacc, 0, 8ull * kBLines * kAccStride * slots, c->stream));
    int rc = capture(b);
    if (rc != GV_OK) {
        gv_batch_destroy(b);
        return rc;
    }
    *out = b;
    return GV_OK;
}

int gv_batch_destroy(gv_batch_t b) {
    if (!b) return GV_ERR_ARG;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* ptrs[] = {b->d_desc, b->d_st, b->d_half[0], b->d_half[1], b->d_anc_beg, b->d_cnt,
                    b->d_anc_i, b->d_tiles, b->d_glob, b->d_gbits, b->d_sacc, b->d_kb,
                    b->d_st_stage, b->d_bits_out, b->d_wend, b->d_wslot};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->h_desc) (void)hipHostFree(b->h_desc);
    if (b->h_st) (void)hipHostFree(b->h_st);
    delete b;
    return GV_OK;
}

int gv_batch_set_tune(gv_batch_t b, const gv_tune_t* t) {
    if (!b || !tune_ok(t)) return GV_ERR_ARG;
    int rc = sync(b);
    if (rc != GV_OK) return rc;
    b->tune = *t;
    return capture(b);
}

int gv_batch_set_row_cap(gv_batch_t b, uint64_t row_cap) {
    if (!b || row_cap == 0 || row_cap > b->row_cap_alloc) return GV_ERR_ARG;
    int rc = sync(b);
    if (rc != GV_OK) return rc;
    b->row_cap = row_cap;
    return capture(b);
}

int gv_batch_admit(gv_batch_t b, const gv_query_t* q, uint64_t tag, uint32_t* slot) {
    return gv_batch_admit_leaves(b, q, nullptr, tag, slot);
}

int gv_batch_admit_leaves(gv_batch_t b, const gv_query_t* q, const gv_leaves_t* leaves, uint64_t tag,
                          uint32_t* slot) {
    return gv_batch_admit_tuned(b, q, leaves, 0, tag, slot);
}

int gv_batch_admit_tuned(gv_batch_t b, const gv_query_t* q, const gv_leaves_t* leaves, uint32_t ipt,
                         uint64_t tag, uint32_t* slot) {
    return gv_batch_admit_part(b, q, leaves, ipt, tag, 0, 1, slot);
}

int gv_batch_admit_part(gv_batch_t b, const gv_query_t* q, const gv_leaves_t* leaves, uint32_t ipt,
                        uint64_t tag, uint32_t part, uint32_t nparts, uint32_t* slot) {
    if (nparts == 0 || part >= nparts) return GV_ERR_ARG;
    if (b && ipt && (ipt > b->tune.ipt || (ipt & (ipt - 1)))) return GV_ERR_ARG;
    if (!b || !q || !slot || q->k < 2 || q->k > GV_MAXK || q->label0 >= b->g->nlabels)
        return GV_ERR_ARG;
    if (leaves && (!valid_leaves(leaves, b->g) || leaves->k != q->k)) return GV_ERR_ARG;
    for (uint32_t t = 1; t < q->k; ++t) {
        const gv_step_t& s = q->step[t];
        if (s.nback < 1 || s.nback >= GV_MAXK || s.label >= b->g->nlabels || s.ncoll >= GV_MAXK)
            return GV_ERR_ARG;
        for (int i = 0; i < s.nback; ++i)
            if (s.back_col[i] >= t || s.back_part[i] >= b->g->nparts) return GV_ERR_ARG;
        for (int i = 0; i < s.ncoll; ++i)
            if (s.coll_col[i] >= t) return GV_ERR_ARG;
        if (s.restrict_col != GV_NONE && s.restrict_col >= t) return GV_ERR_ARG;
    }
    uint32_t s = 0;
    while (s < b->nslots && b->used[s]) ++s;
    if (s == b->nslots) return GV_ERR_CAPACITY;
    // the staging entry may still be read by a copy issued since the last
    // synchronisation (only if the slot was released without a poll between)
    if (b->flush_epoch[s] == b->epoch) {
        int rc = sync(b);
        if (rc != GV_OK) return rc;
    }
    BDesc& d = b->h_desc[s];
    memset(&d, 0, sizeof(d));
    d.q = *q;
    d.tile = 64u * ipt;
    d.start_off = (uint32_t)b->g->lvoff[q->label0];
    d.start_n = (uint32_t)(b->g->lvoff[q->label0 + 1] - b->g->lvoff[q->label0]);
    if (leaves) {
        d.has_leaves = 1;
        d.leaves = *leaves;
    }
    BState& st = b->h_st[s];
    memset(&st, 0, sizeof(st));
    st.tag = tag;
    st.state = GV_SLOT_RUNNING;
    st.width = 1;
    st.rows = d.start_n;
    // the first step from the label-pair edge table: a step that extends
    // (not the counted last step, not a fused tail, not a leaf core of width
    // 1) with one back edge, to column 0, and no restriction or the one below
    // column 0 of a same-label step
    const gv_graph* G = b->g;
    const gv_step_t& s1 = q->step[1];
    const bool extends = q->k >= 3 && !(q->k == 3 && q->step[2].nback == 1) && !(leaves && leaves->width == 1);
    const uint32_t la = q->label0, lb = s1.label, NL = G->nlabels;
    if (b->pair_start && extends && s1.nback == 1 && s1.back_col[0] == 0 &&
        (s1.restrict_col == GV_NONE || (s1.restrict_col == 0 && la == lb)) && !G->pe_off.empty()) {
        const uint32_t p = s1.back_part[0];
        const uint64_t e0 = G->pe_off[((uint64_t)p * NL + la) * NL + lb];
        const uint64_t e1 = G->pe_off[((uint64_t)p * NL + la) * NL + lb + 1];
        uint64_t r0 = e0, r1 = e1;
        const uint32_t* base = G->d_pe;
        if (s1.restrict_col == 0) {
            r0 = G->pel_off[(uint64_t)p * NL + la];
            r1 = G->pel_off[(uint64_t)p * NL + la + 1];
            base = G->d_pel;
        }
        d.init_rows = base + 2 * r0;
        st.width = 2;
        st.rows = r1 - r0;
        // the first step's algorithmic bytes: every start row read (4) with
        // its row-pointer pair (8), its label-lb segment (4 per entry) and
        // each row it writes (8)
        st.bytes = 12ull * d.start_n + 4ull * (e1 - e0) + 8ull * (r1 - r0);
    }
    if (nparts > 1) {
        // rows [n * part / nparts, n * (part + 1) / nparts) of the first
        // table; the closed-form first step's bytes go with piece 0
        const uint64_t n = st.rows;
        const uint64_t lo = (uint64_t)((unsigned __int128)n * part / nparts);
        const uint64_t hi = (uint64_t)((unsigned __int128)n * (part + 1) / nparts);
        if (st.width == 1) {
            d.start_off += (uint32_t)lo;
            d.start_n = (uint32_t)(hi - lo);
        } else {
            d.init_rows += 2 * lo;
        }
        st.rows = hi - lo;
        if (part != 0) st.bytes = 0;
        d.piece = 1;
    }
    st.serial = b->next_serial++;
    d.init_width = st.width;
    d.init_n = st.rows;
    d.init_bytes = st.bytes;
    d.leader = kNoSlot;
    if (b->share && !d.piece) {
        uint32_t p;
        const uint32_t w = pick_leader(b, s, d, p);
        if (w) {
            d.leader = p;
            d.fork_width = w;
            d.leader_serial = b->h_st[p].serial;
            st.wait = 1;
            st.bytes = 0;   // the leader's counters at the shared width are taken over
        }
    }
    b->hwidth[s] = st.width;
    b->hstate[s] = GV_SLOT_RUNNING;
    b->dirty[s] = 1;
    b->used[s] = 1;
    *slot = s;
    return GV_OK;
}

int gv_batch_run(gv_batch_t b, uint32_t iterations) {
    if (!b) return GV_ERR_ARG;
    int rc = flush(b);
    if (rc != GV_OK) return rc;
    for (uint32_t i = 0; i < iterations; ++i) {
        if (b->ctx->prof) {
            // profiling: the same kernels launched one by one, each timed,
            // and the iteration's extents read back
            launch_iteration(b, true);
            HIPCHK(hipGetLastError());
            uint64_t g[4];
            HIPCHK(hipMemcpyAsync(g, b->d_glob, sizeof(g), hipMemcpyDeviceToHost, b->ctx->stream));
            int rc2 = sync(b);
            if (rc2 != GV_OK) return rc2;
            b->hist.push_back(g[0]);
            b->hist.push_back(g[1]);
        } else {
            HIPCHK(hipGraphLaunch(b->exec, b->ctx->stream));
        }
    }
    return GV_OK;
}

int gv_batch_poll_bits(gv_batch_t b, gv_slot_info_t* info, uint32_t* bits, uint32_t nwords) {
    if (!b || !info || (bits && nwords != b->gwords)) return GV_ERR_ARG;
    int rc = flush(b);
    if (rc != GV_OK) return rc;
    std::vector<BState> h(b->nslots);
    HIPCHK(hipMemcpyAsync(h.data(), b->d_st, sizeof(BState) * b->nslots, hipMemcpyDeviceToHost,
                          b->ctx->stream));
    if (bits)
        HIPCHK(hipMemcpyAsync(bits, b->d_gbits, 4ull * b->gwords * b->nslots, hipMemcpyDeviceToHost,
                              b->ctx->stream));
    rc = sync(b);
    if (rc != GV_OK) return rc;
    for (uint32_t s = 0; s < b->nslots; ++s) {
        if (b->used[s]) {
            b->hwidth[s] = h[s].width;
            b->hstate[s] = h[s].state;
        }
        info[s].tag = h[s].tag;
        info[s].state = b->used[s] ? h[s].state : GV_SLOT_FREE;
        info[s].width = h[s].width;
        info[s].rows = h[s].rows;
        info[s].count = h[s].count;
        info[s].alg_bytes = h[s].bytes;
    }
    return GV_OK;
}

int gv_batch_poll(gv_batch_t b, gv_slot_info_t* info) { return gv_batch_poll_bits(b, info, nullptr, 0); }

int gv_batch_release(gv_batch_t b, uint32_t s, uint32_t* bits, uint32_t nwords) {
    if (!b || s >= b->nslots || !b->used[s]) return GV_ERR_ARG;
    if (bits && nwords && nwords > b->gwords) return GV_ERR_ARG;
    if (b->dirty[s]) {
        // never reached the device: nothing to undo there
        b->dirty[s] = 0;
        b->used[s] = 0;
        b->hstate[s] = GV_SLOT_FREE;
        return GV_OK;
    }
    // the slot stops here whatever its state: its device state becomes free
    SlotList L;
    L.n = 1;
    L.s[0] = (uint16_t)s;
    std::vector<uint32_t> all;
    if (bits && nwords) all.resize(b->gwords);
    int rc = release_list(b, L, bits && nwords ? all.data() : nullptr);
    if (rc != GV_OK) return rc;
    if (bits && nwords) memcpy(bits, all.data(), 4ull * nwords);
    b->used[s] = 0;
    b->hstate[s] = GV_SLOT_FREE;
    return GV_OK;
}

int gv_batch_profile(gv_batch_t b, int on) {
    if (!b) return GV_ERR_ARG;
    int rc = sync(b);
    if (rc != GV_OK) return rc;
    if (on && !b->d_kb && hipMalloc(&b->d_kb, kKbBytes) != hipSuccess) return GV_ERR_NOMEM;
    if (on) HIPCHK(hipMemsetAsync(b->d_kb, 0, kKbBytes, b->ctx->stream));
    b->kb_on = on != 0;
    return capture(b);
}

int gv_batch_profile_bytes(gv_batch_t b, uint64_t* size_bytes, uint64_t* walk_bytes) {
    if (!b) return GV_ERR_ARG;
    uint64_t v[2] = {0, 0};
    if (b->d_kb) {
        std::vector<uint64_t> h(kKbBytes / 8);
        HIPCHK(hipMemcpyAsync(h.data(), b->d_kb, kKbBytes, hipMemcpyDeviceToHost, b->ctx->stream));
        int rc = sync(b);
        if (rc != GV_OK) return rc;
        for (uint32_t i = 0; i < kBLines; ++i) {
            v[0] += h[i * kAccStride];
            v[1] += h[i * kAccStride + 1];
        }
    }
    if (size_bytes) *size_bytes = v[0];
    if (walk_bytes) *walk_bytes = v[1];
    return GV_OK;
}

int gv_batch_pair_start(gv_batch_t b, int on) {
    if (!b) return GV_ERR_ARG;
    b->pair_start = on != 0;
    return GV_OK;
}

int gv_batch_share(gv_batch_t b, int on) {
    if (!b) return GV_ERR_ARG;
    b->share = on != 0;
    return GV_OK;
}

int gv_batch_fuse_sizing(gv_batch_t b, int on) {
    if (!b) return GV_ERR_ARG;
    int rc = sync(b);
    if (rc != GV_OK) return rc;
    b->fuse = on != 0;
    return capture(b);
}

int gv_batch_fused_rows(gv_batch_t b, uint64_t* rows) {
    if (!b || !rows) return GV_ERR_ARG;
    uint64_t g[8];
    HIPCHK(hipMemcpyAsync(g, b->d_glob, sizeof(g), hipMemcpyDeviceToHost, b->ctx->stream));
    int rc = sync(b);
    if (rc != GV_OK) return rc;
    *rows = g[7];
    return GV_OK;
}

int gv_batch_share_stats(gv_batch_t b, uint64_t* shared, uint64_t* restarted) {
    if (!b) return GV_ERR_ARG;
    uint64_t g[8];
    HIPCHK(hipMemcpyAsync(g, b->d_glob, sizeof(g), hipMemcpyDeviceToHost, b->ctx->stream));
    int rc = sync(b);
    if (rc != GV_OK) return rc;
    if (shared) *shared = g[4];
    if (restarted) *restarted = g[5];
    return GV_OK;
}

int gv_batch_history(gv_batch_t b, uint64_t* rows, uint64_t* candidates, uint32_t cap, uint32_t* n) {
    if (!b || !n) return GV_ERR_ARG;
    *n = (uint32_t)(b->hist.size() / 2);
    for (uint32_t i = 0; i < *n && i < cap; ++i) {
        if (rows) rows[i] = b->hist[2 * i];
        if (candidates) candidates[i] = b->hist[2 * i + 1];
    }
    if (rows || candidates) b->hist.clear();
    return GV_OK;
}

int gv_batch_release_many(gv_batch_t b, uint32_t n, const uint32_t* slots, uint32_t* bits,
                          uint32_t nwords) {
    if (!b || (n && !slots) || (bits && nwords != b->gwords)) return GV_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i)
        if (slots[i] >= b->nslots || !b->used[slots[i]]) return GV_ERR_ARG;
    // one kernel frees every listed slot that reached the device and gathers
    // their bitsets; one copy brings those back
    SlotList L;
    L.n = 0;
    std::vector<uint32_t> pos;   // output position of each listed slot
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t s = slots[i];
        if (b->dirty[s]) {   // never reached the device: no bits, nothing to undo
            if (bits) memset(bits + (uint64_t)i * nwords, 0, 4ull * nwords);
            b->dirty[s] = 0;
        } else {
            L.s[L.n++] = (uint16_t)s;
            pos.push_back(i);
        }
        b->used[s] = 0;
        b->hstate[s] = GV_SLOT_FREE;
    }
    const bool direct = L.n == n;   // no dirty slot skipped: positions are 0..n-1
    std::vector<uint32_t> gathered;
    if (bits && !direct) gathered.resize((uint64_t)L.n * nwords);
    int rc = release_list(b, L, bits ? (direct ? bits : gathered.data()) : nullptr);
    if (rc != GV_OK) return rc;
    if (bits && !direct)
        for (uint32_t k = 0; k < L.n; ++k)
            memcpy(bits + (uint64_t)pos[k] * nwords, gathered.data() + (uint64_t)k * nwords, 4ull * nwords);
    return GV_OK;
}

int gv_batch_slot_iterations(gv_batch_t b, uint32_t slot, uint32_t* iterations) {
    if (!b || slot >= b->nslots || !b->used[slot] || !iterations) return GV_ERR_ARG;
    const BDesc& d = b->h_desc[slot];
    uint32_t last = counted_width(d);
    if (d.leader != kNoSlot) last = std::max(last, d.fork_width);
    const uint32_t w = b->hwidth[slot];
    *iterations = last >= w ? last - w + 1 : 1;
    return GV_OK;
}

int gv_batch_info(gv_batch_t b, uint32_t* slots, uint64_t* slot_bytes, uint32_t* free_slots) {
    if (!b) return GV_ERR_ARG;
    if (slots) *slots = b->nslots;
    if (slot_bytes) *slot_bytes = b->slot_bytes;
    if (free_slots) {
        uint32_t n = 0;
        for (auto u : b->used) n += u ? 0 : 1;
        *free_slots = n;
    }
    return GV_OK;
}

}  // extern "C"
