/*
 * geneva_hip.h -- C ABI of the gfx950 embedding-extension library
 * (libgeneva_hip.so).
 *
 * This layer replaces the reference's kernel launchers
 * (include/miningPattern.h:1-31, include/reduction.h:1-4, implemented in
 * src/miningPattern.cu:3350-3752 and src/reduction.cu:266-296) and the CUDA
 * runtime calls its host driver makes (src/gpumine.cpp:868-1072).  Every
 * argument is a plain pointer, size or POD struct; every entry point returns a
 * status (0 = GV_OK) instead of void, and never aborts the process.
 *
 * Model.  A partial embedding table is row-major uint32[nrows][row_len]; column
 * j holds the data vertex of the j-th query vertex in match order (data ids
 * start at 1, as in the reference).  One extension step adds column row_len.
 * The step descriptor (gv_step_t) carries everything the reference spreads
 * over auxArray, restriction flags and the following reduction phases: the
 * label of the new vertex, every earlier query neighbour (its column and edge
 * label partition), the restriction column and the same-label columns that
 * enforce injectivity.  The back edges the reference verifies in separate 'R'
 * phases (src/reduction.cu:115-264) are checked during the extension as
 * sorted-set membership, so no unverified row is ever written.
 *
 * Threading: a context owns one HIP stream on one device; calls on one context
 * must not overlap.  Calls on different contexts may run concurrently.
 */
#ifndef GENEVA_HIP_H
#define GENEVA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GV_OK 0
#define GV_ERR_ARG -1       /* invalid argument */
#define GV_ERR_HIP -2       /* HIP runtime error (see gv_last_error) */
#define GV_ERR_NOMEM -3     /* device allocation failed */
#define GV_ERR_CAPACITY -4  /* output would exceed the caller's capacity */
#define GV_ERR_FORMAT -5    /* malformed partition block */
#define GV_ERR_NODEV -6     /* no HIP device */

#define GV_MAXK 16          /* query vertices per pattern */
#define GV_NONE 0xffu

typedef struct gv_ctx* gv_ctx_t;
typedef struct gv_graph* gv_graph_t;
typedef struct gv_frontier* gv_frontier_t;

/* One extension step: the vertex placed in column `col` (= input row_len). */
typedef struct {
    uint32_t label;              /* dense vertex label of the new vertex */
    uint32_t col;                /* column it occupies */
    uint8_t nback;               /* earlier query neighbours, 1..GV_MAXK-1 */
    uint8_t back_col[GV_MAXK];   /* their columns */
    uint8_t back_part[GV_MAXK];  /* edge-label partition of each back edge */
    uint8_t restrict_col;        /* new vertex id < row[restrict_col]; GV_NONE = none */
    uint8_t ncoll;               /* earlier columns holding the same vertex label */
    uint8_t coll_col[GV_MAXK];
} gv_step_t;

/* Leaves of a query: the vertices in columns `width` .. k-1 each hang by a
 * single edge from a core column (< width).  Leaves with the same anchor
 * column, partition and label form a group; groups have pairwise different
 * labels, the first leaf of a group may be restricted to a core column and
 * every later one is restricted to the group's previous leaf (the chain the
 * planner's automorphism breaking gives).  A core row then has
 * prod_g C(s_g, m_g) embeddings, s_g = the group's segment below the bound
 * minus the core values of its label found there. */
typedef struct {
    uint8_t anchor_col;          /* core column the group's leaves hang from */
    uint8_t part;                /* edge-label partition of their edges */
    uint8_t m;                   /* leaves in the group */
    uint8_t restrict_col;        /* first leaf < row[restrict_col]; GV_NONE = none */
    uint32_t label;              /* dense vertex label of the leaves */
    uint8_t ncoll;               /* core columns holding the same label */
    uint8_t coll_col[GV_MAXK];
} gv_leaf_group_t;

#define GV_MAX_LEAF_GROUPS 8
typedef struct {
    uint32_t width;              /* core width */
    uint32_t k;                  /* query vertices */
    uint32_t ngroups;
    uint8_t group_of[GV_MAXK];   /* group of the leaf in column width + j */
    gv_leaf_group_t group[GV_MAX_LEAF_GROUPS];
} gv_leaves_t;

/* A compiled query: k vertices in match order; step[t] adds column t. */
typedef struct {
    uint32_t k;
    uint32_t label0;             /* dense label of column 0 */
    gv_step_t step[GV_MAXK];
} gv_query_t;

/* ---- errors --------------------------------------------------------------- */
const char* gv_status_string(int status);
/* Text of the last HIP error seen on this thread ("" if none). */
const char* gv_last_error(void);

/* ---- context and device memory (src/gpumine.cpp:868,885-906,952) ---------- */
int gv_ctx_create(int device, gv_ctx_t* out);
int gv_ctx_destroy(gv_ctx_t ctx);
int gv_ctx_synchronize(gv_ctx_t ctx);
int gv_device_count(int* n);
int gv_mem_get_info(gv_ctx_t ctx, uint64_t* free_bytes, uint64_t* total_bytes);
int gv_malloc(gv_ctx_t ctx, uint64_t bytes, void** dptr);
int gv_free(gv_ctx_t ctx, void* dptr);
/* Stream-ordered copies on the context's stream; *_d2h waits for completion.
 * Replace cudaMemcpyAsync in the planner upload (src/genMatchOrder.cpp:768)
 * and the counter read-backs (src/gpumine.cpp:395,574,807,952). */
int gv_memcpy_h2d(gv_ctx_t ctx, void* dst_dev, const void* src_host, uint64_t bytes);
int gv_memcpy_d2h(gv_ctx_t ctx, void* dst_host, const void* src_dev, uint64_t bytes);

/* ---- resident data graph -----------------------------------------------------
 * Built from exactly what loadDataBase returns (src/loadDataBase.cpp:309-365):
 * vLabel (vLabel[0] = max id) and the edge-label partition blocks in the
 * .mygraph layout (tools/genDatabase.cpp:437-580; skip-index prefixes are
 * recognised and dropped).  block_edge_labels[i] names the edge label of
 * blocks[i].  The device copy is a label-segmented CSR: for partition p,
 * vertex v and dense label l, the neighbours of v with label l are
 * nbr[seg(p,v,l) .. seg(p,v,l+1)), sorted ascending.  This replaces the
 * interval lookup and FIND_LABEL_LIMIT search (include/common.h:27-75). */
int gv_graph_create(gv_ctx_t ctx, const uint32_t* vLabel, uint32_t nblocks,
                    const uint32_t* const* blocks, const uint32_t* block_edge_labels,
                    gv_graph_t* out);
int gv_graph_destroy(gv_graph_t g);
/* Optional owner table for a database of graphs: graph_of[v] for v in 0..maxid. */
int gv_graph_set_owner(gv_graph_t g, const uint32_t* graph_of, uint32_t ngraphs);
/* Number of owner graphs (0 when no owner table is set). */
int gv_graph_owners(gv_graph_t g, uint32_t* ngraphs);
/* Dense index of an original vertex label / partition of an edge label;
 * GV_ERR_ARG when the graph has none. */
int gv_graph_label_index(gv_graph_t g, uint32_t label, uint32_t* dense);
int gv_graph_part_index(gv_graph_t g, uint32_t edge_label, uint32_t* part);
/* Sizes: vertices (max id + 1), dense labels, partitions, neighbour entries,
 * device bytes held. */
int gv_graph_info(gv_graph_t g, uint64_t* nv1, uint32_t* nlabels, uint32_t* nparts,
                  uint64_t* nnbr, uint64_t* dev_bytes);
/* Largest neighbour list of any vertex (bounds every label segment). */
int gv_graph_max_degree(gv_graph_t g, uint32_t* max_degree);
/* Label statistics of partition `part`: vertices per dense label
 * (nlabels entries) and directed edges per (source label, neighbour label)
 * (nlabels*nlabels entries, row-major).  Either pointer may be null. */
int gv_graph_label_stats(gv_graph_t g, uint32_t part, uint64_t* vertices_by_label,
                         uint64_t* edges_by_label_pair);
/* The one-column start table of a label: every vertex carrying it, ascending. */
int gv_graph_start_rows(gv_graph_t g, uint32_t dense_label, const uint32_t** rows_dev,
                        uint64_t* nrows);

/* ---- frontier sizing --------------------------------------------------------
 * A frontier holds, for up to `capacity` input rows, the chosen anchor list
 * (the shortest back-neighbour list), its length after the restriction cut,
 * and the inclusive prefix sum of those lengths: the candidate index space
 * [0, total) of one extension step. */
int gv_frontier_create(gv_ctx_t ctx, uint64_t capacity, gv_frontier_t* out);
int gv_frontier_destroy(gv_frontier_t f);

/* Size the extension of `rows` by `step` (replaces the per-row label search
 * and restriction cut of src/miningPattern.cu:3249-3324).  *total receives
 * the number of candidates (stream-synchronising read). */
int gv_size_rows(gv_ctx_t ctx, gv_graph_t g, const gv_step_t* step, const uint32_t* rows_dev,
                 uint32_t row_len, uint64_t nrows, gv_frontier_t f, uint64_t* total);

/* Extension (replaces initEmb_2V_NoHash, initEmb_3V_1and2_NoHash, extEmb_1V_NoHash,
 * the four extEmb_2V_* launchers and the reductionPhase that follows them;
 * include/miningPattern.h:1-31): candidates [cand_begin, cand_end) of the
 * sized frontier are checked against every back edge, the restriction and
 * injectivity, and survivors are appended to new_rows_dev as rows of
 * row_len+1 uints.  Fails with GV_ERR_CAPACITY (nothing lost: the call can be
 * repeated on a smaller range) when more than max_rows would be written; since
 * survivors never exceed candidates, cand_end-cand_begin <= max_rows always
 * succeeds.  *rows_written is read back before returning. */
int gv_ext_emb(gv_ctx_t ctx, gv_graph_t g, const gv_step_t* step, const uint32_t* rows_dev,
               uint32_t row_len, uint64_t nrows, gv_frontier_t f, uint64_t cand_begin,
               uint64_t cand_end, uint32_t* new_rows_dev, uint64_t max_rows,
               uint64_t* rows_written);

/* Count-only final phase (the isLastPhase branch of the reference launchers):
 * candidates [cand_begin, cand_end) that satisfy `step` are counted into the
 * context's uint64 accumulator; when `tail` is non-null each survivor is
 * instead extended by the single-back-edge step `tail` and the number of valid
 * tail vertices is counted without enumerating them.  Matched graph ids (owner
 * of column 0) are OR-ed into the context's bitset.  Asynchronous. */
int gv_count_emb(gv_ctx_t ctx, gv_graph_t g, const gv_step_t* step, const gv_step_t* tail,
                 const uint32_t* rows_dev, uint32_t row_len, uint64_t nrows, gv_frontier_t f,
                 uint64_t cand_begin, uint64_t cand_end);

/* Count-only final phase without a candidate walk, for a last materialised
 * step with a single back edge (its survivors per row are one label segment
 * cut at the restriction, minus the row's same-label values in it).  Row r's
 * count is, by gv_count_rows_mode(step, tail, row_len):
 *   0  (tail == NULL): the survivors s;
 *   1  tail independent of the survivor (its back edge, restriction and
 *      same-label columns all in the row): s * T(row);
 *   2  tail draws from the survivor's own segment and lies below it: s(s-1)/2.
 * -1 = not applicable (use gv_size_rows + gv_count_emb; gv_count_rows then
 * returns GV_ERR_ARG).  Counts, algorithmic bytes and graph bits accumulate
 * exactly as gv_count_emb's do.  Replaces, like gv_count_emb, the isLastPhase
 * branch of extEmb_1V_NoHash (include/miningPattern.h:29-31).  Asynchronous. */
int gv_count_rows_mode(const gv_step_t* step, const gv_step_t* tail, uint32_t row_len);
int gv_count_rows(gv_ctx_t ctx, gv_graph_t g, const gv_step_t* step, const gv_step_t* tail,
                  const uint32_t* rows_dev, uint32_t row_len, uint64_t nrows);

/* A whole query on the plain path with every extent kept in device memory:
 * (leaves, when non-null, ends it at width leaves->width with gv_count_leaves)
 * the step kernels read their row counts from the previous step's output
 * counter, so nothing is read back until the end (one synchronisation).
 * Start rows: start_rows[0..nstart).  For each materialised width w (2 ..
 * the counted step's input width), tables[w] holds up to table_rows[w] rows;
 * frontiers[w] (capacity >= that width's rows; frontiers[1] >= nstart) holds
 * the sizing arrays of step w.  count_by_rows: use gv_count_rows for a last
 * step it applies to.  Count, bytes and graph bits accumulate into the
 * context's accumulators (gv_counters_read; count / alg_bytes / bits, when
 * non-null, receive them too).  GV_ERR_CAPACITY: some table outgrew its
 * capacity; nothing was counted past that point and the accumulators must be
 * reset before the query is mined again with the chunked step entries.
 * Replaces the phase loop of src/gpumine.cpp:993-1061 with its per-launch
 * read-backs. */
int gv_mine_levels(gv_ctx_t ctx, gv_graph_t g, const gv_query_t* q, const uint32_t* start_rows,
                   uint64_t nstart, uint32_t* const* tables, const uint64_t* table_rows,
                   const gv_frontier_t* frontiers, int count_by_rows, const gv_leaves_t* leaves,
                   uint64_t* count, uint64_t* alg_bytes, uint32_t* bits, uint32_t nwords);

/* Count-only final phases for a query whose last vertices are leaf groups
 * (gv_leaves_t): one thread per core row (rows of d->width columns), no
 * candidate walk and nothing materialised past the core.  Counts, algorithmic
 * bytes (those of every row the enumeration would hold) and graph bits
 * accumulate like gv_count_emb's.  GV_ERR_ARG when the descriptor is not
 * valid for this graph (including max_degree^m >= 2^63 for some group).
 * Replaces the isLastPhase extension phases of src/gpumine.cpp:993-1061 for
 * such queries.  Asynchronous. */
int gv_count_leaves(gv_ctx_t ctx, gv_graph_t g, const gv_leaves_t* d, const uint32_t* rows_dev,
                    uint64_t nrows);

/* Accumulators of a context: embedding count, algorithmic bytes and matched
 * graph bitset (ngraphs bits).  reset zeroes them; read synchronises. */
int gv_counters_reset(gv_ctx_t ctx, gv_graph_t g);
int gv_counters_read(gv_ctx_t ctx, uint64_t* count, uint64_t* alg_bytes,
                     uint32_t* graph_bits, uint32_t nwords);

/* Overlap-safe device copy of n uints (src/reduction.cu:280-296). */
int gv_copy_partial_emb(gv_ctx_t ctx, const uint32_t* src_dev, uint32_t* dst_dev, uint64_t n);

/* ---- batched execution ------------------------------------------------------
 * Many queries advance together: every iteration moves each running query one
 * vertex further, each at its own depth, in one fixed sequence of kernels
 * (prepare, size, scan, extend/count, retire) that reads all sizes from
 * device memory.  The sequence is captured once in a hipGraph and replayed, so
 * an iteration costs one graph launch and no host synchronisation.  A query
 * occupies a slot: two private row tables of `slot_bytes` each (input and
 * output, swapped every iteration).  A query whose table would exceed its slot
 * stops with state GV_SLOT_OVERFLOW and its partial counts are discarded; the
 * caller re-runs it on the chunked plain path.  Up to 1024 slots. */
typedef struct gv_batch* gv_batch_t;

typedef struct {
    uint32_t ipt;             /* candidates per lane per extension tile: 1, 2, 4 or 8 */
    uint32_t walk_blocks_per_cu;  /* persistent grid of the extension kernel */
} gv_tune_t;

#define GV_SLOT_FREE 0
#define GV_SLOT_RUNNING 1
#define GV_SLOT_DONE 2
#define GV_SLOT_OVERFLOW 3

typedef struct {
    uint64_t tag;             /* caller's id, given at admission */
    uint32_t state;           /* GV_SLOT_* */
    uint32_t width;           /* columns placed */
    uint64_t rows;            /* rows of the current table */
    uint64_t count;           /* embeddings counted (valid when DONE) */
    uint64_t alg_bytes;       /* algorithmic bytes (same figure as the plain path) */
} gv_slot_info_t;

/* row_cap: rows the whole batch may size in one iteration (queries beyond it
 * wait for the next iteration).  GV_ERR_ARG for a graph with a neighbour list
 * of 2^22 entries or more. */
int gv_batch_create(gv_ctx_t ctx, gv_graph_t g, uint32_t slots, uint64_t slot_bytes,
                    uint64_t row_cap, const gv_tune_t* tune, gv_batch_t* out);
int gv_batch_destroy(gv_batch_t b);
/* Rows sized per iteration from now on (at most the row_cap the batch was
 * created with; re-captures the iteration).  Slots whose table alone exceeds
 * it stop with GV_SLOT_OVERFLOW. */
int gv_batch_set_row_cap(gv_batch_t b, uint64_t row_cap);
/* Re-capture the iteration with other tuning parameters. */
int gv_batch_set_tune(gv_batch_t b, const gv_tune_t* tune);
/* Place a compiled query in a free slot (stream-ordered; effective at the next
 * iteration).  GV_ERR_CAPACITY when no slot is free. */
int gv_batch_admit(gv_batch_t b, const gv_query_t* q, uint64_t tag, uint32_t* slot);
/* As gv_batch_admit; leaves (nullable, see gv_leaves_t) makes the slot count
 * its leaf groups per core row when its table reaches leaves->width. */
int gv_batch_admit_leaves(gv_batch_t b, const gv_query_t* q, const gv_leaves_t* leaves,
                          uint64_t tag, uint32_t* slot);
/* As gv_batch_admit_leaves, with this slot's candidates per lane per walk
 * tile (a power of two up to the batch's gv_tune_t.ipt; 0 = the batch's):
 * the per-query-class tuning of the wavefront split. */
int gv_batch_admit_tuned(gv_batch_t b, const gv_query_t* q, const gv_leaves_t* leaves, uint32_t ipt,
                         uint64_t tag, uint32_t* slot);
/* As gv_batch_admit_tuned, for piece `part` of `nparts` of the query: the
 * slot starts from rows [n * part / nparts, n * (part + 1) / nparts) of the
 * query's first table (its start vertices, or its label-pair edge rows), so
 * the counts, algorithmic bytes and graph bitsets of the nparts pieces add up
 * (bitsets by union) to the whole query's.  Pieces never share tables.  For
 * a query too large for one slot. */
int gv_batch_admit_part(gv_batch_t b, const gv_query_t* q, const gv_leaves_t* leaves, uint32_t ipt,
                        uint64_t tag, uint32_t part, uint32_t nparts, uint32_t* slot);
/* Launch `iterations` iterations (asynchronous). */
int gv_batch_run(gv_batch_t b, uint32_t iterations);
/* Copy every slot's state to info[0..slots) (synchronises). */
int gv_batch_poll(gv_batch_t b, gv_slot_info_t* info);
/* As gv_batch_poll, and every slot's matched-graph bitset into
 * bits[slot * nwords ..] (nwords = ceil(graphs / 32), all slots, one copy). */
int gv_batch_poll_bits(gv_batch_t b, gv_slot_info_t* info, uint32_t* bits, uint32_t nwords);
/* Free a slot (finished or not: a running slot stops).  graph_bits (nullable)
 * receives its matched graph ids, which synchronises; without it the call
 * only enqueues the reset. */
int gv_batch_release(gv_batch_t b, uint32_t slot, uint32_t* graph_bits, uint32_t nwords);
/* Free n slots at once; bits (nullable) receives slot slots[i]'s matched graph
 * ids at bits[i * nwords ..] (nwords = ceil(graphs / 32)), with one
 * synchronisation for all of them. */
int gv_batch_release_many(gv_batch_t b, uint32_t n, const uint32_t* slots, uint32_t* bits,
                          uint32_t nwords);
/* Per-kernel algorithmic bytes of the batch engine: on = 1 re-captures the
 * iteration with counters of the bytes B charged to k_b_size (rows and back
 * lists read, rows counted there) and to k_b_walk (rows written, lists of the
 * counted walk), zeroed here; on = 0 re-captures without them.  While the
 * context's profiling (gv_profile_enable) is on, gv_batch_run launches the
 * iteration's kernels one by one, each bracketed by HIP events of its kind
 * (k_b_size: sizing, k_b_scan_*: scan, k_b_walk: extension, the others:
 * batch). */
int gv_batch_profile(gv_batch_t b, int on);
int gv_batch_profile_bytes(gv_batch_t b, uint64_t* size_bytes, uint64_t* walk_bytes);
/* Queries admitted from now on start on the graph's label-pair edge table
 * when their first step only lists the edges between two labels (one back
 * edge to column 0; no restriction, or the one below column 0 of a
 * same-label step) and is not their counted last step: the slot starts at
 * width 2 with that table as its input, the first iteration is saved, and
 * the step's algorithmic bytes are added in closed form.  On by default. */
int gv_batch_pair_start(gv_batch_t b, int on);
/* Prefix sharing (on by default): a query admitted while a running slot's
 * query has the same start label and the same first steps up to a table of
 * width d >= 3 (the leader extending from that width, the follower extending
 * from it or counting its last step on it) waits for that slot's table
 * of width d and extends it, in the iteration the leader does, instead of
 * building its own copy; it takes over the leader's counters at that point
 * (the shared steps' algorithmic bytes).  A follower whose leader ran empty
 * or overflowed before width d ends the same way; one whose leader went past
 * width d or left its slot, or whose shared table cannot fit an iteration
 * with its leader, starts on its own.  Applies to queries admitted from now
 * on. */
int gv_batch_share(gv_batch_t b, int on);
/* Tables shared and followers restarted on their own since creation
 * (synchronises). */
int gv_batch_share_stats(gv_batch_t b, uint64_t* shared, uint64_t* restarted);
/* Sizing groups (on by default): slots that size the same input table in an
 * iteration (the same start list or label-pair range, a leader and the
 * followers reading its table) are sized tile by tile together, each tile of
 * the group's slots by one workgroup in a row, so the rows and neighbour
 * lists one slot reads are still cached for the next.  Results are the same
 * either way.  Synchronises and captures the iteration again. */
int gv_batch_fuse_sizing(gv_batch_t b, int on);
/* Rows sized as part of another slot's sizing group since creation
 * (synchronises). */
int gv_batch_fused_rows(gv_batch_t b, uint64_t* rows);
/* Rows sized and candidates walked by each iteration run while the
 * context's profiling was on (oldest first; cleared when read into rows or
 * candidates). */
int gv_batch_history(gv_batch_t b, uint64_t* rows, uint64_t* candidates, uint32_t cap, uint32_t* n);
/* Iterations a running slot needs from the width it had at admission or at
 * the last poll until it finishes, when it is never deferred: its counted
 * width (the width of the table its last step, a fused tail or its leaf
 * groups are counted on; for a follower at least the shared width) minus
 * that width, plus one. */
int gv_batch_slot_iterations(gv_batch_t b, uint32_t slot, uint32_t* iterations);
/* Slots, rows a slot table holds at a given width, free slots. */
int gv_batch_info(gv_batch_t b, uint32_t* slots, uint64_t* slot_bytes, uint32_t* free_slots);

/* Per-kernel profiling on the context's stream.  on = 1: every launch is
 * bracketed by HIP events; on = 2: in addition each kernel kind counts the
 * bytes its algorithm reads and writes (each access once) and its work units
 * (launch variants that carry that arithmetic); 0: off.  Kinds: 0 sizing
 * (k_size), 1 scan (k_scan_*), 2 extension (k_walk, writing rows), 3 counting
 * (k_walk, last step), 4 batch iteration (one graph launch), 5 per-row
 * counting (k_count_rows, k_count_leaves), 6 parent intersections of a
 * prefix-shared counted step (their sizing, k_isect, k_child_size; no work
 * units or requested bytes), 7 pendant chains (k_chain_mark,
 * k_chain_weights, k_count_chain).  read synchronises and returns the
 * launch count, the summed device milliseconds and the requested bytes since
 * the last reset. */
#define GV_KIND_SIZE 0
#define GV_KIND_SCAN 1
#define GV_KIND_EXTEND 2
#define GV_KIND_COUNT 3
#define GV_KIND_BATCH 4
#define GV_KIND_ROWS 5
#define GV_KIND_ISECT 6
#define GV_KIND_CHAIN 7
int gv_profile_enable(gv_ctx_t ctx, int on);
int gv_profile_reset(gv_ctx_t ctx);
int gv_profile_read(gv_ctx_t ctx, int kind, uint64_t* launches, double* total_ms,
                    uint64_t* req_bytes);
/* Work units of a kind since the last reset: rows (size, scan, rows kinds) or
 * candidates (extend, count); 0 for batch iterations (extents on the device). */
int gv_profile_units(gv_ctx_t ctx, int kind, uint64_t* units);
/* Algorithmic bytes B (the reference figure, accumulated exactly as for the
 * counters) of a kind since the last reset, with on = 2. */
int gv_profile_alg_bytes(gv_ctx_t ctx, int kind, uint64_t* bytes);

/* Every profiled launch since the last reset, in launch order: its kind and
 * device milliseconds (up to cap of them; *n receives the number). */
int gv_profile_records(gv_ctx_t ctx, int* kinds, float* ms, uint64_t cap, uint64_t* n);

/* Timing on the context's stream (HIP events): record a mark, and the
 * milliseconds between two marks (synchronises on the second). */
int gv_mark(gv_ctx_t ctx, int slot);
int gv_elapsed_ms(gv_ctx_t ctx, int slot_a, int slot_b, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* GENEVA_HIP_H */
