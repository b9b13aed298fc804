/*
 * geneva_host.h -- C ABI of the host library (libgeneva_host.so): data sets,
 * the database loader, the match-order planner and the mining drivers built
 * on libgeneva_hip.so.  Python binds it with ctypes (geneva_amd/__init__.py).
 *
 * Handles are opaque.  Functions returning int give 0 on success and a
 * negative GV_* status otherwise; gh_last_error() describes the last failure
 * on the calling thread.
 */
#ifndef GENEVA_HOST_H
#define GENEVA_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gh_dataset* gh_dataset_t;
typedef struct gh_engine* gh_engine_t;

const char* gh_last_error(void);

/* ---- data sets -------------------------------------------------------------
 * A data set is a labelled graph (vertex ids 1..nv), an optional owner graph
 * per vertex (databases of graphs) and a list of query texts. */
gh_dataset_t gh_ds_gen_database(uint32_t ngraphs, uint32_t nmin, uint32_t nmax, double avg_deg,
                                double ba_fraction, uint32_t ba_m, uint32_t nlabels,
                                double zipf_s, uint64_t seed);
gh_dataset_t gh_ds_gen_chung_lu(uint32_t nv, double gamma, double mean_deg, double wmax,
                                uint32_t nlabels, double zipf_s, uint64_t seed);
/* Load through loadDataBase (src/loadDataBase.cpp:309-365); the edge label of
 * the k-th stored partition is k. */
gh_dataset_t gh_ds_load(const char* mygraph_path, const char* metadata_path);
/* Graph from arrays: vlabel[0..nv] (index 0 unused), m edges (a, b, label). */
gh_dataset_t gh_ds_from_arrays(uint32_t nv, const uint32_t* vlabel, uint64_t m,
                               const uint32_t* edges, const uint32_t* graph_of,
                               uint32_t ngraphs);
void gh_ds_free(gh_dataset_t ds);
/* min_deg: resample until every query vertex has at least min_deg query
 * neighbours (0 = no condition). */
int gh_ds_add_walk_queries(gh_dataset_t ds, uint32_t nq, uint32_t kmin, uint32_t kmax,
                           int induced, uint32_t min_extra, uint64_t seed, uint32_t min_deg);
int gh_ds_add_shape_queries(gh_dataset_t ds, uint64_t seed);
int gh_ds_add_query(gh_dataset_t ds, const char* text);
uint32_t gh_ds_nv(gh_dataset_t ds);
const uint32_t* gh_ds_vlabel(gh_dataset_t ds);          /* nv+1 entries */
uint64_t gh_ds_nedges(gh_dataset_t ds);
const uint32_t* gh_ds_edges(gh_dataset_t ds);           /* 3 per edge: a, b, label */
const uint32_t* gh_ds_graph_of(gh_dataset_t ds);        /* nv+1 entries */
uint32_t gh_ds_ngraphs(gh_dataset_t ds);
uint32_t gh_ds_nqueries(gh_dataset_t ds);
const char* gh_ds_query(gh_dataset_t ds, uint32_t i);
/* Write prefix.g, prefix.metadata, prefix.mygraph, prefix.reorder (prefix-free
 * groups unless skip_index) and prefix_<i>.query. */
int gh_ds_write(gh_dataset_t ds, const char* prefix, int skip_index);

/* Renumber the data set's vertices so that every edge-label partition's
 * source vertices form at most 256 runs of consecutive ids, the limit of the
 * partition block (tools/genDatabase.cpp:451-460; the reference reorders for
 * the same purpose, tools/genDatabase.cpp:259-413): vertices ordered by their
 * edge-label signature as a Gray code (reorderForIntervals in
 * loadDataBase.h).  Vertex labels, edges and owner graphs move with their
 * vertices; the original ids are kept (gh_ds_old_ids, and the third column of
 * the .metadata gh_ds_write writes).  Returns the largest run count, or a
 * negative status when some partition would still need more than 256. */
int gh_ds_reorder_intervals(gh_dataset_t ds);
/* Source runs each edge label's partition needs under the current ids:
 * labels[i], runs[i] for i < min(cap, returned count). */
int gh_ds_partition_runs(gh_dataset_t ds, uint32_t* labels, uint32_t* runs, uint32_t cap);
/* Original id of each vertex after gh_ds_reorder_intervals (nv+1 entries),
 * or NULL. */
const uint32_t* gh_ds_old_ids(gh_dataset_t ds);

/* ---- planner ---------------------------------------------------------------
 * Plan a query with genMatchOrder (single_v = 0) or genMatchOrderSingleV and
 * print it in the layout the reference cpumine prints
 * (src/cpumine.cpp:661-674).  Returns the text length (0 on error). */
uint64_t gh_plan_dump(const char* query_text, int single_v, char* out, uint64_t cap);
/* Two-vertex extension phases of the DV plan (tools/classifyExtPhase.cpp:27-38);
 * negative on error. */
int gh_dv_phase_count(const char* query_text);

/* ---- engine: resident graph + plain path -------------------------------- */
gh_engine_t gh_engine_create(gh_dataset_t ds, int device, uint64_t level_bytes, int* status);
void gh_engine_free(gh_engine_t e);
/* Device bytes of the resident graph. */
uint64_t gh_engine_graph_bytes(gh_engine_t e);
/* Mine one query on the plain path (one query per launch sequence).
 * graph_bits (nullable) receives ceil(ngraphs/32) words. */
int gh_engine_mine(gh_engine_t e, const char* query_text, int single_v, uint64_t* count,
                   uint64_t* alg_bytes, double* device_ms, uint32_t* graph_bits,
                   uint32_t nwords, uint64_t* launches);
/* One shard of the start vertices: indices start_begin, start_begin+stride, ...
 * below start_end (0 = all) of the id-sorted list of vertices carrying the
 * first query vertex's label.  Counts of the shards of a partition add up to
 * the whole count (the unit a multi-GPU run would hand each device). */
int gh_engine_mine_shard(gh_engine_t e, const char* query_text, int single_v,
                         uint64_t start_begin, uint64_t start_end, uint64_t stride,
                         uint64_t* count, uint64_t* alg_bytes, double* device_ms,
                         uint32_t* graph_bits, uint32_t nwords, uint64_t* launches);

/* Plain path: stop each query after `seconds` of wall time (0 = no limit).
 * Under a limit a query runs on the chunked path, launch by launch; a query
 * stopped at the limit returns the exact count, bytes and graph ids of the
 * embeddings whose first-extension candidate lies in the part mined, and
 * gh_engine_last_progress reports finished = 0 and that part's share of the
 * first extension's candidates. */
int gh_engine_time_limit(gh_engine_t e, double seconds);
int gh_engine_last_progress(gh_engine_t e, int* finished, double* fraction);

/* Plain path: count a last step with a single back edge per row
 * (gv_count_rows, default on) or by walking its candidates (off). */
int gh_engine_count_by_rows(gh_engine_t e, int on);

/* ---- batched path ------------------------------------------------------------
 * Configure the batch engine (slots, bytes per slot table, rows sized per
 * iteration, candidates per lane per tile, extension blocks per CU); the
 * engine is (re)built on the next batched call. */
int gh_engine_batch_config(gh_engine_t e, uint32_t slots, uint64_t slot_bytes, uint64_t row_cap,
                           uint32_t ipt, uint32_t blocks_per_cu);
/* Mine n queries together.  counts/alg_bytes: n entries; graph_bits: n*nwords
 * (nullable); stats (nullable, 8 doubles): batched, overflowed (re-run on the
 * plain path), graph launches, device ms of the batch iterations, wall ms of
 * the batch run, ms spent planning the queries, wall ms of the plain-path
 * re-runs, splits of queries too large for a slot into pieces of their first
 * table (gv_batch_admit_part). */
int gh_engine_batch_mine(gh_engine_t e, const char* const* queries, uint32_t n, int single_v,
                         uint64_t* counts, uint64_t* alg_bytes, uint32_t* graph_bits,
                         uint32_t nwords, double* stats);
/* Autotune the batch engine on a sample; reports the chosen candidates per
 * lane and blocks per CU and its device ms. */
int gh_engine_autotune(gh_engine_t e, const char* const* queries, uint32_t n, uint32_t* ipt,
                       uint32_t* blocks_per_cu, double* best_ms);
/* Rows the batch engine sizes per iteration (after autotune: its choice). */
int gh_engine_batch_row_cap(gh_engine_t e, uint64_t* rows);
/* The per-query-class choices of the last autotune as text, one
 * "class:candidates per lane:device ms;" entry per class (class = k<vertices>
 * c<steps with several back edges>[l if leaf groups]).  Returns its length. */
uint64_t gh_engine_autotune_classes(gh_engine_t e, char* out, uint64_t cap);
/* Serve n requests arriving at arrival_s[i] seconds (non-decreasing).
 * mode 0: continuous batching with admission under mem_budget bytes of
 * estimated peak tables (requests that leave the batch engine are mined on
 * the plain path by a worker thread on a second context); mode 1: plain
 * path, one request at a time.  latency_ms and counts receive n entries;
 * graph_bits (nullable) receives each request's matched graph ids,
 * nwords = ceil(ngraphs/32) words per request; stats (nullable, 6 doubles):
 * wall seconds, iterations, most requests in flight, requests re-run on the
 * plain path after overflowing a slot, requests sent to the plain path
 * without entering the batch engine, most requests on the batch engine at
 * once. */
int gh_engine_serve(gh_engine_t e, const char* const* queries, const double* arrival_s, uint32_t n,
                    int mode, uint64_t mem_budget, double* latency_ms, uint64_t* counts,
                    uint32_t* graph_bits, uint32_t nwords, double* stats);

/* ---- the reference's phase driver over the reference-named launchers ------
 * Mines the query in query_file over a .mygraph / .metadata pair the way the
 * reference drivers do (src/gpumine.cpp:866-1072 with single_v = 0,
 * src/gpumine_singleV.cpp:707-928 with single_v = 1): loadDataBase, the
 * reference planner, one device arena of arena_bytes (0 = 1 GiB) on a context
 * of its own, every phase through the launchers of geneva_launchers.hpp with
 * a counter read-back after each, overflow rows moved to host memory and the
 * launch resumed with isContinue.  single_v: 0 DV plan, 1 SV plan, -1 the
 * kind chosen from the DV-phase count (DV when the plan has a two-vertex
 * phase; gh_phase_single_v reports the choice).  max_rows (0 = none) caps maxRowNum so the
 * overflow path can be exercised on small inputs.  The result holds the
 * final rows (keep_rows), spilled parts included; column c holds the data
 * vertex of query vertex order[c]. */
typedef struct gh_phase_run* gh_phase_run_t;
int gh_phase_mine(const char* mygraph, const char* metadata, const char* query_file, int single_v, int device,
                  uint64_t arena_bytes, uint32_t max_rows, int keep_rows, gh_phase_run_t* out);
uint64_t gh_phase_count(gh_phase_run_t r);
uint32_t gh_phase_width(gh_phase_run_t r);
const uint32_t* gh_phase_order(gh_phase_run_t r);     /* width entries */
const uint32_t* gh_phase_rows(gh_phase_run_t r);      /* count * width entries, or NULL */
uint64_t gh_phase_launches(gh_phase_run_t r);
uint64_t gh_phase_spills(gh_phase_run_t r);
int gh_phase_single_v(gh_phase_run_t r);
void gh_phase_free(gh_phase_run_t r);

/* Batch engine: per-kernel algorithmic bytes (gv_batch_profile) on or off,
 * and the bytes counted since it was turned on. */
int gh_engine_batch_profile(gh_engine_t e, int on);
int gh_engine_batch_profile_bytes(gh_engine_t e, uint64_t* size_bytes, uint64_t* walk_bytes);
/* Batch engine: first steps from the label-pair edge tables on or off
 * (gv_batch_pair_start). */
int gh_engine_batch_pair_start(gh_engine_t e, int on);
/* Batch engine: prefix sharing between queries on or off (gv_batch_share),
 * and the tables shared and followers restarted so far. */
int gh_engine_batch_share(gh_engine_t e, int on);
int gh_engine_batch_share_stats(gh_engine_t e, uint64_t* shared, uint64_t* restarted);
/* Batch engine: sizing groups on or off (gv_batch_fuse_sizing), and the rows
 * sized with another slot's group so far (gv_batch_fused_rows). */
int gh_engine_batch_fuse_sizing(gh_engine_t e, int on);
int gh_engine_batch_fused_rows(gh_engine_t e, uint64_t* rows);
/* Rows sized and candidates walked per profiled batch iteration
 * (gv_batch_history). */
int gh_engine_batch_history(gh_engine_t e, uint64_t* rows, uint64_t* candidates, uint32_t cap, uint32_t* n);
/* Every profiled launch in order: kind and device ms (gv_profile_records). */
int gh_engine_profile_records(gh_engine_t e, int* kinds, float* ms, uint64_t cap, uint64_t* n);

/* Per-kernel profiling of the engine's context (see gv_profile_* in
 * geneva_hip.h): enable/disable, reset, and read one kernel kind. */
int gh_engine_profile(gh_engine_t e, int on);
int gh_engine_profile_reset(gh_engine_t e);
int gh_engine_profile_read(gh_engine_t e, int kind, uint64_t* launches, double* total_ms,
                           uint64_t* req_bytes);
int gh_engine_profile_units(gh_engine_t e, int kind, uint64_t* units);
int gh_engine_profile_alg_bytes(gh_engine_t e, int kind, uint64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* GENEVA_HOST_H */
