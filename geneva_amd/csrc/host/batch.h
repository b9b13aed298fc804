// Batched and served mining over the batch engine of libgeneva_hip.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../../include/geneva_hip.h"
#include "miner.h"

namespace geneva {

// Vertex and edge counts by label, per edge-label partition.
struct GraphStats {
    uint32_t nlabels = 0;
    std::vector<uint64_t> vertices;                 // by dense label
    std::vector<std::vector<uint64_t>> pairEdges;   // [part][a*nl+b]
    static GraphStats of(gv_graph_t g);
};

// Estimate of the largest row table the plan materialises, in bytes: the start
// table is the label's vertex count, each extension multiplies by the average
// number of neighbours of the new vertex's label seen from its sparsest back
// neighbour's label.  The count-only last step (and a single-edge last step
// fused into the one before it, or trailing leaf groups counted on their core)
// materialises nothing.
uint64_t estimatePeakBytes(const CompiledQuery& cq, const GraphStats& st);

struct BatchOptions {
    uint32_t slots = 256;
    uint64_t slotBytes = 64ull << 20;   // per table, two tables per slot
    uint64_t rowCap = 1ull << 28;       // rows sized per iteration, whole batch
    gv_tune_t tune{4, 16};
};

struct BatchStats {
    uint64_t batched = 0;       // finished on the batch engine
    uint64_t overflowed = 0;    // re-run on the plain path
    uint64_t graphLaunches = 0;
    double deviceMs = 0;        // batch iterations, device time
    double wallMs = 0;
    double fallbackMs = 0;      // wall time of the plain-path re-runs
    uint64_t pieces = 0;        // splits of overflowed queries into pieces
};

class BatchMiner {
public:
    BatchMiner(gv_ctx_t ctx, gv_graph_t g, PlainMiner* fallback, const BatchOptions& opt);
    ~BatchMiner();
    BatchMiner(const BatchMiner&) = delete;
    BatchMiner& operator=(const BatchMiner&) = delete;

    int init();
    int setTune(const gv_tune_t& t);
    // Rows sized per iteration (at most options().rowCap, the allocation).
    int setRowCap(uint64_t rows);
    uint64_t rowCap() const { return rowCap_ ? rowCap_ : opt_.rowCap; }
    const BatchOptions& options() const { return opt_; }
    gv_batch_t handle() const { return b_; }
    gv_graph_t graph() const { return g_; }
    PlainMiner* fallback() const { return fallback_; }
    // words of a matched-graph bitset (0 without an owner table); valid after init()
    uint32_t graphWords() const { return nwords_; }

    // Mine every query; results in input order.  Queries with a plan error
    // get status GV_ERR_ARG; overflowing ones are re-run on the plain path.
    int run(const std::vector<CompiledQuery>& qs, std::vector<MineResult>& out, BatchStats* st);

    // Place a query (or piece `part` of `nparts` of its first table) in a
    // free slot with its class's candidates per lane (classTune);
    // GV_ERR_CAPACITY when no slot is free.
    int admit(const CompiledQuery& cq, uint64_t tag, uint32_t* slot, uint32_t part = 0,
              uint32_t nparts = 1);

    // Query class of the per-class tuning: vertices, steps with more than
    // one back edge, trailing leaf groups.
    static std::string classKey(const CompiledQuery& cq);
    // candidates per lane per walk tile of each class (0 or absent: the
    // batch's own)
    std::map<std::string, uint32_t>& classTune() { return classIpt_; }

private:
    gv_ctx_t ctx_;
    gv_graph_t g_;
    PlainMiner* fallback_;
    BatchOptions opt_;
    gv_batch_t b_ = nullptr;
    uint32_t nwords_ = 0;
    uint64_t rowCap_ = 0;
    std::map<std::string, uint32_t> classIpt_;
    // pieces a query needed the last time it outgrew its slot (keyed by its
    // compiled steps): later runs admit it in that many pieces at once
    std::map<std::string, uint32_t> pieceHint_;
};

// Autotuner: times the batch engine under each candidate tuning on a sample of
// queries and keeps the fastest (candidates per lane of the walk kernel x its
// blocks per CU; the rows sized per iteration stay the allocation's), then,
// for each query class of the sample with at least four queries, times that class
// alone under each candidates-per-lane value up to the chosen one and keeps
// the fastest for the class (slots of the class run their walk tiles at that
// size).  Remembered per graph, engine options and sample shape.
struct TuneResult {
    gv_tune_t tune{4, 16};
    uint64_t rowCap = 0;
    double bestMs = 0;
    std::vector<std::pair<gv_tune_t, double>> tried;
    // per query class: the candidates per lane chosen and its device ms
    std::map<std::string, std::pair<uint32_t, double>> classes;
};
int autotune(BatchMiner& bm, const std::vector<CompiledQuery>& sample, TuneResult& out);

// Serving: requests with arrival times (seconds from the start) are mined as
// they arrive.  Batched: every loop turn admits the requests that have
// arrived, while slots are free and the summed peak-table estimate of the
// running requests stays under memBudget, then runs one iteration, in which
// every running request advances by one vertex; requests leave as they
// finish.  Requests that do not go through the batch engine (one that arrives
// alone while nothing runs, one whose estimate exceeds a slot table, one
// whose table overflowed its slot) are mined on the plain path: by `side`,
// a PlainMiner on its own context (own HIP stream) driven by a worker thread,
// so admission and iterations go on meanwhile; without `side`, inline by the
// engine's fallback.  Plain: requests are mined one at a time in arrival
// order.
struct ServeResult {
    std::vector<double> latencyMs;   // arrival to result, per request
    std::vector<uint64_t> counts;
    std::vector<uint32_t> graphBits; // nwords per request (matched graph ids)
    uint32_t nwords = 0;
    double wallS = 0;                // first arrival to last result
    uint64_t iterations = 0;
    uint64_t maxConcurrent = 0;      // requests in flight (batch engine + plain worker)
    uint64_t maxBatched = 0;         // requests on the batch engine at once
    uint64_t overflowed = 0;     // left the batch engine for the plain path
    uint64_t direct = 0;         // went to the plain path without entering the batch engine
};
int serveBatched(BatchMiner& bm, const std::vector<CompiledQuery>& reqs,
                 const std::vector<double>& arrivalS, uint64_t memBudget, const GraphStats& gs,
                 ServeResult& out, PlainMiner* side = nullptr);
int servePlain(PlainMiner& pm, const std::vector<CompiledQuery>& reqs,
               const std::vector<double>& arrivalS, ServeResult& out);

}  // namespace geneva
