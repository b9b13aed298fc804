// Plain mining path: one query at a time, phase by phase, through the C ABI of
// libgeneva_hip (the role of src/gpumine.cpp:866-1072).
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../../include/geneva_hip.h"
#include "genMatchOrder.h"

namespace geneva {

// Vertices and labelled edges of a query pattern, parsed from .query text.
struct QueryPattern {
    uint32_t vnum = 0;
    std::vector<uint32_t> labels;
    std::vector<uint32_t> adj;  // vnum*vnum edge labels, 0xffffffff = none
};
bool parseQueryPattern(const std::string& text, QueryPattern& q, std::string* err);

// A planned query lowered to device step descriptors.
struct CompiledQuery {
    gv_query_t q{};
    bool empty = false;               // a label or edge label the graph lacks: count is 0
    std::vector<uint32_t> order;      // query vertex of each column
    uint32_t dvPhases = 0;            // two-vertex extension phases in the plan
    uint32_t nphases = 0;
    bool hasLeaves = false;           // trailing leaf groups (leafPlan) for this graph
    gv_leaves_t leaves{};
};

// Plan `text` with the DV planner (or the SV planner) and lower the plan: the
// column order is the order in which extension phases introduce vertices;
// each step lists every earlier query neighbour (extension source and the
// pairs of the reduction phases that follow), the restriction and the
// same-label columns.
bool planAndCompile(const std::string& text, bool singleV, gv_graph_t g, CompiledQuery& out,
                    std::string* err);

// Shortcut used when the planner output is already at hand.
bool compilePlan(const std::vector<Phase>& mo, const std::map<uint, uint>& restricts,
                 const std::map<uint, uint>& queryLabel, const QueryPattern& qp, gv_graph_t g,
                 CompiledQuery& out, std::string* err);

struct MineResult {
    int status = GV_OK;
    std::string error;
    uint64_t count = 0;
    uint64_t algBytes = 0;
    double ms = 0;               // device time, first launch to last completion
    uint64_t launches = 0;       // extension/count launches
    uint64_t chunks = 0;         // extension chunks (capacity splits)
    uint64_t peakRows = 0;       // largest materialised table (rows)
    uint64_t overflowRetries = 0;   // whole-query runs redone chunked
    bool finished = true;        // false: stopped at the deadline (counts are partial)
    double fraction = 1.0;       // share of the first extension's candidates mined
    std::vector<uint32_t> graphBits;
};

// The narrowest core (gv_leaves_t::width) whose remaining vertices form leaf
// groups that gv_count_leaves counts per core row; false if there is none.
bool leafPlan(const gv_query_t& q, uint32_t maxDeg, gv_leaves_t& d);

class PlainMiner {
public:
    // levelBytes: device memory per table level (rows + sizing arrays).
    PlainMiner(gv_ctx_t ctx, gv_graph_t g, uint64_t levelBytes);
    ~PlainMiner();
    PlainMiner(const PlainMiner&) = delete;
    PlainMiner& operator=(const PlainMiner&) = delete;

    int mine(const CompiledQuery& cq, MineResult& out);
    int mineText(const std::string& text, bool singleV, MineResult& out);

    // One shard of the start vertices: indices begin, begin+stride, ... below
    // end (0 = all) of the id-sorted list of vertices carrying the first
    // query vertex's label.  Counts over the shards of a partition add up to
    // the whole count; (0, 0, 1) is the whole query.
    struct Shard {
        uint64_t begin = 0, end = 0, stride = 1;
    };
    int mine(const CompiledQuery& cq, const Shard& sh, MineResult& out);

    // Count a last step with a single back edge per row (gv_count_rows)
    // instead of walking its candidates (on by default).
    void setCountByRows(bool on) { countByRows_ = on; }

    // Stop a query once `seconds` of wall time have passed (0 = no limit).
    // With a limit the query runs on the chunked path, whose launches are
    // bounded (a level's capacity, or kCountChunk candidates), and the host
    // waits for each launch; at the limit the query stops after the launch
    // in flight and MineResult holds the exact count of the embeddings whose
    // first-extension candidate lies in the part mined (finished = false,
    // fraction = that part's share of the first extension's candidates).
    void setTimeLimit(double seconds) { limitS_ = seconds; }

private:
    struct Level {
        uint32_t* rows = nullptr;   // table of width `col` (rows produced by step col-1)
        uint64_t cap = 0;
        gv_frontier_t fr = nullptr;  // sizing arrays for extending this table
        uint64_t frCap = 0;
    };
    int ensureFrontier(uint32_t width, uint64_t n);
    int ensureRows(uint32_t width);
    int run(const gv_query_t& q, uint32_t width, const uint32_t* rows, uint64_t n,
            MineResult& out);
    int runLevels(const CompiledQuery& cq, const uint32_t* rows, uint64_t n, MineResult& out);

    gv_ctx_t ctx_;
    gv_graph_t g_;
    uint64_t levelBytes_;
    Level lv_[GV_MAXK + 1];
    bool countByRows_ = true;
    double limitS_ = 0;
    double deadline_ = 0;           // steady-clock seconds; 0 = none
    uint64_t l1Total_ = 0, l1Done_ = 0;
    // counters after the last completed first-extension chunk (time limit)
    uint64_t snapCount_ = 0, snapBytes_ = 0;
    std::vector<uint32_t> snapBits_;
    bool timedOut() const;
    int snapshot();
    uint32_t* shardRows_ = nullptr;   // start table of a strided shard
    uint64_t shardCap_ = 0;
};

}  // namespace geneva
