// The reference's phase-by-phase driver over the reference-named launchers
// (include/geneva_launchers.hpp): loadDataBase, the reference planner, one
// device arena, a counter read-back after every launch, overflow rows moved
// to host memory and the launch resumed with isContinue.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/geneva_hip.h"

namespace geneva {

struct PhaseDriverOptions {
    uint64_t arenaBytes = 1ull << 30;   // the driver's device arena
    uint32_t maxRows = 0;               // cap on maxRowNum (0 = the reference's bound only)
    bool keepRows = true;               // return the final rows
};

struct PhaseRunResult {
    uint64_t count = 0;                 // final rows, spilled parts included
    uint32_t width = 0;                 // columns of a final row
    std::vector<uint32_t> order;        // query vertex of each column
    std::vector<uint32_t> rows;         // count * width (keepRows)
    uint64_t launches = 0;              // launcher calls (one counter read-back each)
    uint64_t spills = 0;                // tables or parts moved to host memory
    bool singleV = false;               // the plan kind that ran
    std::string error;
};

// Mine the query in queryFile over the .mygraph / .metadata pair with the DV
// plan (planKind 0, genMatchOrder), the SV plan (1, genMatchOrderSingleV) or
// the kind chosen from the DV-phase count (-1, choosePlanKind).
int runPhaseDriver(gv_ctx_t ctx, const std::string& mygraph, const std::string& metadata,
                   const std::string& queryFile, int planKind, const PhaseDriverOptions& opt,
                   PhaseRunResult& out);

// The plan kind for a query (tools/classifyExtPhase.cpp:27-38 classifies
// queries by calDVPhaseCount): DV when its DV plan has a two-vertex phase,
// SV otherwise.
int choosePlanKind(const std::string& queryFile, bool& singleV, std::string* err);

}  // namespace geneva
