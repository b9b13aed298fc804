// Match-order planner: query pattern -> phase list, restrictions and caching flags.
//
// Same interface as the reference planner (include/genMatchOrder.h:3-10,
// include/genMatchOrderSingleV.h:3-10, include/calDVPhaseCount.h); the device
// upload of the first edge-label partition goes through the C-ABI
// (gv_memcpy_h2d) instead of cudaMemcpyAsync (src/genMatchOrder.cpp:766-769).
#pragma once
#include <cstdint>
#include <istream>
#include <map>
#include <string>
#include <vector>

typedef unsigned int uint;

// One phase of a match order (include/genMatchOrder.h:3-9).
//   phaseType 'E': pairs = (src, ext) or (src1, ext1, src2, ext2)
//   phaseType 'R': pairs = (a1, b1, a2, b2, ...) back edges to verify
class Phase {
public:
    char phaseType = 0;
    uint edgeLabel = 0;
    std::vector<uint> pairs;
    std::vector<uint8_t> reducePos;
};

// Single-vertex variant (include/genMatchOrderSingleV.h:3-9).
class PhaseSingleV {
public:
    char phaseType = 0;
    uint edgeLabel = 0;
    std::vector<uint> pairs;
    std::vector<uint8_t> reducePos;
};

// DV planner (src/genMatchOrder.cpp:671-996). When baseAddr_dev is non-null the
// first edge-label partition (minus its 5-uint header) is copied to
// baseAddr_dev+stride on the current upload context and stride is advanced.
void genMatchOrder(std::string queryFileName, std::vector<Phase>& matchOrder,
                   std::map<uint64_t, uint>& repeatFlag, std::map<uint, uint>& restricts,
                   std::map<uint, std::map<uint, bool>>& recordLabelStat,
                   std::map<uint, uint>& queryLabel,
                   std::map<uint, uint*>& allEdgeLabelPartitions, uint* baseAddr_dev,
                   uint& stride);

// SV planner (src/genMatchOrderSingleV.cpp:678+): every two-vertex extension is
// emitted as two single-vertex phases and every multi-pair reduction as one
// phase per pair.
void genMatchOrderSingleV(std::string queryFileName, std::vector<PhaseSingleV>& matchOrder,
                          std::map<uint64_t, uint>& repeatFlag,
                          std::map<uint, uint>& restricts,
                          std::map<uint, std::map<uint, bool>>& recordLabelStat,
                          std::map<uint, uint>& queryLabel,
                          std::map<uint, uint*>& allEdgeLabelPartitions, uint* baseAddr_dev,
                          uint& stride);

// Number of two-vertex ('E' with 4 pair entries) phases
// (tools/classifyExtPhase.cpp:27-38).
int calDVPhaseCount(const std::vector<Phase>& matchOrder);

namespace geneva {

// Error raised for malformed queries (more than 32 vertices, ids outside
// 0..vnum-1, a disconnected pattern).
struct PlanError {
    std::string what;
};

// Stream forms of the two planners; the file forms above open the file and call
// these. `upload` (may be null) receives (dst_dev, src_host, nbytes) for the
// partition copy. `skipDigitAndT` selects the line-skipping rule of the GPU
// planner ('#', 't', digit-led lines) instead of the CPU copy ('#' only,
// src/genMatchOrderCPU.cpp:691).
typedef void (*UploadFn)(void* dst_dev, const void* src_host, uint64_t nbytes);

void planDV(std::istream& in, bool skipDigitAndT, std::vector<Phase>& matchOrder,
            std::map<uint64_t, uint>& repeatFlag, std::map<uint, uint>& restricts,
            std::map<uint, std::map<uint, bool>>& recordLabelStat,
            std::map<uint, uint>& queryLabel, std::map<uint, uint*>* partitions,
            uint* baseAddr_dev, uint* stride, UploadFn upload);

void planSV(std::istream& in, bool skipDigitAndT, std::vector<PhaseSingleV>& matchOrder,
            std::map<uint64_t, uint>& repeatFlag, std::map<uint, uint>& restricts,
            std::map<uint, std::map<uint, bool>>& recordLabelStat,
            std::map<uint, uint>& queryLabel, std::map<uint, uint*>* partitions,
            uint* baseAddr_dev, uint* stride, UploadFn upload);

// Upload hook used by the file-form planners.
void setPlannerUpload(UploadFn fn);

}  // namespace geneva
