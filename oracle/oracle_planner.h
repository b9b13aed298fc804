// ORACLE (test infrastructure only): never linked into the product library.
//
// Restatement of the CPU copy of the reference planner
// (src/genMatchOrderCPU.cpp:1-983), kept separate from the product planner so
// that the two are checked against each other and against plan dumps of the
// reference's own cpumine.  Written as one function with local state.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace oracle {

struct OPhase {
    char type;                  // 'E' or 'R'
    unsigned edgeLabel;
    std::vector<unsigned> pairs;
};

struct OPlan {
    std::vector<OPhase> phases;
    std::map<unsigned, unsigned> restricts;   // key -> value, 0xffffffff = none
    std::map<unsigned, unsigned> queryLabel;  // vertex -> label
    std::vector<unsigned> adj;                // vnum*vnum edge labels, 0xffffffff = none
    unsigned vnum = 0;
    size_t nAut = 0;                          // automorphisms found (identity excluded)
};

// Parse a .query text and plan it.  Returns false (with err) on malformed input.
bool planQuery(const std::string& text, OPlan& plan, std::string& err);

// Text dump in the layout the reference cpumine prints
// (src/cpumine.cpp:661-674): "restrict: k<v" lines then "phase i type T edge
// label:L" + "svid:a evid:b " lines.
std::string dumpPlan(const OPlan& plan);

}  // namespace oracle
