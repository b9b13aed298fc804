// Database loader and on-disk formats of the labelled data graph.
//
// loadDataBase keeps the reference signature (include/loadDataBase.h:1):
// result[k] for k = 1..K is the k-th stored edge-label partition block (uint32
// array whose element 0 is its own length), result[0xffffffff] is vLabel
// (vLabel[0] = max vertex id, vLabel[id] = label).
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

typedef unsigned int uint;

std::map<uint, uint*> loadDataBase(std::string filename, std::string metafilename);

namespace geneva {

// Release every array returned by loadDataBase / buildDataBase.
void freeDataBase(std::map<uint, uint*>& db);

// One undirected edge of a data graph (vertex ids >= 1, edge label >= 1).
struct Edge {
    uint a, b, label;
};

// A labelled data graph held in memory: vertex ids 1..nv, vlabel[id] its label
// (vlabel[0] unused), edges listed once per undirected pair.
struct DataGraph {
    uint nv = 0;
    std::vector<uint> vlabel;  // size nv+1
    std::vector<Edge> edges;
    // original id of each vertex (size nv+1) after reorderForIntervals; empty
    // while ids are the original ones (written as the .metadata third column)
    std::vector<uint> oldId;
};

// New vertex ids under which every edge-label partition's source vertices
// form at most 256 runs of consecutive ids, the limit of the partition block
// (tools/genDatabase.cpp:451-460; the reference reorders ids for the same
// purpose, tools/genDatabase.cpp:259-413).  Each vertex's signature is the set
// of edge labels it has edges of; vertices are ordered by the rank of their
// signature read as a binary-reflected Gray code (bit order: the partition
// with the most sources in the highest bit), then by id, and isolated
// vertices come first.  In that order the i-th largest partition's bit
// (i = 0 for the largest) is set in at most max(1, 2^(i-1)) runs over all
// 2^K signatures, and dropping absent signatures only merges runs: always
// within 256 for K <= 10 partitions, and for any K when the signatures
// present are few.  newOf[old] = new id (newOf[0] = 0).  Returns the largest run count; err is set and 0
// returned when some partition would still need more than 256 runs or
// there are more than 64 edge labels.
uint reorderForIntervals(const DataGraph& g, std::vector<uint>& newOf, std::string* err);

// The graph under new ids (vertex labels and edges moved; oldId records the
// original ids).
DataGraph relabel(const DataGraph& g, const std::vector<uint>& newOf);

// Source runs (intervals) each edge label's partition block would need.
std::map<uint, uint> partitionRuns(const DataGraph& g);

// The partition block of one edge label in the .mygraph layout
// (tools/genDatabase.cpp:437-580): u0..u4 header, array0 (interval starts and
// cumulative lengths), array1 (per-source offsets), empty array2, array3
// (neighbours grouped by ascending neighbour label, ids ascending, with the
// two-level skip-index prefix on groups longer than 128,
// tools/genDatabase.cpp:415-436,741-743).  Sources are the vertices with at
// least one edge of this label; their ids must form at most 256 runs of
// consecutive ids (the reference drops runs past 256; this writer refuses).
// Returns a new[]-allocated array (element 0 = total uints), released by
// freeDataBase like loadDataBase's blocks.
uint* buildPartitionBlock(const DataGraph& g, uint edgeLabel, std::string* err);

// Same result as loadDataBase for an in-memory graph: one block per non-empty
// edge label in ascending label order, keyed 1..K, plus vLabel.  `keyOfLabel`
// (optional) receives the map edge label -> partition key.
std::map<uint, uint*> buildDataBase(const DataGraph& g, std::map<uint, uint>* keyOfLabel,
                                    std::string* err);

// Text and binary writers of the reference's file formats.
bool writeGraphText(const std::string& path, const DataGraph& g);         // .g
bool writeMetadata(const std::string& path, const DataGraph& g);          // .metadata
bool writeMyGraph(const std::string& path, const DataGraph& g, std::string* err);  // .mygraph
// .reorder (tools/genDatabase.cpp:826-845): label partitions 0..maxEdgeLabel,
// each source's neighbours grouped by label.  withSkipIndex reproduces the
// as-shipped file (skip-index prefixes inside the groups).
bool writeReorder(const std::string& path, const DataGraph& g, bool withSkipIndex);
// Parse a .g file (tools/genDatabase.cpp:771-800): 'v id label' lines, then
// 'e a b label' lines; 't' lines ignored.  Vertex ids are kept as given.
bool readGraphText(const std::string& path, DataGraph& g, std::string* err);

// Skip-index length of a label group whose stored length (prefix included) is
// L (src/miningPattern.cu:3289-3290).
inline uint skipIndexLen(uint L) {
    if (L > 130 && L <= 4128) return (L + 128) / 129;
    if (L > 4128) return 32;
    return 0;
}

// Decode one partition block back to per-source neighbour lists with the skip
// index removed: calls fn(source id, const uint* nbrs, count) for each source
// and label group in stored order.  Uses vLabel to find group boundaries.
// Returns false when the block is malformed.
bool forEachLabelGroup(const uint* block, const uint* vLabel, uint maxid,
                       void (*fn)(void* user, uint src, uint label, const uint* nbrs,
                                  uint count),
                       void* user, std::string* err);

}  // namespace geneva
