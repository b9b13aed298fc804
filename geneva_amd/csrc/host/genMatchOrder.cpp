// Match-order planner restated for the host library.
//
// The algorithm is the reference's, step for step, because the phase order and
// the restriction set it yields fix the embedding count (a restriction
// f(key) < f(value) removes automorphic copies).  Tie-breaking follows the same
// containers: std::map iteration order and libstdc++ std::sort with the same
// comparators on the same input order.  What differs from the reference:
//   * planner state lives in one object per call instead of file-scope globals
//     (src/genMatchOrder.cpp:60-65), so many queries can be planned in one
//     process;
//   * the partition upload goes through an upload hook (C-ABI) rather than
//     cudaMemcpyAsync (src/genMatchOrder.cpp:766-769,805-808,820-823);
//   * the single-vertex planner records the index of the reduction phase it
//     just pushed (the reference records one past it,
//     src/genMatchOrderSingleV.cpp:366-388, which lets the later orientation
//     pass swap the pairs of an unrelated phase or index past the end).
#include "genMatchOrder.h"

#include <algorithm>
#include <fstream>
#include <list>

namespace {

struct QueryNeigInfo {  // src/genMatchOrder.cpp:14-26
    uint neigs = 0;     // bit n set <=> vertex n is a neighbour
    // edge label -> {number of neighbours, neighbour bitmask}
    std::map<uint, std::vector<uint>> edgeLabelNeigs;
};

struct QueryEdge {  // src/genMatchOrder.cpp:27-32
    uint svid, evid, edgeLabel;
};

struct TriInfo {  // src/genMatchOrder.cpp:34-40 (value-initialised)
    uint edgeIndex = 0, triCount = 0, triScore = 0, maxScoreVID = 0;
};

struct NeigInfo {  // src/genMatchOrder.cpp:41-49
    uint commonVNum = 0, svid = 0, evid = 0, edgeLabel = 0, evLabel = 0, evNum = 0;
};

struct PhaseKeyCount {  // src/genMatchOrder.cpp:50-57
    uint64_t key = 0;
    uint count = 0;
    std::vector<uint64_t> keyList;
    std::vector<uint> phaseList;  // (phase index, pair index) pairs
};

// src/genMatchOrder.cpp:66-68
bool sortDescend(const TriInfo& a, const TriInfo& b) { return a.triCount > b.triCount; }

// src/genMatchOrder.cpp:126-150: more back edges, then more same-edge-label
// neighbours, then smaller edge label, source id, end-vertex label.
bool multiLevelSort(const NeigInfo& a, const NeigInfo& b) {
    if (a.commonVNum != b.commonVNum) return a.commonVNum > b.commonVNum;
    if (a.evNum != b.evNum) return a.evNum > b.evNum;
    if (a.edgeLabel != b.edgeLabel) return a.edgeLabel < b.edgeLabel;
    if (a.svid != b.svid) return a.svid < b.svid;
    return a.evLabel < b.evLabel;
}

// src/genMatchOrder.cpp:897-904
bool byEdgeThenEndLabel(const NeigInfo& a, const NeigInfo& b) {
    if (a.edgeLabel != b.edgeLabel) return a.edgeLabel < b.edgeLabel;
    return a.evLabel < b.evLabel;
}

void addRecordLabel(std::map<uint, std::map<uint, bool>>& stat, uint64_t key) {
    uint vlabel = (uint)(key & 0xffffffffull);
    uint edgeLabel = (uint)((key >> 32) & 0xffffffffull);
    auto it = stat.find(edgeLabel);
    if (it == stat.end()) {
        stat[edgeLabel][vlabel] = true;
    } else if (it->second.find(vlabel) == it->second.end()) {
        it->second[vlabel] = true;
    }
}

template <class PhaseT, bool SV>
class Planner {
public:
    Planner(std::vector<PhaseT>& mo, std::map<uint64_t, uint>& rf, std::map<uint, uint>& rs,
            std::map<uint, std::map<uint, bool>>& rls, std::map<uint, uint>& ql)
        : matchOrder(mo), repeatFlag(rf), restricts(rs), recordLabelStat(rls), queryLabel(ql) {}

    void run(std::istream& in, bool skipDigitAndT, std::map<uint, uint*>* partitions,
             uint* baseAddr_dev, uint* stride, geneva::UploadFn upload);

private:
    std::vector<PhaseT>& matchOrder;
    std::map<uint64_t, uint>& repeatFlag;
    std::map<uint, uint>& restricts;
    std::map<uint, std::map<uint, bool>>& recordLabelStat;
    std::map<uint, uint>& queryLabel;

    uint vnum = 0;
    uint partialQueryVidMap = 0;
    std::list<uint> partialQueryVid;
    std::vector<int> vidToIndex;
    std::vector<uint> vertexNeig;
    std::vector<uint> adjMatrix;
    std::vector<QueryNeigInfo> query;
    std::vector<QueryEdge> edges;
    std::map<uint64_t, uint> extendPhaseKeyCount;
    std::map<uint64_t, PhaseKeyCount> reducePhaseKeyCount;
    std::map<uint64_t, uint> othersKeyCount;

    uint ql(uint v) { return queryLabel[v]; }
    uint adj(uint a, uint b) const { return adjMatrix[a * vnum + b]; }

    void addToQueryGraph(uint svid, uint evid, uint elabel);
    void countTriangle(uint edgeIndex, std::vector<TriInfo>& out);
    void initialNoTriangle(uint& svid, uint& evid1, uint& evid2);
    void visit(uint src, uint ext);
    void bumpExtendKey(uint64_t key);
    void addExtPhase(uint svid1, uint extvid1, uint svid2, uint extvid2, uint sameEdgeLabelNum);
    void recordReducePair(uint svid1, uint extvid1, uint phaseIndex, uint pairIndex);
    void addReducePhase(const std::vector<uint>& pairs);
    bool isEquivalent(uint a, uint b);
    bool isPossibleAut(const std::vector<uint>& cand, uint newID);
    void genAutomorphism(std::vector<uint> cand, bool isIdentity,
                         std::vector<std::vector<uint>>& out);
    void genRestriction(std::vector<std::vector<uint>> automorphisms);
    void rearrangeReducePhase();
    void uploadPartition(uint edgelabel, std::map<uint, uint*>* partitions, uint* baseAddr_dev,
                         uint* stride, geneva::UploadFn upload);
};

// src/genMatchOrder.cpp:70-86
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::addToQueryGraph(uint svid, uint evid, uint elabel) {
    query[svid].neigs |= (1u << evid);
    auto& neig = query[svid].edgeLabelNeigs;
    auto it = neig.find(elabel);
    if (it != neig.end()) {
        it->second[0] += 1;
        it->second[1] |= (1u << evid);
    } else {
        neig.emplace(elabel, std::vector<uint>{1u, 1u << evid});
    }
}

// src/genMatchOrder.cpp:89-116: triangles through an edge, scored by how many
// edge labels (x10) and vertex labels agree; the best-scoring third vertex wins.
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::countTriangle(uint edgeIndex, std::vector<TriInfo>& out) {
    const QueryEdge& e = edges[edgeIndex];
    uint tri = vertexNeig[e.svid] & vertexNeig[e.evid];
    out[edgeIndex].edgeIndex = edgeIndex;
    out[edgeIndex].triCount = (uint)__builtin_popcount(tri);
    if (tri == 0) return;
    int prei = -1, prescore = -1;
    for (uint i = 0; i < 32; ++i) {
        if (((1u << i) & tri) == 0) continue;
        uint e1l = adj(e.svid, i), e2l = adj(e.evid, i), e3l = e.edgeLabel;
        int edgescore = 10 * ((e1l == e2l) + (e1l == e3l) + (e2l == e3l));
        uint v1l = ql(e.svid), v2l = ql(e.evid), v3l = ql(i);
        int vertexscore = (v1l == v2l) + (v1l == v3l) + (v2l == v3l);
        if (vertexscore + edgescore > prescore) {
            prescore = vertexscore + edgescore;
            prei = (int)i;
        }
    }
    out[edgeIndex].triScore = (uint)prescore;
    out[edgeIndex].maxScoreVID = (uint)prei;
}

// src/genMatchOrder.cpp:177-286: start vertex and one or two first extensions
// for a triangle-free query.
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::initialNoTriangle(uint& svid, uint& evid1, uint& evid2) {
    uint globalSvid = 0xffffffffu, globalFirstMax = 0, globalSecondMax = 0;
    uint globalFirstEndVid = 0xffffffffu, globalSecondEndVid = 0xffffffffu;
    std::map<uint64_t, std::vector<uint>> sameEdgeEvidLabel, sameEdgeLabel, sameSvid;
    for (uint i = 0; i < vnum; ++i) {
        uint64_t keysvid = i;
        for (auto& kv : query[i].edgeLabelNeigs) {
            uint64_t keysvidedge = ((uint64_t)kv.first << 20) | keysvid;
            uint neigs = kv.second[1];
            if (__builtin_popcount(neigs) >= 2) {
                std::map<uint, std::vector<uint>> byLabel;
                for (uint k = 0; k < 32; ++k)
                    if (neigs & (1u << k)) byLabel[ql(k)].push_back(k);
                for (auto& lv : byLabel) {
                    uint64_t key3 = ((uint64_t)lv.first << 40) | keysvidedge;
                    if (lv.second.size() >= 2) {
                        sameEdgeEvidLabel.insert(std::make_pair(key3, lv.second));
                    } else {
                        auto& v = sameEdgeLabel[keysvidedge];
                        v.insert(v.end(), lv.second.begin(), lv.second.end());
                    }
                }
            } else {
                for (uint k = 0; k < 32; ++k)
                    if (neigs & (1u << k)) sameSvid[keysvid].push_back(k);
            }
        }
    }
    const bool pairMode = !sameEdgeEvidLabel.empty() || !sameEdgeLabel.empty();
    const std::map<uint64_t, std::vector<uint>>& pick =
        !sameEdgeEvidLabel.empty() ? sameEdgeEvidLabel
                                   : (!sameEdgeLabel.empty() ? sameEdgeLabel : sameSvid);
    for (auto& kv : pick) {
        uint firstMaximum = 0, secondMaximum = 0;
        uint firstEndVid = 0xffffffffu, secondEndVid = 0xffffffffu;
        for (uint evid : kv.second) {
            uint tot = (uint)__builtin_popcount(query[evid].neigs);
            if (tot > firstMaximum) {
                secondMaximum = firstMaximum;
                secondEndVid = firstEndVid;
                firstMaximum = tot;
                firstEndVid = evid;
            } else if (tot > secondMaximum) {
                secondMaximum = tot;
                secondEndVid = evid;
            }
        }
        if (pairMode) {
            if (firstMaximum + secondMaximum > globalFirstMax + globalSecondMax) {
                globalFirstMax = firstMaximum;
                globalFirstEndVid = firstEndVid;
                globalSecondMax = secondMaximum;
                globalSecondEndVid = secondEndVid;
                globalSvid = (uint)(kv.first & 0xfffffull);
            }
        } else if (firstMaximum > globalFirstMax) {
            globalFirstMax = firstMaximum;
            globalFirstEndVid = firstEndVid;
            globalSvid = (uint)(kv.first & 0xfffffull);
        }
    }
    svid = globalSvid;
    evid1 = globalFirstEndVid;
    evid2 = globalSecondEndVid;
}

template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::visit(uint src, uint ext) {
    partialQueryVidMap |= (1u << ext);
    vidToIndex[ext] = (int)partialQueryVid.size();
    partialQueryVid.push_back(ext);
    vertexNeig[src] &= ~(1u << ext);
    vertexNeig[ext] &= ~(1u << src);
}

template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::bumpExtendKey(uint64_t key) {
    auto it = extendPhaseKeyCount.find(key);
    if (it == extendPhaseKeyCount.end())
        extendPhaseKeyCount.emplace(key, 1u);
    else
        it->second += 1;
}

// src/genMatchOrder.cpp:288-334 (DV) and src/genMatchOrderSingleV.cpp:288-337
// (SV: the second pair becomes its own phase).
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::addExtPhase(uint svid1, uint extvid1, uint svid2, uint extvid2,
                                      uint sameEdgeLabelNum) {
    if (sameEdgeLabelNum == 2 && ql(extvid2) < ql(extvid1)) {
        std::swap(svid1, svid2);
        std::swap(extvid1, extvid2);
    }
    PhaseT p1;
    p1.phaseType = 'E';
    p1.pairs = {svid1, extvid1};
    p1.edgeLabel = adj(svid1, extvid1);
    if (SV) matchOrder.push_back(p1);
    bumpExtendKey(((uint64_t)p1.edgeLabel << 32) | ql(extvid1));
    visit(svid1, extvid1);
    if (sameEdgeLabelNum == 2) {
        uint64_t key = ((uint64_t)adj(svid2, extvid2) << 32) | ql(extvid2);
        bumpExtendKey(key);
        visit(svid2, extvid2);
        if (SV) {
            PhaseT p2;
            p2.phaseType = 'E';
            p2.edgeLabel = p1.edgeLabel;
            p2.pairs = {svid2, extvid2};
            matchOrder.push_back(p2);
        } else {
            p1.pairs.push_back(svid2);
            p1.pairs.push_back(extvid2);
        }
    }
    if (!SV) matchOrder.push_back(p1);
}

// Key bookkeeping shared by both reduction forms (src/genMatchOrder.cpp:347-390).
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::recordReducePair(uint svid1, uint extvid1, uint phaseIndex,
                                           uint pairIndex) {
    uint64_t el = adj(svid1, extvid1);
    uint64_t key = (el << 32) | ql(extvid1);
    uint64_t key1 = (el << 32) | ql(svid1);
    if (key == key1) {
        auto it = othersKeyCount.find(key);
        if (it == othersKeyCount.end())
            othersKeyCount.emplace(key, 1u);
        else
            it->second += 1;
        return;
    }
    auto note = [&](uint64_t k, uint64_t other) {
        auto it = reducePhaseKeyCount.find(k);
        if (it == reducePhaseKeyCount.end()) {
            PhaseKeyCount t;
            t.key = k;
            t.count = 1;
            t.keyList.push_back(other);
            t.phaseList.push_back(phaseIndex);
            t.phaseList.push_back(pairIndex);
            reducePhaseKeyCount.emplace(k, t);
        } else {
            it->second.count += 1;
            it->second.keyList.push_back(other);
            it->second.phaseList.push_back(phaseIndex);
            it->second.phaseList.push_back(pairIndex);
        }
    };
    note(key, key1);
    note(key1, key);
}

// src/genMatchOrder.cpp:335-395 (one phase for all pairs) and
// src/genMatchOrderSingleV.cpp:339-401 (one phase per pair).
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::addReducePhase(const std::vector<uint>& pairs) {
    if (SV) {
        for (uint i = 0; i < pairs.size() / 2; ++i) {
            uint a = pairs[i * 2], b = pairs[i * 2 + 1];
            PhaseT r;
            r.phaseType = 'R';
            r.reducePos = {1, 1};
            r.pairs = {a, b};
            r.edgeLabel = adj(a, b);
            matchOrder.push_back(r);
            vertexNeig[a] &= ~(1u << b);
            vertexNeig[b] &= ~(1u << a);
            recordReducePair(a, b, (uint)matchOrder.size() - 1, 0);
        }
        return;
    }
    PhaseT r;
    r.phaseType = 'R';
    r.reducePos.push_back(1);
    for (uint i = 0; i < pairs.size() / 2; ++i) {
        uint a = pairs[i * 2], b = pairs[i * 2 + 1];
        r.pairs.push_back(a);
        r.pairs.push_back(b);
        vertexNeig[a] &= ~(1u << b);
        vertexNeig[b] &= ~(1u << a);
        r.edgeLabel = adj(a, b);
        recordReducePair(a, b, (uint)matchOrder.size(), i);
    }
    r.reducePos.push_back((uint8_t)(pairs.size() / 2));
    matchOrder.push_back(r);
}

// src/genMatchOrder.cpp:397-415
template <class PhaseT, bool SV>
bool Planner<PhaseT, SV>::isEquivalent(uint a, uint b) {
    if (a == b) return true;
    if (ql(a) != ql(b)) return false;
    auto& na = query[a].edgeLabelNeigs;
    auto& nb = query[b].edgeLabelNeigs;
    if (na.size() != nb.size()) return false;
    for (auto ia = na.begin(), ib = nb.begin(); ia != na.end(); ++ia, ++ib)
        if (ia->first != ib->first || ia->second[0] != ib->second[0]) return false;
    return true;
}

// src/genMatchOrder.cpp:418-432: candidate position cand.size() maps to newID.
template <class PhaseT, bool SV>
bool Planner<PhaseT, SV>::isPossibleAut(const std::vector<uint>& cand, uint newID) {
    uint ori = (uint)cand.size();
    if (!isEquivalent(ori, newID)) return false;
    for (uint i = 0; i < cand.size(); ++i)
        if (adj(i, ori) != adj(cand[i], newID)) return false;
    return true;
}

// src/genMatchOrder.cpp:434-463: every label- and edge-label-preserving
// permutation except the identity.
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::genAutomorphism(std::vector<uint> cand, bool isIdentity,
                                          std::vector<std::vector<uint>>& out) {
    auto used = [&](uint v) {
        for (uint x : cand)
            if (x == v) return true;
        return false;
    };
    if (cand.size() == vnum - 1) {
        for (uint i = 0; i < vnum; ++i) {
            if (used(i)) continue;
            if (isIdentity && i == cand.size()) return;
            if (isPossibleAut(cand, i)) {
                cand.push_back(i);
                out.push_back(cand);
            }
            break;
        }
        return;
    }
    for (uint i = 0; i < vnum; ++i) {
        if (used(i)) continue;
        if (isPossibleAut(cand, i)) {
            cand.push_back(i);
            bool idt = isIdentity && (i == cand.size() - 1);
            genAutomorphism(cand, idt, out);
            cand.pop_back();
        }
    }
}

// src/genMatchOrder.cpp:468-507: walk the extension pairs in order; each
// vertex moved by a remaining automorphism gets restriction
// restricts[aut[v]] = max(query id) v, then only the stabiliser of v remains.
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::genRestriction(std::vector<std::vector<uint>> automorphisms) {
    std::vector<bool> isVisited(vnum, false);
    for (auto& ph : matchOrder) {
        if (ph.phaseType == 'R') continue;
        for (uint vid : ph.pairs) {
            if (isVisited[vid]) continue;
            isVisited[vid] = true;
            std::vector<std::vector<uint>> stabilized;
            for (auto& aut : automorphisms) {
                if (aut[vid] == vid) {
                    stabilized.push_back(aut);
                } else {
                    uint key = aut[vid];
                    auto it = restricts.find(key);
                    if (it == restricts.end())
                        restricts.emplace(key, vid);
                    else if (it->second < vid)
                        it->second = vid;
                }
            }
            automorphisms.swap(stabilized);
        }
    }
    for (uint i = 0; i <= vnum; ++i)
        if (restricts.find(i) == restricts.end()) restricts.emplace(i, 0xffffffffu);
}

// src/genMatchOrder.cpp:509-668: orientation of reduction pairs and the
// caching flags (repeatFlag, recordLabelStat).  The reference's running
// maximum never updates (`precount = precount`, :528), so each round selects
// the last key (in map order) whose combined count is positive; every stored
// key keeps count >= 1 until it is erased, so that is the last key.
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::rearrangeReducePhase() {
    uint itercount = (uint)reducePhaseKeyCount.size();
    for (uint j = 0; j < itercount; ++j) {
        auto maxite = reducePhaseKeyCount.end();
        for (auto it = reducePhaseKeyCount.begin(); it != reducePhaseKeyCount.end(); ++it) {
            uint64_t key = it->first;
            uint count = it->second.count;
            auto e = extendPhaseKeyCount.find(key);
            if (e != extendPhaseKeyCount.end()) count += e->second;
            auto o = othersKeyCount.find(key);
            if (o != othersKeyCount.end()) count += o->second;
            if (count > 0) maxite = it;
        }
        if (maxite == reducePhaseKeyCount.end()) break;  // unreachable: counts stay >= 1
        uint64_t key = maxite->first;
        PhaseKeyCount& kc = maxite->second;
        for (size_t i = 0; i < kc.keyList.size(); ++i)
            if (kc.count > 0) kc.count -= 1;
        for (size_t i = 0; i < kc.phaseList.size() / 2; ++i) {
            uint phaseIndex = kc.phaseList[i * 2];
            uint pairIndex = kc.phaseList[i * 2 + 1];
            if (phaseIndex >= matchOrder.size()) continue;
            std::vector<uint>& pairs = matchOrder[phaseIndex].pairs;
            if (pairIndex * 2 + 1 >= pairs.size()) continue;
            uint vid1 = pairs[pairIndex * 2], vid2 = pairs[pairIndex * 2 + 1];
            if ((uint64_t)ql(vid1) == (key & 0xffffffffull)) {
                pairs[pairIndex * 2] = vid2;
                pairs[pairIndex * 2 + 1] = vid1;
            }
        }
        if (repeatFlag.find(key) == repeatFlag.end()) {
            uint c = kc.count;
            repeatFlag.emplace(key, (c << 16) | c);
            if (c > 1) addRecordLabel(recordLabelStat, key);
        } else {
            throw geneva::PlanError{"repeated key in reduction orientation"};
        }
        reducePhaseKeyCount.erase(key);
    }
    auto merge = [&](const std::map<uint64_t, uint>& src) {
        for (auto& kv : src) {
            auto it = repeatFlag.find(kv.first);
            if (it == repeatFlag.end()) {
                uint c = kv.second;
                repeatFlag.emplace(kv.first, (c << 16) | c);
                if (c > 1) addRecordLabel(recordLabelStat, kv.first);
            } else {
                it->second += kv.second;
                if (it->second > 1) addRecordLabel(recordLabelStat, kv.first);
            }
        }
    };
    merge(extendPhaseKeyCount);
    merge(othersKeyCount);
}

template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::uploadPartition(uint edgelabel, std::map<uint, uint*>* partitions,
                                          uint* baseAddr_dev, uint* stride,
                                          geneva::UploadFn upload) {
    if (!baseAddr_dev || !partitions || !stride || !upload) return;
    auto it = partitions->find(edgelabel);
    if (it == partitions->end() || it->second == nullptr) return;
    uint* part = it->second;
    uint len = part[0] - 5;
    upload(baseAddr_dev + *stride, part + 5, (uint64_t)len * sizeof(uint));
    *stride += len;
}

// src/genMatchOrder.cpp:671-996
template <class PhaseT, bool SV>
void Planner<PhaseT, SV>::run(std::istream& in, bool skipDigitAndT,
                              std::map<uint, uint*>* partitions, uint* baseAddr_dev,
                              uint* stride, geneva::UploadFn upload) {
    // Vertex lines until the first 'e' (src/genMatchOrder.cpp:681-698).
    char type;
    while (in.get(type)) {
        if (type == 'v') {
            uint vid, vlabel;
            in >> vid >> vlabel;
            if (queryLabel.find(vid) == queryLabel.end()) queryLabel.emplace(vid, vlabel);
        } else if (type == 'e') {
            break;
        } else if (type == '#' || (skipDigitAndT && (type == 't' || (type >= '0' && type <= '9')))) {
            char tmpc[100];
            in.getline(tmpc, 100);
            if (in.fail() && !in.eof()) in.clear();  // over-long comment: keep reading
        }
    }
    if (in) in.unget();
    else in.clear();
    vnum = (uint)queryLabel.size();
    if (vnum < 2 || vnum > 32) throw geneva::PlanError{"query must have 2..32 vertices"};
    for (auto& kv : queryLabel)
        if (kv.first >= vnum) throw geneva::PlanError{"query vertex ids must be 0..vnum-1"};
    query.assign(vnum, QueryNeigInfo());
    adjMatrix.assign((size_t)vnum * vnum, 0xffffffffu);
    vertexNeig.assign(vnum, 0);
    while (in.get(type)) {
        if (type != 'e') continue;
        uint svid, evid, elabel;
        if (!(in >> svid >> evid >> elabel)) break;
        if (svid >= vnum || evid >= vnum || svid == evid)
            throw geneva::PlanError{"bad query edge"};
        addToQueryGraph(svid, evid, elabel);
        addToQueryGraph(evid, svid, elabel);
        if (adj(svid, evid) == 0xffffffffu) {
            QueryEdge e;
            e.svid = std::min(svid, evid);
            e.evid = std::max(svid, evid);
            e.edgeLabel = elabel;
            edges.push_back(e);
            adjMatrix[svid * vnum + evid] = elabel;
            adjMatrix[evid * vnum + svid] = elabel;
        }
        vertexNeig[svid] |= (1u << evid);
        vertexNeig[evid] |= (1u << svid);
    }
    if (edges.empty()) throw geneva::PlanError{"query has no edges"};
    {  // connectivity (the reference assumes it)
        uint seen = 1, frontier = 1;
        while (frontier) {
            uint nxt = 0;
            for (uint v = 0; v < vnum; ++v)
                if (frontier & (1u << v)) nxt |= vertexNeig[v];
            nxt &= ~seen;
            seen |= nxt;
            frontier = nxt;
        }
        uint all = vnum == 32 ? 0xffffffffu : ((1u << vnum) - 1);
        if (seen != all) throw geneva::PlanError{"query is not connected"};
    }

    std::vector<TriInfo> tri(edges.size());
    for (uint i = 0; i < edges.size(); ++i) countTriangle(i, tri);
    std::sort(tri.begin(), tri.end(), sortDescend);
    vidToIndex.assign(vnum, -1);
    if (tri[0].triCount == 0) {
        uint svid, e1, e2;
        initialNoTriangle(svid, e1, e2);
        uploadPartition(adj(svid, e1), partitions, baseAddr_dev, stride, upload);
        vidToIndex[svid] = 0;
        partialQueryVid.push_back(svid);
        partialQueryVidMap |= (1u << svid);
        addExtPhase(svid, e1, svid, e2, e2 == 0xffffffffu ? 1 : 2);
    } else {
        uint tricount = tri[0].triCount;
        uint i = 1, prei = 0;
        while (i < tri.size() && tri[i].triCount == tricount) {
            if (tri[i].triScore > tri[prei].triScore) prei = i;
            ++i;
        }
        uint score = tri[prei].triScore;
        uint ei = tri[prei].edgeIndex;
        uint svid = edges[ei].svid, evid = edges[ei].evid, third = tri[prei].maxScoreVID;
        uint src = svid, x1 = evid, x2 = third;
        if (score == 33 || score == 31 || score == 30 || (score >= 10 && score <= 13)) {
            if (score == 31) {
                if (ql(svid) == ql(evid)) { src = third; x1 = svid; x2 = evid; }
                else if (ql(svid) == ql(third)) { src = evid; x1 = svid; x2 = third; }
                else if (ql(evid) == ql(third)) { src = svid; x1 = evid; x2 = third; }
            } else if (score >= 10 && score <= 13) {
                uint e1l = adj(svid, third), e2l = adj(evid, third), e3l = edges[ei].edgeLabel;
                if (e1l == e2l) { src = third; x1 = svid; x2 = evid; }
                else if (e1l == e3l) { src = svid; x1 = evid; x2 = third; }
                else if (e2l == e3l) { src = evid; x1 = svid; x2 = third; }
            }
            vidToIndex[src] = 0;
            partialQueryVid.push_back(src);
            partialQueryVidMap |= (1u << src);
            uploadPartition(adj(src, x1), partitions, baseAddr_dev, stride, upload);
            addExtPhase(src, x1, src, x2, 2);
            addReducePhase({x1, x2});
        } else if (score <= 3) {
            vidToIndex[src] = 0;
            partialQueryVid.push_back(src);
            partialQueryVidMap |= (1u << src);
            uploadPartition(adj(src, x1), partitions, baseAddr_dev, stride, upload);
            addExtPhase(src, x1, src, x2, 1);
            addExtPhase(src, x2, src, x2, 1);
            addReducePhase({x1, x2});
        } else {
            throw geneva::PlanError{"unexpected triangle score"};
        }
    }

    // Greedy order over the remaining vertices (src/genMatchOrder.cpp:837-976).
    while (partialQueryVid.size() < vnum) {
        std::vector<NeigInfo> info;
        std::map<uint, std::vector<uint>> sameEdgeLabelEvid;
        for (uint s : partialQueryVid) {
            uint neigs = vertexNeig[s];
            auto& labelNeig = query[s].edgeLabelNeigs;
            for (uint i = 0; i < 32; ++i) {
                if ((neigs & (1u << i)) == 0) continue;
                NeigInfo n;
                n.svid = s;
                n.evid = i;
                n.edgeLabel = adj(s, i);
                n.commonVNum = (uint)__builtin_popcount(vertexNeig[i] & partialQueryVidMap) - 1u;
                n.evLabel = ql(i);
                n.evNum = (uint)__builtin_popcount(neigs & labelNeig[n.edgeLabel][1]);
                info.push_back(n);
                if (((1u << i) & partialQueryVidMap) == 0) {
                    auto& v = sameEdgeLabelEvid[n.edgeLabel];
                    v.push_back(s);
                    v.push_back(i);
                }
            }
        }
        if (info.empty()) throw geneva::PlanError{"planner found no frontier"};
        std::sort(info.begin(), info.end(), multiLevelSort);
        if (info[0].commonVNum > 0) {
            uint s = info[0].svid, x = info[0].evid;
            addExtPhase(s, x, 0, 0, 1);
            std::vector<NeigInfo> back;
            uint neigpos = vertexNeig[x] & partialQueryVidMap;
            for (uint i = 0; i < 32; ++i) {
                if ((neigpos & (1u << i)) == 0) continue;
                NeigInfo n;
                n.svid = x;
                n.evid = i;
                n.edgeLabel = adj(x, i);
                n.evLabel = ql(i);
                back.push_back(n);
            }
            std::sort(back.begin(), back.end(), byEdgeThenEndLabel);
            std::vector<uint> prev;
            for (size_t i = 0; i < back.size(); ++i) {
                if (!prev.empty() && back[i].edgeLabel != adj(prev[0], prev[1])) {
                    addReducePhase(prev);
                    prev.clear();
                }
                prev.push_back(back[i].svid);
                prev.push_back(back[i].evid);
            }
            addReducePhase(prev);
        } else {
            bool found = false;
            for (auto& kv : sameEdgeLabelEvid) {
                std::vector<uint>& ev = kv.second;
                for (uint i = 0; i < ev.size() / 2 && !found; ++i) {
                    uint a = ev[i * 2 + 1];
                    for (uint j = i + 1; j < ev.size() / 2; ++j) {
                        uint b = ev[j * 2 + 1];
                        if (adj(a, b) != 0xffffffffu) {
                            found = true;
                            addExtPhase(ev[i * 2], a, ev[j * 2], b, 2);
                            addReducePhase({a, b});
                            break;
                        }
                    }
                }
                if (found) break;
            }
            if (!found) {
                if (info[0].evNum > 1) {
                    addExtPhase(info[0].svid, info[0].evid, info[1].svid, info[1].evid, 2);
                } else {
                    bool pairFound = false;
                    for (size_t i = 0; i + 1 < info.size(); ++i) {
                        if (info[i].edgeLabel == info[i + 1].edgeLabel) {
                            addExtPhase(info[i].svid, info[i].evid, info[i + 1].svid,
                                        info[i + 1].evid, 2);
                            pairFound = true;
                            break;
                        }
                    }
                    if (!pairFound) addExtPhase(info[0].svid, info[0].evid, 0, 0, 1);
                }
            }
        }
    }
    std::vector<std::vector<uint>> auts;
    genAutomorphism(std::vector<uint>(), true, auts);
    rearrangeReducePhase();
    genRestriction(auts);
}

geneva::UploadFn g_upload = nullptr;

}  // namespace

namespace geneva {

void setPlannerUpload(UploadFn fn) { g_upload = fn; }

void planDV(std::istream& in, bool skipDigitAndT, std::vector<Phase>& matchOrder,
            std::map<uint64_t, uint>& repeatFlag, std::map<uint, uint>& restricts,
            std::map<uint, std::map<uint, bool>>& recordLabelStat,
            std::map<uint, uint>& queryLabel, std::map<uint, uint*>* partitions,
            uint* baseAddr_dev, uint* stride, UploadFn upload) {
    Planner<Phase, false> p(matchOrder, repeatFlag, restricts, recordLabelStat, queryLabel);
    p.run(in, skipDigitAndT, partitions, baseAddr_dev, stride, upload);
}

void planSV(std::istream& in, bool skipDigitAndT, std::vector<PhaseSingleV>& matchOrder,
            std::map<uint64_t, uint>& repeatFlag, std::map<uint, uint>& restricts,
            std::map<uint, std::map<uint, bool>>& recordLabelStat,
            std::map<uint, uint>& queryLabel, std::map<uint, uint*>* partitions,
            uint* baseAddr_dev, uint* stride, UploadFn upload) {
    Planner<PhaseSingleV, true> p(matchOrder, repeatFlag, restricts, recordLabelStat,
                                  queryLabel);
    p.run(in, skipDigitAndT, partitions, baseAddr_dev, stride, upload);
}

}  // namespace geneva

void genMatchOrder(std::string queryFileName, std::vector<Phase>& matchOrder,
                   std::map<uint64_t, uint>& repeatFlag, std::map<uint, uint>& restricts,
                   std::map<uint, std::map<uint, bool>>& recordLabelStat,
                   std::map<uint, uint>& queryLabel,
                   std::map<uint, uint*>& allEdgeLabelPartitions, uint* baseAddr_dev,
                   uint& stride) {
    std::ifstream f(queryFileName);
    if (!f.is_open()) throw geneva::PlanError{"can not open query file " + queryFileName};
    geneva::planDV(f, true, matchOrder, repeatFlag, restricts, recordLabelStat, queryLabel,
                   &allEdgeLabelPartitions, baseAddr_dev, &stride, g_upload);
}

void genMatchOrderSingleV(std::string queryFileName, std::vector<PhaseSingleV>& matchOrder,
                          std::map<uint64_t, uint>& repeatFlag,
                          std::map<uint, uint>& restricts,
                          std::map<uint, std::map<uint, bool>>& recordLabelStat,
                          std::map<uint, uint>& queryLabel,
                          std::map<uint, uint*>& allEdgeLabelPartitions, uint* baseAddr_dev,
                          uint& stride) {
    std::ifstream f(queryFileName);
    if (!f.is_open()) throw geneva::PlanError{"can not open query file " + queryFileName};
    geneva::planSV(f, false, matchOrder, repeatFlag, restricts, recordLabelStat, queryLabel,
                   &allEdgeLabelPartitions, baseAddr_dev, &stride, g_upload);
}

int calDVPhaseCount(const std::vector<Phase>& matchOrder) {
    int n = 0;
    for (auto& p : matchOrder)
        if (p.phaseType == 'E' && p.pairs.size() == 4) ++n;
    return n;
}
