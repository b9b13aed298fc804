// Database loader and format writers.
#include "loadDataBase.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

// src/loadDataBase.cpp:309-365.  Differences: a header partition count larger
// than the number of blocks in the file stops at end of file with a warning
// (the shipped Enron.mygraph says 4 and holds 3; the reference then reads a
// fourth block of garbage), and ids missing from the metadata get label
// 0xffffffff instead of uninitialised memory.
std::map<uint, uint*> loadDataBase(std::string filename, std::string metafilename) {
    std::map<uint, uint*> parts;
    std::ifstream graph(filename, std::ios::in | std::ios::binary);
    if (!graph.is_open()) {
        std::cerr << "open " << filename << " wrong" << std::endl;
        return parts;
    }
    std::ifstream meta(metafilename, std::ios::in);
    if (!meta.is_open()) {
        std::cerr << "open metafile wrong" << std::endl;
        return parts;
    }
    std::vector<std::pair<uint, uint>> vl;
    uint newvid, label, oldvid, maxid = 0;
    while (meta >> newvid >> label >> oldvid) {
        vl.emplace_back(newvid, label);
        if (newvid > maxid) maxid = newvid;
    }
    uint allocv = ((maxid + 1 + 15) >> 4) << 4;
    uint* vLabel = (uint*)malloc(sizeof(uint) * allocv);
    for (uint i = 0; i < allocv; ++i) vLabel[i] = 0xffffffffu;
    vLabel[0] = maxid;
    // std::map insert keeps the first label seen for a repeated id
    std::vector<bool> seen(maxid + 1, false);
    for (auto& p : vl) {
        if (p.first == 0 || seen[p.first]) continue;
        seen[p.first] = true;
        vLabel[p.first] = p.second;
    }
    uint totpart = 0;
    graph.read((char*)&totpart, sizeof(uint));
    for (uint j = 1; j <= totpart; ++j) {
        uint n = 0;
        if (!graph.read((char*)&n, sizeof(uint)) || n < 5) {
            std::cerr << "loadDataBase: header lists " << totpart << " partitions, file holds "
                      << (j - 1) << std::endl;
            break;
        }
        uint alloc = ((n + 15) >> 4) << 4;
        uint* blk = new uint[alloc];
        blk[0] = n;
        if (!graph.read((char*)&blk[1], sizeof(uint) * (n - 1))) {
            std::cerr << "loadDataBase: partition " << j << " truncated" << std::endl;
            delete[] blk;
            break;
        }
        parts.emplace(j, blk);
    }
    parts.emplace(0xffffffffu, vLabel);
    return parts;
}

namespace geneva {

void freeDataBase(std::map<uint, uint*>& db) {
    for (auto& kv : db) {
        if (kv.first == 0xffffffffu)
            free(kv.second);
        else
            delete[] kv.second;
    }
    db.clear();
}

namespace {

// Per-source adjacency of one edge label: for every vertex, neighbours sorted
// by (label, id), deduplicated.
struct LabelAdj {
    std::vector<uint64_t> off;  // nv+2
    std::vector<uint> nbr;
};

LabelAdj buildAdj(const DataGraph& g, uint edgeLabel) {
    LabelAdj a;
    a.off.assign(g.nv + 2, 0);
    for (auto& e : g.edges) {
        if (e.label != edgeLabel || e.a == e.b) continue;
        a.off[e.a + 1]++;
        a.off[e.b + 1]++;
    }
    for (uint v = 1; v <= g.nv + 1; ++v) a.off[v] += a.off[v - 1];
    a.nbr.resize(a.off[g.nv + 1]);
    std::vector<uint64_t> pos(a.off.begin(), a.off.end() - 1);
    for (auto& e : g.edges) {
        if (e.label != edgeLabel || e.a == e.b) continue;
        a.nbr[pos[e.a]++] = e.b;
        a.nbr[pos[e.b]++] = e.a;
    }
    // sort each list by (label, id) and drop duplicates
    std::vector<uint64_t> newoff(g.nv + 2, 0);
    uint64_t w = 0;
    for (uint v = 0; v <= g.nv; ++v) {
        newoff[v] = w;
        uint* b = a.nbr.data() + a.off[v];
        uint* e = a.nbr.data() + a.off[v + 1];
        std::sort(b, e, [&](uint x, uint y) {
            uint lx = g.vlabel[x], ly = g.vlabel[y];
            return lx != ly ? lx < ly : x < y;
        });
        uint* u = std::unique(b, e);
        for (uint* p = b; p < u; ++p) a.nbr[w++] = *p;
    }
    newoff[g.nv + 1] = w;
    a.nbr.resize(w);
    a.off.swap(newoff);
    return a;
}

// tools/genDatabase.cpp:415-436
void appendWithSkipIndex(std::vector<uint>& out, const uint* grp, uint s) {
    if (s > 128) {
        if (s <= 32 * 128) {
            uint blocks = s / 128;
            for (uint i = 1; i <= blocks; ++i) out.push_back(grp[i * 128 - 1]);
            if (s % 128 > 0) out.push_back(grp[s - 1]);
        } else {
            uint bl = s / 32;
            for (uint i = 1; i <= 31; ++i) out.push_back(grp[i * bl - 1]);
            out.push_back(grp[s - 1]);
        }
    }
    out.insert(out.end(), grp, grp + s);
}

}  // namespace

uint* buildPartitionBlock(const DataGraph& g, uint edgeLabel, std::string* err) {
    LabelAdj a = buildAdj(g, edgeLabel);
    std::vector<uint> srcs;
    for (uint v = 1; v <= g.nv; ++v)
        if (a.off[v + 1] > a.off[v]) srcs.push_back(v);
    if (srcs.empty()) {
        if (err) *err = "empty partition";
        return nullptr;
    }
    std::vector<std::pair<uint, uint>> runs;  // (start, len)
    for (uint v : srcs) {
        if (!runs.empty() && runs.back().first + runs.back().second == v)
            runs.back().second++;
        else
            runs.emplace_back(v, 1u);
    }
    if (runs.size() > 256) {
        if (err) *err = "more than 256 source intervals";
        return nullptr;
    }
    const uint n = (uint)runs.size(), S = (uint)srcs.size();
    std::vector<uint> nb;  // array3
    std::vector<uint> a1;  // array1
    a1.reserve(S + 1);
    for (uint v : srcs) {
        a1.push_back((uint)nb.size());
        const uint* p = a.nbr.data() + a.off[v];
        const uint* e = a.nbr.data() + a.off[v + 1];
        while (p < e) {
            const uint* q = p;
            while (q < e && g.vlabel[*q] == g.vlabel[*p]) ++q;
            appendWithSkipIndex(nb, p, (uint)(q - p));
            p = q;
        }
    }
    a1.push_back((uint)nb.size());
    uint64_t tot = 5ull + 2 * n + 2 + (S + 1) + nb.size();
    if (tot >= 0xffffffffull) {
        if (err) *err = "partition too large for 32-bit offsets";
        return nullptr;
    }
    uint* blk = new uint[(size_t)tot];   // freed by freeDataBase, as loadDataBase's blocks
    blk[0] = (uint)tot;
    blk[1] = S;
    blk[2] = n;
    blk[3] = 2 * n + 2 + (S + 1);
    blk[4] = blk[3];
    uint pre = 0;
    for (uint i = 0; i < n; ++i) {
        blk[5 + i] = runs[i].first;
        pre += runs[i].second;
        blk[5 + n + 1 + i] = pre;
    }
    blk[5 + n] = 0;
    blk[5 + 2 * n + 1] = 0;
    memcpy(blk + 5 + 2 * n + 2, a1.data(), sizeof(uint) * a1.size());
    memcpy(blk + 5 + blk[4], nb.data(), sizeof(uint) * nb.size());
    return blk;
}

std::map<uint, uint*> buildDataBase(const DataGraph& g, std::map<uint, uint>* keyOfLabel,
                                    std::string* err) {
    std::map<uint, uint*> db;
    std::vector<uint> labels;
    for (auto& e : g.edges) labels.push_back(e.label);
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
    uint key = 1;
    for (uint l : labels) {
        uint* b = buildPartitionBlock(g, l, err);
        if (!b) {
            freeDataBase(db);
            return db;
        }
        db.emplace(key, b);
        if (keyOfLabel) (*keyOfLabel)[l] = key;
        ++key;
    }
    uint allocv = ((g.nv + 1 + 15) >> 4) << 4;
    uint* vLabel = (uint*)malloc(sizeof(uint) * allocv);
    for (uint i = 0; i < allocv; ++i) vLabel[i] = 0xffffffffu;
    vLabel[0] = g.nv;
    for (uint v = 1; v <= g.nv; ++v) vLabel[v] = g.vlabel[v];
    db.emplace(0xffffffffu, vLabel);
    return db;
}

bool writeGraphText(const std::string& path, const DataGraph& g) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return false;
    fprintf(f, "t # 0\n");
    for (uint v = 1; v <= g.nv; ++v) fprintf(f, "v %u %u\n", v, g.vlabel[v]);
    for (auto& e : g.edges) fprintf(f, "e %u %u %u\n", e.a, e.b, e.label);
    fclose(f);
    return true;
}

bool writeMetadata(const std::string& path, const DataGraph& g) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return false;
    for (uint v = 1; v <= g.nv; ++v)
        fprintf(f, "%u %u %u\n", v, g.vlabel[v], g.oldId.size() > v ? g.oldId[v] : v);
    fclose(f);
    return true;
}

namespace {
// each vertex's edge-label signature: bit b set when it has an edge of the
// b-th label of `labels`
std::vector<uint64_t> signatures(const DataGraph& g, const std::vector<uint>& labels) {
    std::vector<uint64_t> sig(g.nv + 1, 0);
    for (auto& e : g.edges) {
        if (e.a == e.b) continue;
        const uint64_t bit = 1ull << (std::lower_bound(labels.begin(), labels.end(), e.label) - labels.begin());
        sig[e.a] |= bit;
        sig[e.b] |= bit;
    }
    return sig;
}
std::vector<uint> edgeLabels(const DataGraph& g) {
    std::vector<uint> labels;
    for (auto& e : g.edges)
        if (e.a != e.b) labels.push_back(e.label);
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
    return labels;
}
}  // namespace

std::map<uint, uint> partitionRuns(const DataGraph& g) {
    const std::vector<uint> labels = edgeLabels(g);
    std::map<uint, uint> runs;
    if (labels.size() > 64) return runs;
    const std::vector<uint64_t> sig = signatures(g, labels);
    for (size_t b = 0; b < labels.size(); ++b) {
        uint r = 0;
        bool in = false;
        for (uint v = 1; v <= g.nv; ++v) {
            const bool s = (sig[v] >> b) & 1;
            if (s && !in) ++r;
            in = s;
        }
        runs[labels[b]] = r;
    }
    return runs;
}

uint reorderForIntervals(const DataGraph& g, std::vector<uint>& newOf, std::string* err) {
    const std::vector<uint> labels = edgeLabels(g);
    const size_t K = labels.size();
    if (K > 64) {
        if (err) *err = "more than 64 edge labels";
        return 0;
    }
    const std::vector<uint64_t> sig = signatures(g, labels);
    // bit positions: the label with the most sources takes the highest bit
    std::vector<uint64_t> srcs(K, 0);
    for (uint v = 1; v <= g.nv; ++v)
        for (size_t b = 0; b < K; ++b) srcs[b] += (sig[v] >> b) & 1;
    std::vector<size_t> byCount(K);
    for (size_t b = 0; b < K; ++b) byCount[b] = b;
    std::stable_sort(byCount.begin(), byCount.end(), [&](size_t x, size_t y) { return srcs[x] > srcs[y]; });
    std::vector<uint> bitOf(K);
    for (size_t i = 0; i < K; ++i) bitOf[byCount[i]] = (uint)(K - 1 - i);
    // rank of a signature read as a Gray code: prefix XOR from the top bit
    auto rank = [&](uint64_t s) {
        uint64_t code = 0;
        for (size_t b = 0; b < K; ++b)
            if ((s >> b) & 1) code |= 1ull << bitOf[b];
        uint64_t r = code;
        for (uint sh = 1; sh < 64; sh <<= 1) r ^= r >> sh;
        return r;
    };
    std::vector<std::pair<uint64_t, uint>> key;
    key.reserve(g.nv);
    for (uint v = 1; v <= g.nv; ++v) key.emplace_back(rank(sig[v]), v);
    std::sort(key.begin(), key.end());
    newOf.assign(g.nv + 1, 0);
    for (uint i = 0; i < key.size(); ++i) newOf[key[i].second] = i + 1;
    uint worst = 0;
    for (auto& kv : partitionRuns(relabel(g, newOf))) worst = std::max(worst, kv.second);
    if (worst > 256) {
        if (err) *err = "a partition needs " + std::to_string(worst) + " source intervals after reordering";
        return 0;
    }
    return worst;
}

DataGraph relabel(const DataGraph& g, const std::vector<uint>& newOf) {
    DataGraph h;
    h.nv = g.nv;
    h.vlabel.assign(g.nv + 1, 0);
    h.oldId.assign(g.nv + 1, 0);
    for (uint v = 1; v <= g.nv; ++v) {
        h.vlabel[newOf[v]] = g.vlabel[v];
        h.oldId[newOf[v]] = g.oldId.size() > v ? g.oldId[v] : v;
    }
    h.edges.reserve(g.edges.size());
    for (auto& e : g.edges) h.edges.push_back({newOf[e.a], newOf[e.b], e.label});
    return h;
}

bool writeMyGraph(const std::string& path, const DataGraph& g, std::string* err) {
    std::map<uint, uint*> db = buildDataBase(g, nullptr, err);
    if (db.empty()) return false;
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        freeDataBase(db);
        return false;
    }
    uint totpart = (uint)db.size() - 1;
    fwrite(&totpart, sizeof(uint), 1, f);
    for (auto& kv : db) {
        if (kv.first == 0xffffffffu) continue;
        fwrite(kv.second, sizeof(uint), kv.second[0], f);
    }
    fclose(f);
    freeDataBase(db);
    return true;
}

bool writeReorder(const std::string& path, const DataGraph& g, bool withSkipIndex) {
    uint maxLabel = 0;
    for (auto& e : g.edges) maxLabel = std::max(maxLabel, e.label);
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return false;
    fprintf(f, "%u\n", maxLabel + 1);
    for (uint l = 0; l <= maxLabel; ++l) {
        LabelAdj a = buildAdj(g, l);
        uint nsrc = 0;
        for (uint v = 1; v <= g.nv; ++v)
            if (a.off[v + 1] > a.off[v]) ++nsrc;
        fprintf(f, "%u\n", nsrc);
        for (uint v = 1; v <= g.nv; ++v) {
            const uint* p = a.nbr.data() + a.off[v];
            const uint* e = a.nbr.data() + a.off[v + 1];
            if (p == e) continue;
            uint ngroups = 0;
            for (const uint* q = p; q < e;) {
                const uint* r = q;
                while (r < e && g.vlabel[*r] == g.vlabel[*q]) ++r;
                ++ngroups;
                q = r;
            }
            fprintf(f, "%u %u\n", v, ngroups);
            while (p < e) {
                const uint* q = p;
                while (q < e && g.vlabel[*q] == g.vlabel[*p]) ++q;
                std::vector<uint> grp;
                if (withSkipIndex)
                    appendWithSkipIndex(grp, p, (uint)(q - p));
                else
                    grp.assign(p, q);
                fprintf(f, "%u %zu\n", g.vlabel[*p], grp.size());
                for (uint x : grp) fprintf(f, "%u ", x);
                fprintf(f, "\n");
                p = q;
            }
        }
    }
    fclose(f);
    return true;
}

bool readGraphText(const std::string& path, DataGraph& g, std::string* err) {
    std::ifstream in(path);
    if (!in.is_open()) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    g = DataGraph();
    std::string line;
    std::vector<std::pair<uint, uint>> vs;
    uint maxid = 0;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line[0] == 'v') {
            std::istringstream ss(line.substr(1));
            uint id, lab;
            if (ss >> id >> lab) {
                vs.emplace_back(id, lab);
                maxid = std::max(maxid, id);
            }
        } else if (line[0] == 'e') {
            std::istringstream ss(line.substr(1));
            uint a, b, l;
            if (ss >> a >> b >> l) {
                g.edges.push_back({a, b, l});
                maxid = std::max(maxid, std::max(a, b));
            }
        }
    }
    g.nv = maxid;
    g.vlabel.assign(maxid + 1, 0xffffffffu);
    for (auto& p : vs) g.vlabel[p.first] = p.second;
    return true;
}

bool forEachLabelGroup(const uint* block, const uint* vLabel, uint maxid,
                       void (*fn)(void*, uint, uint, const uint*, uint), void* user,
                       std::string* err) {
    const uint tot = block[0], S = block[1], n = block[2], u4 = block[4];
    if (n == 0 || n > 256 || 5ull + 2 * n + 2 + S + 1 > tot || 5ull + u4 > tot) {
        if (err) *err = "malformed partition header";
        return false;
    }
    const uint* a0 = block + 5;
    const uint* a1 = block + 5 + 2 * n + 2;
    const uint* a3 = block + 5 + u4;
    const uint64_t nnb = tot - 5 - u4;
    for (uint j = 0; j < n; ++j) {
        uint start = a0[j], lo = a0[n + j], hi = a0[n + 1 + j];
        if (hi < lo || hi > S) {
            if (err) *err = "malformed interval table";
            return false;
        }
        for (uint k = lo; k < hi; ++k) {
            uint src = start + (k - lo);
            uint b = a1[k], e = a1[k + 1];
            if (e < b || e > nnb || src == 0 || src > maxid) {
                if (err) *err = "malformed neighbour offsets";
                return false;
            }
            uint p = b;
            while (p < e) {
                if (a3[p] == 0 || a3[p] > maxid) {
                    if (err) *err = "neighbour id out of range";
                    return false;
                }
                uint lab = vLabel[a3[p]];
                uint q = p;
                while (q < e && a3[q] <= maxid && vLabel[a3[q]] == lab) ++q;
                uint L = q - p, pre = skipIndexLen(L);
                fn(user, src, lab, a3 + p + pre, L - pre);
                p = q;
            }
        }
    }
    return true;
}

}  // namespace geneva
