// ORACLE (test infrastructure only).  Nothing in the product library links,
// loads or calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker.
//
// CPU restatement of the reference's cpumine (src/cpumine.cpp) on prefix-free
// adjacency with uint64 totals.  Count semantics (src/cpumine.cpp:54-584): a
// map f from query vertices to data vertices is counted when it is injective,
// keeps vertex labels, maps every query edge (u,w,l) onto a data edge of label
// l, and satisfies f(u) < f(restricts[u]) for every restricted u.
//
// Modes
//   faithful : phase by phase over materialised row tables, as cpumine does
//              (genInit :54-196, genExtEmb :198-439, reduction :486-584);
//              small inputs only.
//   count    : depth-first over the plan's vertex order, OpenMP over start
//              vertices, count-only last step; also accumulates the
//              algorithmic byte figure B (see below).
//   brute    : every label-preserving injective map, checked against all
//              query edges and restrictions; tiny inputs only.  Shares nothing
//              with the other two modes except the restriction map.
//
// Algorithmic bytes.  For every extension of a partial row m of length L to a
// new query vertex u with earlier query neighbours w_1..w_j:
//   B(m) = 4L + sum_t (8 + 4 |N_{l(u)}(f(w_t))|) + [u not last] 4 (L+1) survivors(m)
// where N_{l}(x) is x's label-l neighbour list under the edge's label.  Rows of
// length 1 are all data vertices carrying the label of the first vertex.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "oracle_planner.h"

namespace oracle {

struct Adj {  // one edge label: neighbours of v sorted by (label, id), no duplicates
    std::vector<uint64_t> off;      // nv+2
    std::vector<unsigned> nbr;
    std::vector<uint64_t> goff;     // per vertex: label groups [goff[v], goff[v+1])
    std::vector<unsigned> glabel;   // label of group
    std::vector<uint64_t> gbeg;     // first neighbour index of group
};

struct Graph {
    unsigned nv = 0;
    std::vector<unsigned> vlabel;            // nv+1
    std::map<unsigned, Adj> byEdgeLabel;
    std::map<unsigned, std::vector<unsigned>> byVertexLabel;  // label -> ids ascending
};

namespace {

void buildAdj(Graph& g, unsigned el, const std::vector<std::pair<unsigned, unsigned>>& es) {
    Adj& a = g.byEdgeLabel[el];
    a.off.assign(g.nv + 2, 0);
    for (auto& e : es) { a.off[e.first + 1]++; a.off[e.second + 1]++; }
    for (unsigned v = 1; v <= g.nv + 1; ++v) a.off[v] += a.off[v - 1];
    std::vector<unsigned> tmp(a.off[g.nv + 1]);
    std::vector<uint64_t> pos(a.off.begin(), a.off.end() - 1);
    for (auto& e : es) { tmp[pos[e.first]++] = e.second; tmp[pos[e.second]++] = e.first; }
    a.nbr.clear();
    std::vector<uint64_t> off2(g.nv + 2, 0);
    a.goff.assign(g.nv + 2, 0);
    for (unsigned v = 0; v <= g.nv; ++v) {
        off2[v] = a.nbr.size();
        a.goff[v] = a.glabel.size();
        auto b = tmp.begin() + a.off[v], e = tmp.begin() + a.off[v + 1];
        std::sort(b, e, [&](unsigned x, unsigned y) {
            return g.vlabel[x] != g.vlabel[y] ? g.vlabel[x] < g.vlabel[y] : x < y;
        });
        e = std::unique(b, e);
        for (auto p = b; p != e; ++p) {
            if (p == b || g.vlabel[*p] != g.vlabel[*(p - 1)]) {
                a.glabel.push_back(g.vlabel[*p]);
                a.gbeg.push_back(a.nbr.size());
            }
            a.nbr.push_back(*p);
        }
    }
    off2[g.nv + 1] = a.nbr.size();
    a.goff[g.nv + 1] = a.glabel.size();
    a.off.swap(off2);
}

// Label-l neighbours of v: [b, e) into a.nbr.
inline void seg(const Adj& a, unsigned v, unsigned l, uint64_t& b, uint64_t& e) {
    for (uint64_t k = a.goff[v]; k < a.goff[v + 1]; ++k) {
        if (a.glabel[k] == l) {
            b = a.gbeg[k];
            e = (k + 1 < a.goff[v + 1]) ? a.gbeg[k + 1] : a.off[v + 1];
            return;
        }
    }
    b = e = 0;
}

inline bool contains(const Adj& a, uint64_t b, uint64_t e, unsigned x) {
    return std::binary_search(a.nbr.begin() + b, a.nbr.begin() + e, x);
}

struct Step {              // extension to vertex order[t]
    unsigned u, label;
    unsigned anchorCol = 0;          // column of the plan's source vertex for u
    const Adj* anchorAdj = nullptr;
    std::vector<unsigned> backCol;        // all earlier neighbours (anchor included)
    std::vector<const Adj*> backAdj;
    int restrictCol = -1;                 // f(u) < row[restrictCol]
};

struct Compiled {
    std::vector<unsigned> order;
    std::vector<unsigned> colOf;          // query vertex -> column
    std::vector<Step> steps;              // steps[t] for t = 1..k-1
    unsigned label0;
    bool ok = true;                       // a needed edge label is absent -> count 0
};

Compiled compile(const Graph& g, const OPlan& P) {
    Compiled c;
    const unsigned n = P.vnum;
    c.colOf.assign(n, 0xffffffffu);
    std::vector<unsigned> src(n, 0xffffffffu);
    for (auto& ph : P.phases) {
        if (ph.type != 'E') continue;
        for (size_t i = 0; i + 1 < ph.pairs.size(); i += 2) {
            unsigned s = ph.pairs[i], x = ph.pairs[i + 1];
            if (c.order.empty()) { c.colOf[s] = 0; c.order.push_back(s); }
            if (c.colOf[x] == 0xffffffffu) {
                c.colOf[x] = (unsigned)c.order.size();
                c.order.push_back(x);
                src[x] = s;
            }
        }
    }
    c.label0 = P.queryLabel.at(c.order[0]);
    c.steps.resize(n);
    for (unsigned t = 1; t < n; ++t) {
        Step& s = c.steps[t];
        s.u = c.order[t];
        s.label = P.queryLabel.at(s.u);
        for (unsigned w = 0; w < n; ++w) {
            unsigned el = P.adj[s.u * n + w];
            if (el == 0xffffffffu || c.colOf[w] >= t) continue;
            auto it = g.byEdgeLabel.find(el);
            const Adj* a = it == g.byEdgeLabel.end() ? nullptr : &it->second;
            if (!a) c.ok = false;
            s.backCol.push_back(c.colOf[w]);
            s.backAdj.push_back(a);
            if (w == src[s.u]) { s.anchorCol = c.colOf[w]; s.anchorAdj = a; }
        }
        unsigned r = P.restricts.count(s.u) ? P.restricts.at(s.u) : 0xffffffffu;
        if (r != 0xffffffffu) s.restrictCol = (int)c.colOf[r];
    }
    return c;
}

struct Acc {
    uint64_t count = 0, bytes = 0;
};

// Depth-first count over the vertex order (cpumine semantics).
void dfs(const Graph& g, const Compiled& c, unsigned t, unsigned* row, Acc& acc,
         bool withBytes) {
    const unsigned k = (unsigned)c.order.size();
    const Step& s = c.steps[t];
    const bool last = (t + 1 == k);
    uint64_t ab = 0, ae = 0;
    seg(*s.anchorAdj, row[s.anchorCol], s.label, ab, ae);
    if (withBytes) {
        acc.bytes += 4ull * t;
        for (size_t i = 0; i < s.backCol.size(); ++i) {
            uint64_t b, e;
            seg(*s.backAdj[i], row[s.backCol[i]], s.label, b, e);
            acc.bytes += 8 + 4 * (e - b);
        }
    }
    const unsigned bound = s.restrictCol >= 0 ? row[s.restrictCol] : 0xffffffffu;
    const std::vector<unsigned>& nb = s.anchorAdj->nbr;
    uint64_t cut = std::lower_bound(nb.begin() + ab, nb.begin() + ae, bound) - nb.begin();
    if (last && s.backCol.size() == 1) {
        // one neighbour list: everything below the bound except mapped vertices
        uint64_t n = cut - ab;
        for (unsigned j = 0; j < t; ++j)
            if (row[j] < bound && contains(*s.anchorAdj, ab, cut, row[j])) --n;
        acc.count += n;
        return;
    }
    uint64_t surv = 0;
    for (uint64_t p = ab; p < cut; ++p) {
        unsigned x = nb[p];
        bool okc = true;
        for (unsigned j = 0; j < t && okc; ++j) okc = row[j] != x;
        for (size_t i = 0; i < s.backCol.size() && okc; ++i) {
            if (s.backCol[i] == s.anchorCol && s.backAdj[i] == s.anchorAdj) continue;
            uint64_t b, e;
            seg(*s.backAdj[i], row[s.backCol[i]], s.label, b, e);
            okc = contains(*s.backAdj[i], b, e, x);
        }
        if (!okc) continue;
        ++surv;
        if (!last) {
            row[t] = x;
            dfs(g, c, t + 1, row, acc, withBytes);
        }
    }
    if (last) acc.count += surv;
    else if (withBytes) acc.bytes += 4ull * (t + 1) * surv;
}

// cpumine's phase-by-phase execution on in-memory tables.  finalRows
// (nullable) receives the last table with its columns put in query-vertex
// order (column u = the data vertex of query vertex u).
uint64_t faithful(const Graph& g, const OPlan& P, const std::vector<unsigned>* graphOf,
                  std::vector<uint8_t>* bits, std::vector<unsigned>* finalRows = nullptr) {
    std::map<unsigned, unsigned> pos;  // posOfVisitedV
    auto ql = [&](unsigned v) { return P.queryLabel.at(v); };
    auto restr = [&](unsigned v) {
        auto it = P.restricts.find(v);
        return it == P.restricts.end() ? 0xffffffffu : it->second;
    };
    auto adjOf = [&](unsigned el) -> const Adj* {
        auto it = g.byEdgeLabel.find(el);
        return it == g.byEdgeLabel.end() ? nullptr : &it->second;
    };
    std::vector<unsigned> rows;
    unsigned len = 0;
    const OPhase& f = P.phases[0];
    const Adj* a0 = adjOf(f.edgeLabel);
    if (!a0) return 0;
    // genInit (src/cpumine.cpp:54-196): sources of the partition with the right label
    if (f.pairs.size() == 2) {
        unsigned s = f.pairs[0], x = f.pairs[1];
        pos[s] = 0; pos[x] = 1; len = 2;
        bool rx = restr(x) == s;
        for (unsigned v = 1; v <= g.nv; ++v) {
            if (a0->off[v + 1] == a0->off[v] || g.vlabel[v] != ql(s)) continue;
            uint64_t b, e;
            seg(*a0, v, ql(x), b, e);
            for (uint64_t p = b; p < e; ++p)
                if (!rx || v > a0->nbr[p]) { rows.push_back(v); rows.push_back(a0->nbr[p]); }
        }
    } else {
        unsigned s = f.pairs[0], x1 = f.pairs[1], x2 = f.pairs[3];
        pos[s] = 0; pos[x1] = 1; pos[x2] = 2; len = 3;
        bool r1 = restr(x1) != 0xffffffffu;
        bool r12 = restr(x2) != 0xffffffffu && restr(x2) == x1;
        bool r2 = !r12 && restr(x2) != 0xffffffffu;
        for (unsigned v = 1; v <= g.nv; ++v) {
            if (a0->off[v + 1] == a0->off[v] || g.vlabel[v] != ql(s)) continue;
            uint64_t b1, e1, b2, e2;
            seg(*a0, v, ql(x1), b1, e1);
            seg(*a0, v, ql(x2), b2, e2);
            if (b1 == e1 || b2 == e2) continue;
            for (uint64_t p = b1; p < e1; ++p)
                for (uint64_t q = b2; q < e2; ++q) {
                    unsigned n1 = a0->nbr[p], n2 = a0->nbr[q];
                    if (n1 == n2) continue;
                    if (r1 && v < n1) continue;
                    if (r2 && v < n2) continue;
                    if (r12 && n1 < n2) continue;
                    rows.push_back(v); rows.push_back(n1); rows.push_back(n2);
                }
        }
    }
    for (size_t pi = 1; pi < P.phases.size(); ++pi) {
        const OPhase& ph = P.phases[pi];
        const Adj* a = adjOf(ph.edgeLabel);
        if (!a) return 0;
        std::vector<unsigned> out;
        const size_t nrows = rows.size() / len;
        if (ph.type == 'E' && ph.pairs.size() == 2) {
            // genExtEmb 1V (src/cpumine.cpp:207-297)
            unsigned s = ph.pairs[0], x = ph.pairs[1];
            unsigned idx = (unsigned)pos.size();
            pos[x] = idx;
            bool rx = restr(x) != 0xffffffffu;
            unsigned rp = rx ? pos.at(restr(x)) : 0;
            for (size_t r = 0; r < nrows; ++r) {
                const unsigned* m = &rows[r * len];
                uint64_t b, e;
                seg(*a, m[pos.at(s)], ql(x), b, e);
                for (uint64_t p = b; p < e; ++p) {
                    unsigned c = a->nbr[p];
                    bool found = false;
                    for (unsigned k = 0; k < len; ++k) if (m[k] == c) { found = true; break; }
                    if (found || (rx && !(m[rp] > c))) continue;
                    out.insert(out.end(), m, m + len);
                    out.push_back(c);
                }
            }
            len += 1;
        } else if (ph.type == 'E') {
            // genExtEmb 2V (src/cpumine.cpp:298-435)
            unsigned s1 = ph.pairs[0], x1 = ph.pairs[1], s2 = ph.pairs[2], x2 = ph.pairs[3];
            unsigned idx = (unsigned)pos.size();
            pos[x1] = idx; pos[x2] = idx + 1;
            bool r1 = restr(x1) != 0xffffffffu;
            bool r12 = restr(x2) != 0xffffffffu && restr(x2) == x1;
            bool r2 = !r12 && restr(x2) != 0xffffffffu;
            for (size_t r = 0; r < nrows; ++r) {
                const unsigned* m = &rows[r * len];
                uint64_t b1, e1, b2, e2;
                seg(*a, m[pos.at(s1)], ql(x1), b1, e1);
                seg(*a, m[pos.at(s2)], ql(x2), b2, e2);
                for (uint64_t p = b1; p < e1; ++p) {
                    unsigned c1 = a->nbr[p];
                    bool found = false;
                    for (unsigned k = 0; k < len; ++k) if (m[k] == c1) { found = true; break; }
                    if (found || (r1 && c1 > m[pos.at(restr(x1))])) continue;
                    for (uint64_t q = b2; q < e2; ++q) {
                        unsigned c2 = a->nbr[q];
                        bool f2 = false;
                        for (unsigned k = 0; k < len; ++k) if (m[k] == c2) { f2 = true; break; }
                        if (f2 || c2 == c1) continue;
                        if (r2 && c2 > m[pos.at(restr(x2))]) continue;
                        if (r12 && c2 > c1) continue;
                        out.insert(out.end(), m, m + len);
                        out.push_back(c1);
                        out.push_back(c2);
                    }
                }
            }
            len += 2;
        } else {
            // reduction (src/cpumine.cpp:486-584)
            for (size_t r = 0; r < nrows; ++r) {
                const unsigned* m = &rows[r * len];
                bool all = true;
                for (size_t j = 0; j + 1 < ph.pairs.size() && all; j += 2) {
                    unsigned sv = m[pos.at(ph.pairs[j])], ev = m[pos.at(ph.pairs[j + 1])];
                    uint64_t b, e;
                    seg(*a, sv, ql(ph.pairs[j + 1]), b, e);
                    all = std::find(a->nbr.begin() + b, a->nbr.begin() + e, ev) !=
                          a->nbr.begin() + e;
                }
                if (all) out.insert(out.end(), m, m + len);
            }
        }
        rows.swap(out);
    }
    if (graphOf && bits)
        for (size_t r = 0; r < rows.size() / len; ++r) (*bits)[(*graphOf)[rows[r * len]]] = 1;
    if (finalRows) {
        finalRows->assign(rows.size(), 0);
        for (size_t r = 0; r < rows.size() / len; ++r)
            for (auto& kv : pos) (*finalRows)[r * len + kv.first] = rows[r * len + kv.second];
    }
    return rows.size() / len;
}

uint64_t brute(const Graph& g, const OPlan& P, const std::vector<unsigned>* graphOf,
               std::vector<uint8_t>* bits) {
    const unsigned n = P.vnum;
    std::vector<unsigned> f(n, 0);
    std::vector<std::vector<unsigned>> cand(n);
    for (unsigned u = 0; u < n; ++u) {
        auto it = g.byVertexLabel.find(P.queryLabel.at(u));
        if (it == g.byVertexLabel.end()) return 0;
        cand[u] = it->second;
    }
    auto hasEdge = [&](unsigned a, unsigned b, unsigned el) {
        auto it = g.byEdgeLabel.find(el);
        if (it == g.byEdgeLabel.end()) return false;
        const Adj& A = it->second;
        return std::find(A.nbr.begin() + A.off[a], A.nbr.begin() + A.off[a + 1], b) !=
               A.nbr.begin() + A.off[a + 1];
    };
    uint64_t count = 0;
    std::function<void(unsigned)> rec = [&](unsigned u) {
        if (u == n) {
            for (auto& kv : P.restricts)
                if (kv.second != 0xffffffffu && kv.first < n && !(f[kv.first] < f[kv.second]))
                    return;
            ++count;
            if (graphOf && bits) (*bits)[(*graphOf)[f[0]]] = 1;
            return;
        }
        for (unsigned x : cand[u]) {
            bool ok = true;
            for (unsigned w = 0; w < u && ok; ++w) {
                if (f[w] == x) ok = false;
                unsigned el = P.adj[u * n + w];
                if (ok && el != 0xffffffffu) ok = hasEdge(f[w], x, el);
            }
            if (!ok) continue;
            f[u] = x;
            rec(u + 1);
        }
    };
    rec(0);
    return count;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

// Graph from an undirected edge list (a, b, edge label) and vertex labels
// vlabel[0..nv] (vlabel[0] ignored).  Self loops and duplicates are dropped.
void* oracle_graph_create(uint32_t nv, const uint32_t* vlabel, uint64_t m,
                          const uint32_t* edges) {
    Graph* g = new Graph();
    g->nv = nv;
    g->vlabel.assign(vlabel, vlabel + nv + 1);
    std::map<unsigned, std::vector<std::pair<unsigned, unsigned>>> byl;
    for (uint64_t i = 0; i < m; ++i) {
        unsigned a = edges[3 * i], b = edges[3 * i + 1], l = edges[3 * i + 2];
        if (a == b || a == 0 || b == 0 || a > nv || b > nv) continue;
        byl[l].emplace_back(a, b);
    }
    for (auto& kv : byl) buildAdj(*g, kv.first, kv.second);
    for (unsigned v = 1; v <= nv; ++v) g->byVertexLabel[g->vlabel[v]].push_back(v);
    return g;
}

void oracle_graph_destroy(void* h) { delete (Graph*)h; }

// Plan dump of a query in the reference cpumine print layout; returns the
// length written (0 on error, message in out).
uint64_t oracle_plan(const char* query, char* out, uint64_t cap) {
    OPlan P;
    std::string err;
    std::string s = planQuery(query, P, err) ? dumpPlan(P) : ("error: " + err);
    uint64_t n = std::min<uint64_t>(s.size(), cap ? cap - 1 : 0);
    if (cap) { memcpy(out, s.data(), n); out[n] = 0; }
    return n;
}

// Mine one query.  mode 0 = count (threads OpenMP workers; start vertices
// restricted to indices startBegin, startBegin+stride, ... below startEnd of
// the id-sorted start list, startEnd = 0 meaning all), 1 = faithful,
// 2 = brute force.
// graphOf (nullable): graph id of every vertex (nv+1 entries); bits receives
// one byte per graph (1 = matched).  Returns 0 on success, negative on a
// malformed query.
int oracle_mine(void* h, const char* query, int mode, int threads, uint64_t startBegin,
                uint64_t startEnd, uint64_t stride, const uint32_t* graphOf, uint32_t ngraphs,
                uint8_t* bits,
                uint64_t* countOut, uint64_t* bytesOut, uint64_t* nstartOut, double* secs) {
    Graph& g = *(Graph*)h;
    OPlan P;
    std::string err;
    if (!planQuery(query, P, err)) return -1;
    std::vector<unsigned> gof;
    std::vector<uint8_t> gb;
    if (graphOf && bits) {
        gof.assign(graphOf, graphOf + g.nv + 1);
        gb.assign(ngraphs, 0);
    }
    auto t0 = std::chrono::steady_clock::now();
    uint64_t count = 0, bytes = 0, nstart = 0;
    if (mode == 1) {
        count = faithful(g, P, gof.empty() ? nullptr : &gof, gof.empty() ? nullptr : &gb);
    } else if (mode == 2) {
        count = brute(g, P, gof.empty() ? nullptr : &gof, gof.empty() ? nullptr : &gb);
    } else {
        Compiled c = compile(g, P);
        auto it = g.byVertexLabel.find(c.label0);
        if (c.ok && it != g.byVertexLabel.end()) {
            const std::vector<unsigned>& starts = it->second;
            uint64_t e = startEnd == 0 ? starts.size() : std::min<uint64_t>(startEnd, starts.size());
            uint64_t b = std::min<uint64_t>(startBegin, e);
            const uint64_t st = stride ? stride : 1;
            nstart = (e - b + st - 1) / st;
            const bool wantBits = !gof.empty();
            int nt = threads > 0 ? threads : omp_get_max_threads();
            std::vector<std::vector<uint8_t>> tb(wantBits ? nt : 0, std::vector<uint8_t>(ngraphs, 0));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16) reduction(+ : count, bytes)
            for (uint64_t j = 0; j < nstart; ++j) {
                const uint64_t i = b + j * st;
                unsigned row[32];
                row[0] = starts[i];
                Acc acc;
                dfs(g, c, 1, row, acc, true);
                count += acc.count;
                bytes += acc.bytes;
                if (wantBits && acc.count) tb[omp_get_thread_num()][gof[row[0]]] = 1;
            }
            for (auto& v : tb)
                for (uint32_t i = 0; i < ngraphs; ++i) gb[i] |= v[i];
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (countOut) *countOut = count;
    if (bytesOut) *bytesOut = bytes;
    if (nstartOut) *nstartOut = nstart;
    if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
    if (bits && !gb.empty()) memcpy(bits, gb.data(), ngraphs);
    return 0;
}

// The embeddings of the faithful mode, one row per embedding with column u
// holding the data vertex of query vertex u, rows sorted ascending
// (lexicographic).  Returns the number of rows; out receives them when
// cap >= rows * k (k = query vertices, written to *kOut).  Negative on a
// malformed query.
int64_t oracle_rows(void* h, const char* query, uint32_t* out, uint64_t cap, uint32_t* kOut) {
    Graph& g = *(Graph*)h;
    OPlan P;
    std::string err;
    if (!planQuery(query, P, err)) return -1;
    std::vector<unsigned> rows;
    const uint64_t n = faithful(g, P, nullptr, nullptr, &rows);
    const unsigned k = P.vnum;
    if (kOut) *kOut = k;
    std::vector<uint64_t> idx(n);
    for (uint64_t i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
        return std::lexicographical_compare(rows.begin() + a * k, rows.begin() + (a + 1) * k,
                                            rows.begin() + b * k, rows.begin() + (b + 1) * k);
    });
    if (out && cap >= n * k)
        for (uint64_t i = 0; i < n; ++i)
            for (unsigned j = 0; j < k; ++j) out[i * k + j] = rows[idx[i] * k + j];
    return (int64_t)n;
}

}  // extern "C"
