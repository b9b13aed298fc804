// Plain mining path over the C ABI.
#include "miner.h"

#include <cmath>
#include <cstring>

#include <algorithm>
#include <chrono>
#include <sstream>

namespace geneva {

namespace {
constexpr uint64_t kCountChunk = 1ull << 31;  // candidates per counting launch
// internal status of a query stopped at its time limit
constexpr int kTimedOut = -100;
double nowS() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

bool PlainMiner::timedOut() const { return deadline_ > 0 && nowS() > deadline_; }

int PlainMiner::snapshot() {
    if (deadline_ <= 0) return GV_OK;
    return gv_counters_read(ctx_, &snapCount_, &snapBytes_, snapBits_.empty() ? nullptr : snapBits_.data(),
                            (uint32_t)snapBits_.size());
}

bool parseQueryPattern(const std::string& text, QueryPattern& q, std::string* err) {
    std::istringstream in(text);
    std::string line;
    std::vector<std::pair<uint32_t, uint32_t>> vs;
    std::vector<uint32_t> es;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream ls(line.substr(1));
        if (line[0] == 'v') {
            uint32_t id, lab;
            if (ls >> id >> lab) vs.emplace_back(id, lab);
        } else if (line[0] == 'e') {
            uint32_t a, b, l;
            if (ls >> a >> b >> l) {
                es.push_back(a);
                es.push_back(b);
                es.push_back(l);
            }
        }
    }
    q = QueryPattern();
    q.vnum = (uint32_t)vs.size();
    if (q.vnum < 2 || q.vnum > GV_MAXK) {
        if (err) *err = "query must have 2.." + std::to_string(GV_MAXK) + " vertices";
        return false;
    }
    q.labels.assign(q.vnum, 0);
    std::vector<bool> seen(q.vnum, false);
    for (auto& p : vs) {
        if (p.first >= q.vnum || seen[p.first]) {
            if (err) *err = "query vertex ids must be 0..n-1, each once";
            return false;
        }
        seen[p.first] = true;
        q.labels[p.first] = p.second;
    }
    q.adj.assign(q.vnum * q.vnum, 0xffffffffu);
    for (size_t i = 0; i < es.size(); i += 3) {
        uint32_t a = es[i], b = es[i + 1], l = es[i + 2];
        if (a >= q.vnum || b >= q.vnum || a == b) {
            if (err) *err = "bad query edge";
            return false;
        }
        if (q.adj[a * q.vnum + b] == 0xffffffffu) q.adj[a * q.vnum + b] = q.adj[b * q.vnum + a] = l;
    }
    return true;
}

bool compilePlan(const std::vector<Phase>& mo, const std::map<uint, uint>& restricts,
                 const std::map<uint, uint>& queryLabel, const QueryPattern& qp, gv_graph_t g,
                 CompiledQuery& out, std::string* err) {
    out = CompiledQuery();
    const uint32_t n = qp.vnum;
    std::vector<uint32_t> col(n, 0xffffffffu);
    std::vector<uint32_t> covered(n * n, 0);
    for (auto& ph : mo) {
        if (ph.phaseType == 'E' && ph.pairs.size() == 4) out.dvPhases++;
        out.nphases++;
        for (size_t i = 0; i + 1 < ph.pairs.size(); i += 2) {
            uint32_t s = ph.pairs[i], x = ph.pairs[i + 1];
            if (s >= n || x >= n) {
                if (err) *err = "plan names a vertex outside the query";
                return false;
            }
            covered[s * n + x]++;
            covered[x * n + s]++;
            if (ph.phaseType != 'E') continue;
            if (out.order.empty()) {
                col[s] = 0;
                out.order.push_back(s);
            }
            if (col[x] == 0xffffffffu) {
                col[x] = (uint32_t)out.order.size();
                out.order.push_back(x);
            }
        }
    }
    if (out.order.size() != n) {
        if (err) *err = "plan does not place every query vertex";
        return false;
    }
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t b = 0; b < n; ++b) {
            bool edge = qp.adj[a * n + b] != 0xffffffffu;
            if ((edge && covered[a * n + b] != 1) || (!edge && covered[a * n + b] != 0)) {
                if (err) *err = "plan does not cover every query edge exactly once";
                return false;
            }
        }
    auto dense = [&](uint32_t label, uint32_t& d) { return gv_graph_label_index(g, label, &d) == GV_OK; };
    gv_query_t& q = out.q;
    q.k = n;
    if (!dense(queryLabel.at(out.order[0]), q.label0)) out.empty = true;
    for (uint32_t t = 1; t < n; ++t) {
        const uint32_t u = out.order[t];
        gv_step_t& s = q.step[t];
        s.col = t;
        s.restrict_col = GV_NONE;
        if (!dense(queryLabel.at(u), s.label)) out.empty = true;
        for (uint32_t w = 0; w < n; ++w) {
            uint32_t el = qp.adj[u * n + w];
            if (el == 0xffffffffu || col[w] >= t) continue;
            uint32_t part = 0;
            if (gv_graph_part_index(g, el, &part) != GV_OK) out.empty = true;
            s.back_col[s.nback] = (uint8_t)col[w];
            s.back_part[s.nback] = (uint8_t)part;
            s.nback++;
        }
        if (s.nback == 0) {
            if (err) *err = "vertex without an earlier neighbour in the plan";
            return false;
        }
        auto it = restricts.find(u);
        if (it != restricts.end() && it->second != 0xffffffffu) {
            if (it->second >= n || col[it->second] >= t) {
                if (err) *err = "restriction refers to a later vertex";
                return false;
            }
            s.restrict_col = (uint8_t)col[it->second];
        }
        for (uint32_t j = 0; j < t; ++j)
            if (queryLabel.at(out.order[j]) == queryLabel.at(u)) s.coll_col[s.ncoll++] = (uint8_t)j;
    }
    uint32_t maxDeg = 0;
    gv_graph_max_degree(g, &maxDeg);
    out.hasLeaves = !out.empty && leafPlan(q, maxDeg, out.leaves);
    return true;
}

bool planAndCompile(const std::string& text, bool singleV, gv_graph_t g, CompiledQuery& out,
                    std::string* err) {
    QueryPattern qp;
    if (!parseQueryPattern(text, qp, err)) return false;
    std::map<uint64_t, uint> repeatFlag;
    std::map<uint, uint> restricts, queryLabel;
    std::map<uint, std::map<uint, bool>> rls;
    try {
        std::istringstream in(text);
        if (singleV) {
            std::vector<PhaseSingleV> mo;
            planSV(in, true, mo, repeatFlag, restricts, rls, queryLabel, nullptr, nullptr, nullptr,
                   nullptr);
            std::vector<Phase> asPhases;
            for (auto& p : mo) {
                Phase q;
                q.phaseType = p.phaseType;
                q.edgeLabel = p.edgeLabel;
                q.pairs = p.pairs;
                q.reducePos = p.reducePos;
                asPhases.push_back(q);
            }
            return compilePlan(asPhases, restricts, queryLabel, qp, g, out, err);
        }
        std::vector<Phase> mo;
        planDV(in, true, mo, repeatFlag, restricts, rls, queryLabel, nullptr, nullptr, nullptr,
               nullptr);
        return compilePlan(mo, restricts, queryLabel, qp, g, out, err);
    } catch (const PlanError& e) {
        if (err) *err = e.what;
        return false;
    }
}

PlainMiner::PlainMiner(gv_ctx_t ctx, gv_graph_t g, uint64_t levelBytes)
    : ctx_(ctx), g_(g), levelBytes_(levelBytes) {}

PlainMiner::~PlainMiner() {
    for (auto& l : lv_) {
        if (l.rows) gv_free(ctx_, l.rows);
        if (l.fr) gv_frontier_destroy(l.fr);
    }
    if (shardRows_) gv_free(ctx_, shardRows_);
}

int PlainMiner::ensureFrontier(uint32_t width, uint64_t n) {
    Level& l = lv_[width];
    if (l.fr && l.frCap >= n) return GV_OK;
    if (l.fr) gv_frontier_destroy(l.fr);
    l.fr = nullptr;
    l.frCap = 0;
    uint64_t cap = std::max<uint64_t>(n, 1u << 16);
    int rc = gv_frontier_create(ctx_, cap, &l.fr);
    if (rc == GV_OK) l.frCap = cap;
    return rc;
}

int PlainMiner::ensureRows(uint32_t width) {
    Level& l = lv_[width];
    if (l.rows) return GV_OK;
    uint64_t cap = levelBytes_ / (4ull * width + 17ull);
    if (cap < 1024) cap = 1024;
    int rc = gv_malloc(ctx_, cap * 4ull * width, (void**)&l.rows);
    if (rc != GV_OK) return rc;
    l.cap = cap;
    return ensureFrontier(width, cap);
}

int PlainMiner::run(const gv_query_t& q, uint32_t width, const uint32_t* rows, uint64_t n,
                    MineResult& out) {
    if (n == 0) return GV_OK;
    const gv_step_t& st = q.step[width];
    const bool last = width + 1 == q.k;
    const bool fusedTail = width + 2 == q.k && q.step[width + 1].nback == 1;
    if (last || fusedTail) {
        // a single back edge: counted per row, no candidate walk
        const gv_step_t* tail = fusedTail ? &q.step[width + 1] : nullptr;
        if (countByRows_ && gv_count_rows_mode(&st, tail, width) >= 0) {
            out.launches++;
            return gv_count_rows(ctx_, g_, &st, tail, rows, width, n);
        }
    }
    int rc = ensureFrontier(width, n);
    if (rc != GV_OK) return rc;
    gv_frontier_t fr = lv_[width].fr;
    uint64_t total = 0;
    rc = gv_size_rows(ctx_, g_, &st, rows, width, n, fr, &total);
    if (rc != GV_OK || total == 0) return rc;
    if (width == 1) l1Total_ = total;
    if (last || fusedTail) {
        // count-only: no output table, so any range works; launches are cut at
        // kCountChunk candidates to keep each kernel short
        const gv_step_t* tail = fusedTail ? &q.step[width + 1] : nullptr;
        for (uint64_t c0 = 0; c0 < total; c0 += kCountChunk) {
            if (timedOut()) return kTimedOut;
            const uint64_t c1 = std::min(total, c0 + kCountChunk);
            out.launches++;
            rc = gv_count_emb(ctx_, g_, &st, tail, rows, width, n, fr, c0, c1);
            if (rc == GV_OK && deadline_ > 0) rc = gv_ctx_synchronize(ctx_);
            if (rc == GV_OK && width == 1) {
                l1Done_ = c1;
                rc = snapshot();
            }
            if (rc != GV_OK) return rc;
        }
        return GV_OK;
    }
    rc = ensureRows(width + 1);
    if (rc != GV_OK) return rc;
    Level& nx = lv_[width + 1];
    // under a time limit the first extension goes in at least 256 pieces, so
    // a stopped query reports its progress in steps of at most 1/256
    const uint64_t piece = deadline_ > 0 && width == 1
                               ? std::min<uint64_t>(nx.cap, std::max<uint64_t>(1, (total + 255) / 256))
                               : nx.cap;
    for (uint64_t c0 = 0; c0 < total; c0 += piece) {
        if (timedOut()) return kTimedOut;
        uint64_t c1 = std::min(total, c0 + piece);
        uint64_t written = 0;
        out.launches++;
        out.chunks++;
        rc = gv_ext_emb(ctx_, g_, &st, rows, width, n, fr, c0, c1, nx.rows, nx.cap, &written);
        if (rc != GV_OK) return rc;
        out.peakRows = std::max(out.peakRows, written);
        rc = run(q, width + 1, nx.rows, written, out);
        if (rc == GV_OK && width == 1) {
            l1Done_ = c1;
            rc = snapshot();
        }
        if (rc != GV_OK) return rc;
    }
    return GV_OK;
}

int PlainMiner::mine(const CompiledQuery& cq, MineResult& out) {
    return mine(cq, Shard(), out);
}

// Leaf groups over the columns width .. k-1 (gv_leaves_t): every vertex there
// hangs by one edge from a core column; vertices sharing anchor, partition and
// label form a group whose first member may be restricted to a core column
// and whose later members are each restricted to the member before (the
// planner's chain); groups carry pairwise different labels.  C(s, m) must be
// exact in 64 bits for s up to the largest degree.
static bool leafGroups(const gv_query_t& q, uint32_t width, uint32_t maxDeg, gv_leaves_t& d) {
    memset(&d, 0, sizeof(d));
    d.width = width;
    d.k = q.k;
    std::vector<std::vector<uint32_t>> members(GV_MAX_LEAF_GROUPS);
    for (uint32_t t = width; t < q.k; ++t) {
        const gv_step_t& s = q.step[t];
        if (s.nback != 1 || s.back_col[0] >= width) return false;
        int gi = -1;
        for (uint32_t i = 0; i < d.ngroups; ++i) {
            if (d.group[i].label != s.label) continue;
            if (d.group[i].anchor_col != s.back_col[0] || d.group[i].part != s.back_part[0]) return false;
            gi = (int)i;
        }
        std::vector<uint32_t> core, leaf;
        for (uint32_t j = 0; j < s.ncoll; ++j) (s.coll_col[j] < width ? core : leaf).push_back(s.coll_col[j]);
        if (gi < 0) {
            if (d.ngroups == GV_MAX_LEAF_GROUPS || !leaf.empty()) return false;
            if (s.restrict_col != GV_NONE && s.restrict_col >= width) return false;
            gv_leaf_group_t& gr = d.group[d.ngroups];
            gr.anchor_col = s.back_col[0];
            gr.part = s.back_part[0];
            gr.label = s.label;
            gr.m = 1;
            gr.restrict_col = s.restrict_col;
            gr.ncoll = (uint8_t)core.size();
            for (size_t j = 0; j < core.size(); ++j) gr.coll_col[j] = (uint8_t)core[j];
            gi = (int)d.ngroups++;
        } else {
            gv_leaf_group_t& gr = d.group[gi];
            if (core.size() != gr.ncoll) return false;
            for (size_t j = 0; j < core.size(); ++j)
                if (core[j] != gr.coll_col[j]) return false;
            if (leaf != members[gi] || s.restrict_col != members[gi].back()) return false;
            gr.m++;
        }
        members[gi].push_back(t);
        d.group_of[t - width] = (uint8_t)gi;
    }
    for (uint32_t i = 0; i < d.ngroups; ++i)
        if (d.group[i].m * std::log2((double)maxDeg + 1.0) >= 63.0) return false;
    return d.ngroups > 0;
}

// The narrowest core whose remaining vertices are leaf groups.
bool leafPlan(const gv_query_t& q, uint32_t maxDeg, gv_leaves_t& d) {
    for (uint32_t w = 1; w < q.k; ++w)
        if (leafGroups(q, w, maxDeg, d)) return true;
    return false;
}

// The whole query with extents kept in device memory (gv_mine_levels): one
// synchronisation at the end.  Tables that outgrow a level are detected on the
// device; the query is then mined again, chunked, by run().
int PlainMiner::runLevels(const CompiledQuery& cq, const uint32_t* rows, uint64_t n,
                          MineResult& out) {
    if (n == 0) return GV_OK;
    const gv_query_t& q = cq.q;
    uint32_t* tables[GV_MAXK + 1] = {};
    uint64_t caps[GV_MAXK + 1] = {};
    gv_frontier_t frs[GV_MAXK + 1] = {};
    int rc = ensureFrontier(1, n);
    if (rc != GV_OK) return rc;
    frs[1] = lv_[1].fr;
    const bool useLeaves = countByRows_ && cq.hasLeaves;
    const gv_leaves_t& leaves = cq.leaves;
    for (uint32_t w = 2; w < q.k; ++w) {
        // the last vertex, when it has one back edge, is counted with the step
        // before it, so width k - 1 is then never materialised; leaf groups
        // are counted on their core
        if (w + 1 == q.k && q.step[w].nback == 1) break;
        if (useLeaves && w > leaves.width) break;
        rc = ensureRows(w);
        if (rc != GV_OK) return rc;
        tables[w] = lv_[w].rows;
        caps[w] = lv_[w].cap;
        frs[w] = lv_[w].fr;
    }
    uint32_t words = 0;
    gv_graph_owners(g_, &words);
    words = (words + 31) / 32;
    uint32_t* bits = out.graphBits.empty() ? nullptr : out.graphBits.data();
    const uint32_t nwords = (uint32_t)out.graphBits.size();
    rc = gv_mine_levels(ctx_, g_, &q, rows, n, tables, caps, frs, countByRows_ ? 1 : 0,
                        useLeaves ? &leaves : nullptr, &out.count, &out.algBytes, bits, nwords);
    out.launches++;
    if (rc != GV_ERR_CAPACITY) return rc;
    // a table outgrew its level: start over, chunked
    out.overflowRetries++;
    rc = gv_counters_reset(ctx_, g_);
    if (rc == GV_OK) rc = run(q, 1, rows, n, out);
    if (rc == GV_OK) rc = gv_counters_read(ctx_, &out.count, &out.algBytes, bits, nwords);
    return rc;
}

int PlainMiner::mine(const CompiledQuery& cq, const Shard& sh, MineResult& out) {
    out = MineResult();
    if (sh.stride == 0) return out.status = GV_ERR_ARG;
    uint32_t ngraphs = 0;
    gv_graph_owners(g_, &ngraphs);
    const uint32_t words = (ngraphs + 31) / 32;
    out.graphBits.assign(words, 0);
    int rc = GV_OK;
    gv_mark(ctx_, 0);
    bool counted = false;   // out holds the results already
    if (!cq.empty) {
        const uint32_t* rows = nullptr;
        uint64_t n = 0;
        rc = gv_graph_start_rows(g_, cq.q.label0, &rows, &n);
        if (rc == GV_OK) {
            const uint64_t e = sh.end == 0 ? n : std::min(sh.end, n);
            const uint64_t b = std::min(sh.begin, e);
            if (sh.stride == 1) {
                rows += b;
                n = e - b;
            } else {
                // gather the shard's vertices into a table of its own
                std::vector<uint32_t> all(e), pick;
                if (e) rc = gv_memcpy_d2h(ctx_, all.data(), rows, e * 4ull);
                for (uint64_t i = b; rc == GV_OK && i < e; i += sh.stride) pick.push_back(all[i]);
                if (rc == GV_OK && pick.size() > shardCap_) {
                    if (shardRows_) gv_free(ctx_, shardRows_);
                    shardRows_ = nullptr;
                    shardCap_ = 0;
                    rc = gv_malloc(ctx_, pick.size() * 4ull, (void**)&shardRows_);
                    if (rc == GV_OK) shardCap_ = pick.size();
                }
                if (rc == GV_OK && !pick.empty())
                    rc = gv_memcpy_h2d(ctx_, shardRows_, pick.data(), pick.size() * 4ull);
                rows = shardRows_;
                n = pick.size();
            }
        }
        if (rc == GV_OK && n && limitS_ > 0) {
            // under a time limit: the chunked path, launch by launch
            deadline_ = nowS() + limitS_;
            l1Total_ = l1Done_ = 0;
            snapCount_ = snapBytes_ = 0;
            snapBits_.assign(out.graphBits.size(), 0);
            uint32_t* bits = out.graphBits.empty() ? nullptr : out.graphBits.data();
            rc = gv_counters_reset(ctx_, g_);
            if (rc == GV_OK) rc = run(cq.q, 1, rows, n, out);
            deadline_ = 0;
            if (rc == kTimedOut) {
                // what the completed first-extension chunks counted
                out.finished = false;
                out.fraction = l1Total_ ? (double)l1Done_ / (double)l1Total_ : 0.0;
                out.count = snapCount_;
                out.algBytes = snapBytes_;
                out.graphBits = snapBits_;
                rc = gv_ctx_synchronize(ctx_);
            } else if (rc == GV_OK) {
                rc = gv_counters_read(ctx_, &out.count, &out.algBytes, bits, (uint32_t)out.graphBits.size());
            }
            counted = rc == GV_OK;
        } else if (rc == GV_OK && n) {
            rc = runLevels(cq, rows, n, out);
            counted = rc == GV_OK;
        }
        if (rc != GV_OK) {
            out.status = rc;
            out.error = gv_last_error();
            return rc;
        }
    }
    gv_mark(ctx_, 1);
    float ms = 0;
    gv_elapsed_ms(ctx_, 0, 1, &ms);
    out.ms = ms;
    if (!counted) {
        // empty query or no start vertex: nothing was counted
        out.count = 0;
        out.algBytes = 0;
    }
    out.status = GV_OK;
    return GV_OK;
}

int PlainMiner::mineText(const std::string& text, bool singleV, MineResult& out) {
    CompiledQuery cq;
    std::string err;
    if (!planAndCompile(text, singleV, g_, cq, &err)) {
        out = MineResult();
        out.status = GV_ERR_ARG;
        out.error = err;
        return GV_ERR_ARG;
    }
    return mine(cq, out);
}

}  // namespace geneva
