// Batched mining over gv_batch_*.
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <mutex>

namespace geneva {

GraphStats GraphStats::of(gv_graph_t g) {
    GraphStats s;
    uint32_t np = 0;
    gv_graph_info(g, nullptr, &s.nlabels, &np, nullptr, nullptr);
    s.vertices.assign(s.nlabels, 0);
    gv_graph_label_stats(g, 0, s.vertices.data(), nullptr);
    for (uint32_t p = 0; p < std::max<uint32_t>(np, 1); ++p) {
        std::vector<uint64_t> e((uint64_t)s.nlabels * s.nlabels, 0);
        gv_graph_label_stats(g, p, nullptr, e.data());
        s.pairEdges.push_back(e);
    }
    return s;
}

uint64_t estimatePeakBytes(const CompiledQuery& cq, const GraphStats& st) {
    if (cq.empty) return 0;
    const gv_query_t& q = cq.q;
    std::vector<uint32_t> colLabel(q.k);
    colLabel[0] = q.label0;
    for (uint32_t t = 1; t < q.k; ++t) colLabel[t] = q.step[t].label;
    double rows = (double)st.vertices[q.label0];
    double peak = 0;
    for (uint32_t t = 1; t < q.k; ++t) {
        const bool last = t + 1 == q.k;
        const bool fused = t + 2 == q.k && q.step[t + 1].nback == 1;
        if (last || fused) break;
        if (cq.hasLeaves && t >= cq.leaves.width) break;   // counted on the core
        const gv_step_t& s = q.step[t];
        double best = 1e300;
        for (uint32_t i = 0; i < s.nback; ++i) {
            const uint32_t a = colLabel[s.back_col[i]];
            const double va = std::max<double>(1.0, (double)st.vertices[a]);
            const double e = (double)st.pairEdges[s.back_part[i]][(uint64_t)a * st.nlabels + s.label];
            best = std::min(best, e / va);
        }
        rows *= best;
        peak = std::max(peak, rows * 4.0 * (t + 1));
    }
    return peak > 1.8e19 ? UINT64_MAX : (uint64_t)peak;
}

BatchMiner::BatchMiner(gv_ctx_t ctx, gv_graph_t g, PlainMiner* fallback, const BatchOptions& opt)
    : ctx_(ctx), g_(g), fallback_(fallback), opt_(opt) {}

BatchMiner::~BatchMiner() {
    if (b_) gv_batch_destroy(b_);
}

int BatchMiner::init() {
    if (b_) return GV_OK;
    uint32_t ng = 0;
    gv_graph_owners(g_, &ng);
    nwords_ = (ng + 31) / 32;
    return gv_batch_create(ctx_, g_, opt_.slots, opt_.slotBytes, opt_.rowCap, &opt_.tune, &b_);
}

int BatchMiner::setTune(const gv_tune_t& t) {
    int rc = gv_batch_set_tune(b_, &t);
    if (rc == GV_OK) opt_.tune = t;
    return rc;
}

std::string BatchMiner::classKey(const CompiledQuery& cq) {
    uint32_t cyc = 0;
    for (uint32_t t = 1; t < cq.q.k; ++t) cyc += cq.q.step[t].nback > 1;
    return "k" + std::to_string(cq.q.k) + "c" + std::to_string(cyc) + (cq.hasLeaves ? "l" : "");
}

int BatchMiner::admit(const CompiledQuery& cq, uint64_t tag, uint32_t* slot) {
    uint32_t ipt = 0;
    if (!classIpt_.empty()) {
        auto it = classIpt_.find(classKey(cq));
        if (it != classIpt_.end() && it->second <= opt_.tune.ipt) ipt = it->second;
    }
    return gv_batch_admit_tuned(b_, &cq.q, cq.hasLeaves ? &cq.leaves : nullptr, ipt, tag, slot);
}

int BatchMiner::run(const std::vector<CompiledQuery>& qs, std::vector<MineResult>& out,
                    BatchStats* st) {
    int rc = init();
    if (rc != GV_OK) return rc;
    auto w0 = std::chrono::steady_clock::now();
    out.assign(qs.size(), MineResult());
    std::deque<size_t> pending;
    for (size_t i = 0; i < qs.size(); ++i) {
        if (qs[i].empty) {
            out[i].graphBits.assign(nwords_, 0);
            continue;
        }
        pending.push_back(i);
    }
    std::vector<size_t> slotQuery(opt_.slots, SIZE_MAX);
    std::vector<size_t> overflow;
    std::vector<gv_slot_info_t> info(opt_.slots);
    std::vector<uint32_t> bits;
    BatchStats local;
    uint32_t running = 0;
    gv_mark(ctx_, 2);
    while (!pending.empty() || running > 0) {
        uint32_t maxk = 0;
        while (!pending.empty()) {
            uint32_t slot = 0;
            const CompiledQuery& cq = qs[pending.front()];
            rc = admit(cq, pending.front(), &slot);
            if (rc == GV_ERR_CAPACITY) break;
            if (rc != GV_OK) return rc;
            slotQuery[slot] = pending.front();
            maxk = std::max(maxk, qs[pending.front()].q.k);
            pending.pop_front();
            ++running;
        }
        // every admitted query needs at most k-1 iterations; deferred ones loop again
        const uint32_t iters = std::max<uint32_t>(1, maxk ? maxk - 1 : 1);
        rc = gv_batch_run(b_, iters);
        if (rc != GV_OK) return rc;
        local.graphLaunches += iters;
        rc = gv_batch_poll(b_, info.data());
        if (rc != GV_OK) return rc;
        // finished slots: their graph bits in one synchronised copy, then free
        std::vector<uint32_t> done, over;
        for (uint32_t s = 0; s < opt_.slots; ++s) {
            if (slotQuery[s] == SIZE_MAX) continue;
            if (info[s].state == GV_SLOT_DONE) done.push_back(s);
            else if (info[s].state == GV_SLOT_OVERFLOW) over.push_back(s);
        }
        bits.assign(done.size() * (size_t)nwords_, 0);
        rc = gv_batch_release_many(b_, (uint32_t)done.size(), done.data(),
                                   nwords_ ? bits.data() : nullptr, nwords_);
        if (rc != GV_OK) return rc;
        for (size_t i = 0; i < done.size(); ++i) {
            const uint32_t s = done[i];
            MineResult& r = out[slotQuery[s]];
            r.count = info[s].count;
            r.algBytes = info[s].alg_bytes;
            r.graphBits.assign(bits.begin() + i * nwords_, bits.begin() + (i + 1) * nwords_);
            slotQuery[s] = SIZE_MAX;
            --running;
            local.batched++;
        }
        rc = gv_batch_release_many(b_, (uint32_t)over.size(), over.data(), nullptr, 0);
        if (rc != GV_OK) return rc;
        for (uint32_t s : over) {
            overflow.push_back(slotQuery[s]);
            slotQuery[s] = SIZE_MAX;
            --running;
        }
    }
    gv_mark(ctx_, 3);
    float ms = 0;
    gv_elapsed_ms(ctx_, 2, 3, &ms);
    local.deviceMs = ms;
    for (size_t qi : overflow) {
        local.overflowed++;
        if (!fallback_) {
            out[qi].status = GV_ERR_CAPACITY;
            continue;
        }
        rc = fallback_->mine(qs[qi], out[qi]);
        if (rc != GV_OK) return rc;
    }
    local.wallMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    if (st) *st = local;
    return GV_OK;
}

namespace {
// Tunings measured in this process, keyed by graph, engine options and the
// sample's shape; shared by every BatchMiner, so guarded by a mutex.
std::mutex& tuneMutex() {
    static std::mutex m;
    return m;
}
struct CachedTune {
    gv_tune_t tune;
    std::map<std::string, uint32_t> classes;
};
std::map<std::string, CachedTune>& tuneCache() {
    static std::map<std::string, CachedTune> m;
    return m;
}
std::string graphKey(gv_graph_t g, const BatchOptions& o) {
    uint64_t nv1 = 0, nnbr = 0;
    uint32_t nl = 0, np = 0, md = 0;
    gv_graph_info(g, &nv1, &nl, &np, &nnbr, nullptr);
    gv_graph_max_degree(g, &md);
    return "g" + std::to_string(nv1) + "/" + std::to_string(nnbr) + "/" + std::to_string(nl) + "/" +
           std::to_string(np) + "/" + std::to_string(md) + " s" + std::to_string(o.slots) + "/" +
           std::to_string(o.slotBytes) + "/" + std::to_string(o.rowCap) + " ";
}
std::string shapeKey(const std::vector<CompiledQuery>& qs) {
    // histogram of query sizes and of multi-edge steps
    uint32_t byk[GV_MAXK + 1] = {0};
    uint32_t cyc = 0;
    for (auto& q : qs) {
        byk[q.q.k]++;
        for (uint32_t t = 1; t < q.q.k; ++t) cyc += q.q.step[t].nback > 1;
    }
    std::string k;
    for (uint32_t i = 0; i <= GV_MAXK; ++i)
        if (byk[i]) k += std::to_string(i) + ":" + std::to_string(byk[i] * 8 / std::max<size_t>(qs.size(), 1)) + ",";
    return k + "c" + std::to_string(cyc * 4 / std::max<size_t>(qs.size(), 1));
}
}  // namespace

int autotune(BatchMiner& bm, const std::vector<CompiledQuery>& sample, TuneResult& out) {
    out = TuneResult();
    const std::string key = graphKey(bm.graph(), bm.options()) + shapeKey(sample);
    {
        std::lock_guard<std::mutex> l(tuneMutex());
        auto it = tuneCache().find(key);
        if (it != tuneCache().end()) {
            out.tune = it->second.tune;
            bm.classTune() = it->second.classes;
            for (auto& kv : it->second.classes) out.classes[kv.first] = {kv.second, 0.0};
            return bm.setTune(out.tune);
        }
    }
    bm.classTune().clear();
    const uint32_t ipts[] = {1, 2, 4, 8};
    const uint32_t bpcs[] = {4, 8, 16, 32};
    double best = 1e300;
    std::vector<MineResult> res;
    for (uint32_t ipt : ipts) {
        for (uint32_t bpc : bpcs) {
            gv_tune_t t{ipt, bpc};
            int rc = bm.setTune(t);
            if (rc != GV_OK) return rc;
            BatchStats st;
            // one untimed pass warms caches and code, the second is measured
            rc = bm.run(sample, res, nullptr);
            if (rc != GV_OK) return rc;
            rc = bm.run(sample, res, &st);
            if (rc != GV_OK) return rc;
            out.tried.push_back({t, st.deviceMs});
            if (st.deviceMs < best) {
                best = st.deviceMs;
                out.tune = t;
            }
        }
    }
    out.bestMs = best;
    int rc = bm.setTune(out.tune);
    if (rc != GV_OK) return rc;
    // per class: candidates per lane up to the chosen kernel's
    std::map<std::string, std::vector<CompiledQuery>> byClass;
    for (auto& q : sample)
        if (!q.empty) byClass[BatchMiner::classKey(q)].push_back(q);
    for (auto& kv : byClass) {
        if (kv.second.size() < 4) continue;
        double cbest = 1e300;
        uint32_t cipt = out.tune.ipt;
        for (uint32_t ipt : ipts) {
            if (ipt > out.tune.ipt) break;
            bm.classTune()[kv.first] = ipt;
            BatchStats st;
            rc = bm.run(kv.second, res, nullptr);
            if (rc == GV_OK) rc = bm.run(kv.second, res, &st);
            if (rc != GV_OK) return rc;
            if (st.deviceMs < cbest) {
                cbest = st.deviceMs;
                cipt = ipt;
            }
        }
        bm.classTune()[kv.first] = cipt;
        out.classes[kv.first] = {cipt, cbest};
    }
    {
        std::lock_guard<std::mutex> l(tuneMutex());
        tuneCache()[key] = CachedTune{out.tune, bm.classTune()};
    }
    return GV_OK;
}

}  // namespace geneva
