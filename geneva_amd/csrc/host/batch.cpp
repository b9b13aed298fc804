// Batched mining over gv_batch_*.
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <mutex>
#include <tuple>

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

int BatchMiner::setRowCap(uint64_t rows) {
    int rc = gv_batch_set_row_cap(b_, rows);
    if (rc == GV_OK) rowCap_ = rows;
    return rc;
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

int BatchMiner::admit(const CompiledQuery& cq, uint64_t tag, uint32_t* slot, uint32_t part,
                      uint32_t nparts) {
    uint32_t ipt = 0;
    if (!classIpt_.empty()) {
        auto it = classIpt_.find(classKey(cq));
        if (it != classIpt_.end() && it->second <= opt_.tune.ipt) ipt = it->second;
    }
    return gv_batch_admit_part(b_, &cq.q, cq.hasLeaves ? &cq.leaves : nullptr, ipt, tag, part, nparts,
                               slot);
}

int BatchMiner::run(const std::vector<CompiledQuery>& qs, std::vector<MineResult>& out,
                    BatchStats* st) {
    int rc = init();
    if (rc != GV_OK) return rc;
    auto w0 = std::chrono::steady_clock::now();
    out.assign(qs.size(), MineResult());
    std::vector<size_t> order;
    for (size_t i = 0; i < qs.size(); ++i) {
        if (qs[i].empty) {
            out[i].graphBits.assign(nwords_, 0);
            continue;
        }
        order.push_back(i);
    }
    // queries that start from the same table (start label and first step)
    // take adjacent slots, whose sizing tiles the engine runs side by side
    auto startKey = [&](size_t i) {
        const gv_query_t& q = qs[i].q;
        const gv_step_t& s1 = q.step[1];
        return std::make_tuple(q.label0, s1.label, s1.nback ? s1.back_part[0] : 0u, s1.restrict_col, s1.nback);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return startKey(a) < startKey(b); });
    // a query whose table outgrows its slot comes back as two pieces of its
    // first table (each piece that overflows splits again), up to kMaxPieces;
    // beyond that it is mined on the plain path
    struct Piece {
        size_t q;
        uint32_t part, nparts;
    };
    constexpr uint32_t kMaxPieces = 64;
    auto queryKey = [&](size_t i) {
        const CompiledQuery& cq = qs[i];
        std::string k;
        auto put = [&](uint32_t v) { k.append(reinterpret_cast<const char*>(&v), sizeof(v)); };
        put(cq.q.k);
        put(cq.q.label0);
        for (uint32_t t = 1; t < cq.q.k; ++t) {
            const gv_step_t& st = cq.q.step[t];
            put(st.label);
            put(st.nback);
            for (uint32_t j = 0; j < st.nback; ++j) put(st.back_col[j] | st.back_part[j] << 8);
            put(st.restrict_col);
            put(st.ncoll);
            for (uint32_t j = 0; j < st.ncoll; ++j) put(st.coll_col[j]);
        }
        put(cq.hasLeaves ? cq.leaves.width : 0);
        return k;
    };
    std::vector<uint32_t> maxParts(qs.size(), 1);
    std::deque<Piece> pending;
    for (size_t i : order) {
        auto it = pieceHint_.find(queryKey(i));
        const uint32_t n = it == pieceHint_.end() ? 1 : it->second;
        for (uint32_t p = 0; p < n; ++p) pending.push_back({i, p, n});
        maxParts[i] = n;
    }
    std::vector<uint32_t> open(qs.size(), 0);   // pieces not finished
    std::vector<uint8_t> toPlain(qs.size(), 0);
    for (size_t i : order) {
        open[i] = maxParts[i];
        out[i].graphBits.assign(nwords_, 0);
    }
    std::vector<Piece> slotPiece(opt_.slots, Piece{SIZE_MAX, 0, 0});
    std::vector<gv_slot_info_t> info(opt_.slots);
    std::vector<uint32_t> bits;
    BatchStats local;
    uint32_t running = 0;
    gv_mark(ctx_, 2);
    while (!pending.empty() || running > 0) {
        while (!pending.empty()) {
            uint32_t slot = 0;
            const Piece pc = pending.front();
            if (toPlain[pc.q]) {   // an earlier piece gave up: the plain path mines it whole
                pending.pop_front();
                continue;
            }
            rc = admit(qs[pc.q], pc.q, &slot, pc.part, pc.nparts);
            if (rc == GV_ERR_CAPACITY) break;
            if (rc != GV_OK) return rc;
            slotPiece[slot] = pc;
            pending.pop_front();
            ++running;
        }
        // as many iterations as the running slot furthest from its end needs
        // (deferred slots need more: they come round again)
        uint32_t iters = 1;
        for (uint32_t s = 0; s < opt_.slots; ++s) {
            if (slotPiece[s].q == SIZE_MAX) continue;
            uint32_t n = 1;
            rc = gv_batch_slot_iterations(b_, s, &n);
            if (rc != GV_OK) return rc;
            iters = std::max(iters, n);
        }
        rc = gv_batch_run(b_, iters);
        if (rc != GV_OK) return rc;
        local.graphLaunches += iters;
        rc = gv_batch_poll(b_, info.data());
        if (rc != GV_OK) return rc;
        // finished slots: their graph bits in one synchronised copy, then free
        std::vector<uint32_t> done, over;
        for (uint32_t s = 0; s < opt_.slots; ++s) {
            if (slotPiece[s].q == SIZE_MAX) continue;
            if (info[s].state == GV_SLOT_DONE) done.push_back(s);
            else if (info[s].state == GV_SLOT_OVERFLOW) over.push_back(s);
        }
        bits.assign(done.size() * (size_t)nwords_, 0);
        rc = gv_batch_release_many(b_, (uint32_t)done.size(), done.data(),
                                   nwords_ ? bits.data() : nullptr, nwords_);
        if (rc != GV_OK) return rc;
        for (size_t i = 0; i < done.size(); ++i) {
            const uint32_t s = done[i];
            const size_t q = slotPiece[s].q;
            MineResult& r = out[q];
            r.count += info[s].count;
            r.algBytes += info[s].alg_bytes;
            for (uint32_t w = 0; w < nwords_; ++w) r.graphBits[w] |= bits[i * nwords_ + w];
            slotPiece[s].q = SIZE_MAX;
            --running;
            if (--open[q] == 0 && !toPlain[q]) local.batched++;
        }
        rc = gv_batch_release_many(b_, (uint32_t)over.size(), over.data(), nullptr, 0);
        if (rc != GV_OK) return rc;
        for (uint32_t s : over) {
            const Piece pc = slotPiece[s];
            slotPiece[s].q = SIZE_MAX;
            --running;
            if (!toPlain[pc.q] && pc.nparts < kMaxPieces) {
                pending.push_front({pc.q, 2 * pc.part + 1, 2 * pc.nparts});
                pending.push_front({pc.q, 2 * pc.part, 2 * pc.nparts});
                ++open[pc.q];
                maxParts[pc.q] = std::max(maxParts[pc.q], 2 * pc.nparts);
                local.pieces++;
            } else {
                toPlain[pc.q] = 1;
                --open[pc.q];
            }
        }
    }
    std::vector<size_t> overflow;
    for (size_t i : order) {
        if (toPlain[i]) overflow.push_back(i);
        else if (maxParts[i] > 1) pieceHint_[queryKey(i)] = maxParts[i];
    }
    gv_mark(ctx_, 3);
    float ms = 0;
    gv_elapsed_ms(ctx_, 2, 3, &ms);
    local.deviceMs = ms;
    const auto f0 = std::chrono::steady_clock::now();
    for (size_t qi : overflow) {
        local.overflowed++;
        out[qi] = MineResult();
        if (!fallback_) {
            out[qi].status = GV_ERR_CAPACITY;
            continue;
        }
        rc = fallback_->mine(qs[qi], out[qi]);
        if (rc != GV_OK) return rc;
    }
    local.fallbackMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - f0).count();
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
    uint64_t rowCap;
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
            out.rowCap = it->second.rowCap;
            bm.classTune() = it->second.classes;
            for (auto& kv : it->second.classes) out.classes[kv.first] = {kv.second, 0.0};
            int rc = bm.setTune(out.tune);
            return rc != GV_OK ? rc : bm.setRowCap(out.rowCap);
        }
    }
    bm.classTune().clear();
    int rc0 = bm.setRowCap(bm.options().rowCap);
    if (rc0 != GV_OK) return rc0;
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
    // rows sized per iteration: the allocation's.  A smaller cap timed on
    // the sample does not carry over to a larger batch, whose first
    // iterations it would defer.
    out.rowCap = bm.options().rowCap;
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
        tuneCache()[key] = CachedTune{out.tune, out.rowCap, bm.classTune()};
    }
    return GV_OK;
}

}  // namespace geneva
