// extern "C" surface of the host library (include/geneva_host.h).
#include "../../../include/geneva_host.h"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>

#include "../../../include/geneva_hip.h"
#include "gen.h"
#include "genMatchOrder.h"
#include "loadDataBase.h"
#include "batch.h"
#include "miner.h"
#include "phase_driver.h"

namespace {
thread_local std::string g_err;
void setErr(const std::string& s) { g_err = s; }
}  // namespace

struct gh_dataset {
    geneva::Dataset ds;
};

namespace {
// Host threads that plan a list of queries: idle between calls, each call's
// queries taken one at a time from a shared counter.
class PlanPool {
public:
    explicit PlanPool(uint32_t n) {
        for (uint32_t t = 0; t < n; ++t) th_.emplace_back([this] { loop(); });
    }
    ~PlanPool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    // Run f(i) for i in [0, n) on the pool and the calling thread.
    void run(uint32_t n, const std::function<void(uint32_t)>& f) {
        {
            std::lock_guard<std::mutex> l(m_);
            f_ = &f;
            n_ = n;
            next_.store(0);
            busy_ = (uint32_t)th_.size();
            ++gen_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [this] { return busy_ == 0; });
        f_ = nullptr;
    }

private:
    void drain() {
        for (uint32_t i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) (*f_)(i);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            drain();
            {
                std::lock_guard<std::mutex> l(m_);
                if (--busy_ == 0) done_.notify_all();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(uint32_t)>* f_ = nullptr;
    uint32_t n_ = 0;
    std::atomic<uint32_t> next_{0};
    uint32_t busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};
}  // namespace

struct gh_engine {
    gv_ctx_t ctx = nullptr;
    gv_graph_t graph = nullptr;
    std::unique_ptr<geneva::PlainMiner> plain;
    std::unique_ptr<geneva::BatchMiner> batch;
    geneva::BatchOptions batchOpt;
    std::unique_ptr<geneva::GraphStats> stats;
    uint32_t ngraphs = 0;
    int device = 0;
    uint64_t levelBytes = 0;
    // the serving loop's plain-path worker: its own context (own stream)
    gv_ctx_t sideCtx = nullptr;
    std::unique_ptr<geneva::PlainMiner> side;
    // progress of the last plain-path query (time limit)
    bool lastFinished = true;
    double lastFraction = 1.0;
    // per-class tuning of the last autotune: "class:ipt:ms;" entries
    std::string tuneTable;
    std::unique_ptr<PlanPool> planPool;   // made on the first list of queries
};

namespace {
int ensureBatch(gh_engine* e) {
    if (!e->batch) {
        e->batch.reset(new geneva::BatchMiner(e->ctx, e->graph, e->plain.get(), e->batchOpt));
        int rc = e->batch->init();
        if (rc != GV_OK) {
            e->batch.reset();
            return rc;
        }
    }
    return GV_OK;
}
// Plans and compiles queries on up to 16 host threads (each plan is
// independent; the graph is only read).
int compileAll(gh_engine* e, const char* const* qs, uint32_t n, int single_v,
               std::vector<geneva::CompiledQuery>& out) {
    out.assign(n, geneva::CompiledQuery());
    std::vector<std::string> err(n);
    std::vector<char> ok(n, 0);
    const std::function<void(uint32_t)> one = [&](uint32_t i) {
        ok[i] = qs[i] && geneva::planAndCompile(qs[i], single_v != 0, e->graph, out[i], &err[i]);
    };
    if (n < 16) {
        for (uint32_t i = 0; i < n; ++i) one(i);
    } else {
        if (!e->planPool)
            e->planPool.reset(new PlanPool(std::min<uint32_t>(15, std::max(1u, std::thread::hardware_concurrency()) - 1)));
        e->planPool->run(n, one);
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!ok[i]) {
            setErr("query " + std::to_string(i) + ": " + err[i]);
            return GV_ERR_ARG;
        }
    return GV_OK;
}
}  // namespace

extern "C" {

const char* gh_last_error(void) { return g_err.c_str(); }

gh_dataset_t gh_ds_gen_database(uint32_t ngraphs, uint32_t nmin, uint32_t nmax, double avg_deg,
                                double ba_fraction, uint32_t ba_m, uint32_t nlabels,
                                double zipf_s, uint64_t seed) {
    if (ngraphs == 0 || nmin < 1 || nmax < nmin || nlabels == 0) {
        setErr("bad database parameters");
        return nullptr;
    }
    gh_dataset* d = new gh_dataset();
    d->ds = geneva::genDatabase(ngraphs, nmin, nmax, avg_deg, ba_fraction, ba_m, nlabels, zipf_s,
                                seed);
    return d;
}

gh_dataset_t gh_ds_gen_chung_lu(uint32_t nv, double gamma, double mean_deg, double wmax,
                                uint32_t nlabels, double zipf_s, uint64_t seed) {
    if (nv < 2 || gamma <= 1.0 || mean_deg <= 0 || nlabels == 0) {
        setErr("bad Chung-Lu parameters");
        return nullptr;
    }
    gh_dataset* d = new gh_dataset();
    d->ds = geneva::genChungLu(nv, gamma, mean_deg, wmax, nlabels, zipf_s, seed);
    return d;
}

namespace {
struct DecodeCtx {
    geneva::DataGraph* g;
    uint32_t label;
};
void collect(void* user, uint32_t src, uint32_t, const uint32_t* nbrs, uint32_t count) {
    DecodeCtx* c = (DecodeCtx*)user;
    for (uint32_t i = 0; i < count; ++i)
        if (src < nbrs[i]) c->g->edges.push_back({src, nbrs[i], c->label});
}
}  // namespace

gh_dataset_t gh_ds_load(const char* mygraph, const char* metadata) {
    std::map<uint, uint*> db = loadDataBase(mygraph, metadata);
    auto it = db.find(0xffffffffu);
    if (it == db.end()) {
        setErr("loadDataBase failed");
        return nullptr;
    }
    gh_dataset* d = new gh_dataset();
    const uint* vLabel = it->second;
    geneva::DataGraph& g = d->ds.g;
    g.nv = vLabel[0];
    g.vlabel.assign(vLabel, vLabel + g.nv + 1);
    g.vlabel[0] = 0;
    for (auto& kv : db) {
        if (kv.first == 0xffffffffu) continue;
        DecodeCtx c{&g, kv.first};
        std::string err;
        if (!geneva::forEachLabelGroup(kv.second, vLabel, g.nv, collect, &c, &err)) {
            setErr("partition " + std::to_string(kv.first) + ": " + err);
            geneva::freeDataBase(db);
            delete d;
            return nullptr;
        }
    }
    d->ds.graphOf.assign(g.nv + 1, 0);
    d->ds.ngraphs = 1;
    geneva::freeDataBase(db);
    return d;
}

gh_dataset_t gh_ds_from_arrays(uint32_t nv, const uint32_t* vlabel, uint64_t m,
                               const uint32_t* edges, const uint32_t* graph_of,
                               uint32_t ngraphs) {
    if (!vlabel || (m && !edges)) {
        setErr("null arrays");
        return nullptr;
    }
    gh_dataset* d = new gh_dataset();
    d->ds.g.nv = nv;
    d->ds.g.vlabel.assign(vlabel, vlabel + nv + 1);
    for (uint64_t i = 0; i < m; ++i)
        d->ds.g.edges.push_back({edges[3 * i], edges[3 * i + 1], edges[3 * i + 2]});
    if (graph_of && ngraphs > 0) {
        d->ds.graphOf.assign(graph_of, graph_of + nv + 1);
        d->ds.ngraphs = ngraphs;
    } else {
        d->ds.graphOf.assign(nv + 1, 0);
        d->ds.ngraphs = 1;
    }
    return d;
}

void gh_ds_free(gh_dataset_t ds) { delete ds; }

int gh_ds_add_walk_queries(gh_dataset_t d, uint32_t nq, uint32_t kmin, uint32_t kmax, int induced,
                           uint32_t min_extra, uint64_t seed, uint32_t min_deg) {
    if (!d || kmin < 2 || kmax < kmin || kmax > GV_MAXK) return GV_ERR_ARG;
    auto q = geneva::genWalkQueries(d->ds, nq, kmin, kmax, induced != 0, min_extra, seed, min_deg);
    d->ds.queries.insert(d->ds.queries.end(), q.begin(), q.end());
    return GV_OK;
}

int gh_ds_add_shape_queries(gh_dataset_t d, uint64_t seed) {
    if (!d) return GV_ERR_ARG;
    auto q = geneva::genShapeQueries(d->ds, seed);
    d->ds.queries.insert(d->ds.queries.end(), q.begin(), q.end());
    return GV_OK;
}

int gh_ds_add_query(gh_dataset_t d, const char* text) {
    if (!d || !text) return GV_ERR_ARG;
    d->ds.queries.emplace_back(text);
    return GV_OK;
}

uint32_t gh_ds_nv(gh_dataset_t d) { return d ? d->ds.g.nv : 0; }
const uint32_t* gh_ds_vlabel(gh_dataset_t d) { return d ? d->ds.g.vlabel.data() : nullptr; }
uint64_t gh_ds_nedges(gh_dataset_t d) { return d ? d->ds.g.edges.size() : 0; }
const uint32_t* gh_ds_edges(gh_dataset_t d) {
    static_assert(sizeof(geneva::Edge) == 12, "Edge must be three packed uint32");
    return d ? reinterpret_cast<const uint32_t*>(d->ds.g.edges.data()) : nullptr;
}
const uint32_t* gh_ds_graph_of(gh_dataset_t d) { return d ? d->ds.graphOf.data() : nullptr; }
uint32_t gh_ds_ngraphs(gh_dataset_t d) { return d ? d->ds.ngraphs : 0; }
uint32_t gh_ds_nqueries(gh_dataset_t d) { return d ? (uint32_t)d->ds.queries.size() : 0; }
const char* gh_ds_query(gh_dataset_t d, uint32_t i) {
    return (d && i < d->ds.queries.size()) ? d->ds.queries[i].c_str() : nullptr;
}

int gh_ds_write(gh_dataset_t d, const char* prefix, int skip_index) {
    if (!d || !prefix) return GV_ERR_ARG;
    std::string p(prefix), err;
    if (!geneva::writeGraphText(p + ".g", d->ds.g) || !geneva::writeMetadata(p + ".metadata", d->ds.g) ||
        !geneva::writeMyGraph(p + ".mygraph", d->ds.g, &err) ||
        !geneva::writeReorder(p + ".reorder", d->ds.g, skip_index != 0)) {
        setErr("write failed: " + err);
        return GV_ERR_ARG;
    }
    for (size_t i = 0; i < d->ds.queries.size(); ++i) {
        FILE* f = fopen((p + "_" + std::to_string(i) + ".query").c_str(), "w");
        if (!f) return GV_ERR_ARG;
        fputs(d->ds.queries[i].c_str(), f);
        fclose(f);
    }
    return GV_OK;
}

int gh_ds_reorder_intervals(gh_dataset_t d) {
    if (!d) return GV_ERR_ARG;
    std::vector<uint> newOf;
    std::string err;
    const uint worst = geneva::reorderForIntervals(d->ds.g, newOf, &err);
    if (worst == 0 && !err.empty()) {
        setErr("reorder: " + err);
        return GV_ERR_FORMAT;
    }
    d->ds.g = geneva::relabel(d->ds.g, newOf);
    if (d->ds.graphOf.size() == newOf.size()) {
        std::vector<uint32_t> go(newOf.size(), 0);
        for (size_t v = 1; v < newOf.size(); ++v) go[newOf[v]] = d->ds.graphOf[v];
        d->ds.graphOf.swap(go);
    }
    return (int)worst;
}

int gh_ds_partition_runs(gh_dataset_t d, uint32_t* labels, uint32_t* runs, uint32_t cap) {
    if (!d) return GV_ERR_ARG;
    auto m = geneva::partitionRuns(d->ds.g);
    uint32_t i = 0;
    for (auto& kv : m) {
        if (i < cap) {
            if (labels) labels[i] = kv.first;
            if (runs) runs[i] = kv.second;
        }
        ++i;
    }
    return (int)i;
}

const uint32_t* gh_ds_old_ids(gh_dataset_t d) {
    return d && !d->ds.g.oldId.empty() ? d->ds.g.oldId.data() : nullptr;
}

uint64_t gh_plan_dump(const char* text, int single_v, char* out, uint64_t cap) {
    if (!text) return 0;
    std::map<uint64_t, uint> rf;
    std::map<uint, uint> restricts, ql;
    std::map<uint, std::map<uint, bool>> rls;
    std::ostringstream o;
    try {
        std::istringstream in(text);
        std::vector<std::pair<char, std::pair<uint, std::vector<uint>>>> phases;
        if (single_v) {
            std::vector<PhaseSingleV> mo;
            geneva::planSV(in, false, mo, rf, restricts, rls, ql, nullptr, nullptr, nullptr, nullptr);
            for (auto& p : mo) phases.push_back({p.phaseType, {p.edgeLabel, p.pairs}});
        } else {
            std::vector<Phase> mo;
            geneva::planDV(in, false, mo, rf, restricts, rls, ql, nullptr, nullptr, nullptr, nullptr);
            for (auto& p : mo) phases.push_back({p.phaseType, {p.edgeLabel, p.pairs}});
        }
        for (auto& kv : restricts) o << "restrict: " << kv.first << "<" << kv.second << "\n";
        for (size_t i = 0; i < phases.size(); ++i) {
            o << "phase " << i << " type " << phases[i].first << " edge label:" << phases[i].second.first
              << "\n";
            const auto& pr = phases[i].second.second;
            for (size_t j = 0; j + 1 < pr.size(); j += 2) o << "svid:" << pr[j] << " evid:" << pr[j + 1] << " ";
            o << "\n";
        }
    } catch (const geneva::PlanError& e) {
        setErr(e.what);
        return 0;
    }
    std::string s = o.str();
    if (out && cap) {
        uint64_t n = std::min<uint64_t>(s.size(), cap - 1);
        memcpy(out, s.data(), n);
        out[n] = 0;
    }
    return s.size();
}

int gh_dv_phase_count(const char* text) {
    if (!text) return GV_ERR_ARG;
    std::map<uint64_t, uint> rf;
    std::map<uint, uint> restricts, ql;
    std::map<uint, std::map<uint, bool>> rls;
    std::vector<Phase> mo;
    try {
        std::istringstream in(text);
        geneva::planDV(in, true, mo, rf, restricts, rls, ql, nullptr, nullptr, nullptr, nullptr);
    } catch (const geneva::PlanError& e) {
        setErr(e.what);
        return GV_ERR_ARG;
    }
    return calDVPhaseCount(mo);
}

gh_engine_t gh_engine_create(gh_dataset_t d, int device, uint64_t level_bytes, int* status) {
    auto fail = [&](int rc, const std::string& msg) -> gh_engine_t {
        if (status) *status = rc;
        setErr(msg);
        return nullptr;
    };
    if (!d) return fail(GV_ERR_ARG, "null data set");
    std::unique_ptr<gh_engine> e(new gh_engine());
    int rc = gv_ctx_create(device, &e->ctx);
    if (rc != GV_OK) return fail(rc, std::string("gv_ctx_create: ") + gv_status_string(rc));
    std::map<uint, uint> keyOfLabel;
    std::string err;
    std::map<uint, uint*> db = geneva::buildDataBase(d->ds.g, &keyOfLabel, &err);
    if (db.empty()) {
        gv_ctx_destroy(e->ctx);
        return fail(GV_ERR_FORMAT, "buildDataBase: " + err);
    }
    std::vector<const uint32_t*> blocks;
    std::vector<uint32_t> labels;
    for (auto& kv : keyOfLabel) {
        labels.push_back(kv.first);
        blocks.push_back(db.at(kv.second));
    }
    rc = gv_graph_create(e->ctx, db.at(0xffffffffu), (uint32_t)blocks.size(), blocks.data(),
                         labels.data(), &e->graph);
    geneva::freeDataBase(db);
    if (rc != GV_OK) {
        gv_ctx_destroy(e->ctx);
        return fail(rc, std::string("gv_graph_create: ") + gv_status_string(rc) + " " + gv_last_error());
    }
    if (d->ds.ngraphs > 1) {
        rc = gv_graph_set_owner(e->graph, d->ds.graphOf.data(), d->ds.ngraphs);
        if (rc != GV_OK) {
            gv_graph_destroy(e->graph);
            gv_ctx_destroy(e->ctx);
            return fail(rc, "gv_graph_set_owner failed");
        }
        e->ngraphs = d->ds.ngraphs;
    }
    e->device = device;
    e->levelBytes = level_bytes ? level_bytes : (1ull << 30);
    e->plain.reset(new geneva::PlainMiner(e->ctx, e->graph, e->levelBytes));
    if (status) *status = GV_OK;
    return e.release();
}

void gh_engine_free(gh_engine_t e) {
    if (!e) return;
    e->batch.reset();
    e->side.reset();
    if (e->sideCtx) gv_ctx_destroy(e->sideCtx);
    e->plain.reset();
    if (e->graph) gv_graph_destroy(e->graph);
    if (e->ctx) gv_ctx_destroy(e->ctx);
    delete e;
}

uint64_t gh_engine_graph_bytes(gh_engine_t e) {
    uint64_t b = 0;
    if (e && e->graph) gv_graph_info(e->graph, nullptr, nullptr, nullptr, nullptr, &b);
    return b;
}

int gh_engine_mine(gh_engine_t e, const char* text, int single_v, uint64_t* count,
                   uint64_t* alg_bytes, double* ms, uint32_t* bits, uint32_t nwords,
                   uint64_t* launches) {
    return gh_engine_mine_shard(e, text, single_v, 0, 0, 1, count, alg_bytes, ms, bits, nwords,
                                launches);
}

int gh_engine_mine_shard(gh_engine_t e, const char* text, int single_v, uint64_t start_begin,
                         uint64_t start_end, uint64_t stride, uint64_t* count, uint64_t* alg_bytes,
                         double* ms, uint32_t* bits, uint32_t nwords, uint64_t* launches) {
    if (!e || !text || stride == 0) return GV_ERR_ARG;
    geneva::MineResult r;
    geneva::CompiledQuery cq;
    std::string err;
    int rc = GV_OK;
    if (!geneva::planAndCompile(text, single_v != 0, e->graph, cq, &err)) {
        r.status = rc = GV_ERR_ARG;
        r.error = err;
    } else {
        rc = e->plain->mine(cq, geneva::PlainMiner::Shard{start_begin, start_end, stride}, r);
    }
    if (rc != GV_OK) {
        setErr(r.error.empty() ? std::string(gv_status_string(rc)) : r.error);
        return rc;
    }
    e->lastFinished = r.finished;
    e->lastFraction = r.fraction;
    if (count) *count = r.count;
    if (alg_bytes) *alg_bytes = r.algBytes;
    if (ms) *ms = r.ms;
    if (launches) *launches = r.launches;
    if (bits && nwords) {
        for (uint32_t i = 0; i < nwords; ++i) bits[i] = i < r.graphBits.size() ? r.graphBits[i] : 0;
    }
    return GV_OK;
}

int gh_engine_time_limit(gh_engine_t e, double seconds) {
    if (!e || seconds < 0) return GV_ERR_ARG;
    e->plain->setTimeLimit(seconds);
    return GV_OK;
}

int gh_engine_last_progress(gh_engine_t e, int* finished, double* fraction) {
    if (!e) return GV_ERR_ARG;
    if (finished) *finished = e->lastFinished ? 1 : 0;
    if (fraction) *fraction = e->lastFraction;
    return GV_OK;
}

int gh_engine_count_by_rows(gh_engine_t e, int on) {
    if (!e) return GV_ERR_ARG;
    e->plain->setCountByRows(on != 0);
    return GV_OK;
}

int gh_engine_batch_config(gh_engine_t e, uint32_t slots, uint64_t slot_bytes, uint64_t row_cap,
                           uint32_t ipt, uint32_t bpc) {
    if (!e) return GV_ERR_ARG;
    e->batch.reset();
    if (slots) e->batchOpt.slots = slots;
    if (slot_bytes) e->batchOpt.slotBytes = slot_bytes;
    if (row_cap) e->batchOpt.rowCap = row_cap;
    if (ipt) e->batchOpt.tune.ipt = ipt;
    if (bpc) e->batchOpt.tune.walk_blocks_per_cu = bpc;
    return GV_OK;
}

int gh_engine_batch_mine(gh_engine_t e, const char* const* queries, uint32_t n, int single_v,
                         uint64_t* counts, uint64_t* alg_bytes, uint32_t* bits, uint32_t nwords,
                         double* stats) {
    if (!e || (n && !queries)) return GV_ERR_ARG;
    std::vector<geneva::CompiledQuery> qs;
    const auto c0 = std::chrono::steady_clock::now();
    int rc = compileAll(e, queries, n, single_v, qs);
    if (rc != GV_OK) return rc;
    const double compileMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
    rc = ensureBatch(e);
    if (rc != GV_OK) {
        setErr(std::string("batch engine: ") + gv_status_string(rc) + " " + gv_last_error());
        return rc;
    }
    std::vector<geneva::MineResult> res;
    geneva::BatchStats st;
    rc = e->batch->run(qs, res, &st);
    if (rc != GV_OK) {
        setErr(std::string("batch run: ") + gv_status_string(rc) + " " + gv_last_error());
        return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (res[i].status != GV_OK) {
            setErr("query " + std::to_string(i) + " failed");
            return res[i].status;
        }
        if (counts) counts[i] = res[i].count;
        if (alg_bytes) alg_bytes[i] = res[i].algBytes;
        if (bits && nwords) {
            uint32_t* dst = bits + (uint64_t)i * nwords;
            const size_t m = std::min<size_t>(res[i].graphBits.size(), nwords);
            if (m) std::memcpy(dst, res[i].graphBits.data(), m * sizeof(uint32_t));
            if (m < nwords) std::memset(dst + m, 0, (nwords - m) * sizeof(uint32_t));
        }
    }
    if (stats) {
        stats[0] = (double)st.batched;
        stats[1] = (double)st.overflowed;
        stats[2] = (double)st.graphLaunches;
        stats[3] = st.deviceMs;
        stats[4] = st.wallMs;
        stats[5] = compileMs;
        stats[6] = st.fallbackMs;
        stats[7] = (double)st.pieces;
    }
    return GV_OK;
}

int gh_engine_autotune(gh_engine_t e, const char* const* queries, uint32_t n, uint32_t* ipt,
                       uint32_t* bpc, double* best_ms) {
    if (!e || (n && !queries)) return GV_ERR_ARG;
    std::vector<geneva::CompiledQuery> qs;
    int rc = compileAll(e, queries, n, 0, qs);
    if (rc != GV_OK) return rc;
    rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    geneva::TuneResult tr;
    rc = geneva::autotune(*e->batch, qs, tr);
    if (rc != GV_OK) return rc;
    e->batchOpt.tune = tr.tune;
    e->tuneTable.clear();
    for (auto& kv : tr.classes)
        e->tuneTable += kv.first + ":" + std::to_string(kv.second.first) + ":" +
                        std::to_string(kv.second.second) + ";";
    if (ipt) *ipt = tr.tune.ipt;
    if (bpc) *bpc = tr.tune.walk_blocks_per_cu;
    if (best_ms) *best_ms = tr.bestMs;
    return GV_OK;
}

int gh_engine_batch_row_cap(gh_engine_t e, uint64_t* rows) {
    if (!e || !rows) return GV_ERR_ARG;
    *rows = e->batch ? e->batch->rowCap() : e->batchOpt.rowCap;
    return GV_OK;
}

uint64_t gh_engine_autotune_classes(gh_engine_t e, char* out, uint64_t cap) {
    if (!e) return 0;
    const std::string& t = e->tuneTable;
    if (out && cap) {
        const uint64_t n = std::min<uint64_t>(t.size(), cap - 1);
        memcpy(out, t.data(), n);
        out[n] = 0;
    }
    return t.size();
}

int gh_engine_serve(gh_engine_t e, const char* const* queries, const double* arrival_s, uint32_t n,
                    int mode, uint64_t mem_budget, double* latency_ms, uint64_t* counts,
                    uint32_t* graph_bits, uint32_t nwords, double* stats) {
    if (!e || !arrival_s || !latency_ms || (n && !queries)) return GV_ERR_ARG;
    for (uint32_t i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) return GV_ERR_ARG;
    std::vector<geneva::CompiledQuery> qs;
    int rc = compileAll(e, queries, n, 0, qs);
    if (rc != GV_OK) return rc;
    std::vector<double> arr(arrival_s, arrival_s + n);
    geneva::ServeResult sr;
    if (mode == 0) {
        rc = ensureBatch(e);
        if (rc != GV_OK) return rc;
        if (!e->stats) e->stats.reset(new geneva::GraphStats(geneva::GraphStats::of(e->graph)));
        if (!e->side) {
            rc = gv_ctx_create(e->device, &e->sideCtx);
            if (rc != GV_OK) return rc;
            e->side.reset(new geneva::PlainMiner(e->sideCtx, e->graph, e->levelBytes));
        }
        rc = geneva::serveBatched(*e->batch, qs, arr, mem_budget, *e->stats, sr, e->side.get());
    } else {
        rc = geneva::servePlain(*e->plain, qs, arr, sr);
    }
    if (rc != GV_OK) {
        setErr(std::string("serve: ") + gv_status_string(rc) + " " + gv_last_error());
        return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        latency_ms[i] = sr.latencyMs[i];
        if (counts) counts[i] = sr.counts[i];
        if (graph_bits && nwords) {
            uint32_t* dst = graph_bits + (size_t)i * nwords;
            for (uint32_t w = 0; w < nwords; ++w)
                dst[w] = w < sr.nwords ? sr.graphBits[(size_t)i * sr.nwords + w] : 0;
        }
    }
    if (stats) {
        stats[0] = sr.wallS;
        stats[1] = (double)sr.iterations;
        stats[2] = (double)sr.maxConcurrent;
        stats[3] = (double)sr.overflowed;
        stats[4] = (double)sr.direct;
        stats[5] = (double)sr.maxBatched;
    }
    return GV_OK;
}

int gh_engine_batch_profile(gh_engine_t e, int on) {
    if (!e) return GV_ERR_ARG;
    int rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    return gv_batch_profile(e->batch->handle(), on);
}

int gh_engine_batch_profile_bytes(gh_engine_t e, uint64_t* size_bytes, uint64_t* walk_bytes) {
    if (!e || !e->batch) return GV_ERR_ARG;
    return gv_batch_profile_bytes(e->batch->handle(), size_bytes, walk_bytes);
}

int gh_engine_batch_pair_start(gh_engine_t e, int on) {
    if (!e) return GV_ERR_ARG;
    int rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    return gv_batch_pair_start(e->batch->handle(), on);
}

int gh_engine_batch_share(gh_engine_t e, int on) {
    if (!e) return GV_ERR_ARG;
    int rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    return gv_batch_share(e->batch->handle(), on);
}

int gh_engine_batch_share_stats(gh_engine_t e, uint64_t* shared, uint64_t* restarted) {
    if (!e) return GV_ERR_ARG;
    int rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    return gv_batch_share_stats(e->batch->handle(), shared, restarted);
}

int gh_engine_batch_fuse_sizing(gh_engine_t e, int on) {
    if (!e) return GV_ERR_ARG;
    int rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    return gv_batch_fuse_sizing(e->batch->handle(), on);
}

int gh_engine_batch_fused_rows(gh_engine_t e, uint64_t* rows) {
    if (!e) return GV_ERR_ARG;
    int rc = ensureBatch(e);
    if (rc != GV_OK) return rc;
    return gv_batch_fused_rows(e->batch->handle(), rows);
}

int gh_engine_batch_history(gh_engine_t e, uint64_t* rows, uint64_t* candidates, uint32_t cap, uint32_t* n) {
    if (!e || !e->batch) return GV_ERR_ARG;
    return gv_batch_history(e->batch->handle(), rows, candidates, cap, n);
}

int gh_engine_profile_records(gh_engine_t e, int* kinds, float* ms, uint64_t cap, uint64_t* n) {
    return e ? gv_profile_records(e->ctx, kinds, ms, cap, n) : GV_ERR_ARG;
}

int gh_engine_profile(gh_engine_t e, int on) {
    return e ? gv_profile_enable(e->ctx, on) : GV_ERR_ARG;
}

int gh_engine_profile_reset(gh_engine_t e) { return e ? gv_profile_reset(e->ctx) : GV_ERR_ARG; }

int gh_engine_profile_read(gh_engine_t e, int kind, uint64_t* launches, double* total_ms,
                           uint64_t* req_bytes) {
    return e ? gv_profile_read(e->ctx, kind, launches, total_ms, req_bytes) : GV_ERR_ARG;
}

int gh_engine_profile_units(gh_engine_t e, int kind, uint64_t* units) {
    return e ? gv_profile_units(e->ctx, kind, units) : GV_ERR_ARG;
}

int gh_engine_profile_alg_bytes(gh_engine_t e, int kind, uint64_t* bytes) {
    return e ? gv_profile_alg_bytes(e->ctx, kind, bytes) : GV_ERR_ARG;
}

struct gh_phase_run {
    geneva::PhaseRunResult r;
};

int gh_phase_mine(const char* mygraph, const char* metadata, const char* query_file, int single_v, int device,
                  uint64_t arena_bytes, uint32_t max_rows, int keep_rows, gh_phase_run_t* out) {
    if (!mygraph || !metadata || !query_file || !out) return GV_ERR_ARG;
    *out = nullptr;
    gv_ctx_t ctx = nullptr;
    int rc = gv_ctx_create(device, &ctx);
    if (rc != GV_OK) {
        setErr(std::string("gv_ctx_create: ") + gv_status_string(rc));
        return rc;
    }
    std::unique_ptr<gh_phase_run> run(new gh_phase_run());
    geneva::PhaseDriverOptions opt;
    if (arena_bytes) opt.arenaBytes = arena_bytes;
    opt.maxRows = max_rows;
    opt.keepRows = keep_rows != 0;
    rc = geneva::runPhaseDriver(ctx, mygraph, metadata, query_file, single_v < 0 ? -1 : (single_v ? 1 : 0),
                                opt, run->r);
    gv_ctx_destroy(ctx);
    if (rc != GV_OK) {
        setErr("phase driver: " + std::string(gv_status_string(rc)) + " " + run->r.error + " " + gv_last_error());
        return rc;
    }
    *out = run.release();
    return GV_OK;
}

uint64_t gh_phase_count(gh_phase_run_t r) { return r ? r->r.count : 0; }
uint32_t gh_phase_width(gh_phase_run_t r) { return r ? r->r.width : 0; }
const uint32_t* gh_phase_order(gh_phase_run_t r) { return r ? r->r.order.data() : nullptr; }
const uint32_t* gh_phase_rows(gh_phase_run_t r) { return r && !r->r.rows.empty() ? r->r.rows.data() : nullptr; }
uint64_t gh_phase_launches(gh_phase_run_t r) { return r ? r->r.launches : 0; }
uint64_t gh_phase_spills(gh_phase_run_t r) { return r ? r->r.spills : 0; }
int gh_phase_single_v(gh_phase_run_t r) { return r ? (r->r.singleV ? 1 : 0) : -1; }
void gh_phase_free(gh_phase_run_t r) { delete r; }

}  // extern "C"
