// Serving loop: continuous batching over the batch engine, and the plain
// one-query-at-a-time server it is compared with.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "batch.h"

namespace geneva {

namespace {
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}
void waitUntil(Clock::time_point t0, double s) {
    for (;;) {
        const double d = s - since(t0);
        if (d <= 0) return;
        if (d > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(d - 1e-4));
    }
}

// Plain-path requests of the batched server, mined in submission order by one
// worker thread on its own PlainMiner (own context and stream), so that the
// serving loop keeps admitting and iterating while they run.
class SideWorker {
public:
    struct Done {
        size_t req;
        int status;
        uint64_t count;
        std::vector<uint32_t> bits;
        double finishS;
    };
    SideWorker(PlainMiner* pm, const std::vector<CompiledQuery>& reqs, Clock::time_point t0)
        : pm_(pm), reqs_(reqs), t0_(t0) {
        if (pm_) th_ = std::thread([this] { loop(); });
    }
    ~SideWorker() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }
    bool enabled() const { return pm_ != nullptr; }
    void submit(size_t r) {
        {
            std::lock_guard<std::mutex> l(m_);
            q_.push_back(r);
            ++pending_;
        }
        cv_.notify_all();
    }
    // requests submitted and not yet collected
    uint32_t pending() const { return pending_.load(); }
    // results waiting to be collected
    size_t doneCount() {
        std::lock_guard<std::mutex> l(m_);
        return done_.size();
    }
    bool idle() const { return pending_.load() == 0; }
    void collect(std::vector<Done>& out) {
        std::lock_guard<std::mutex> l(m_);
        for (auto& d : done_) out.push_back(std::move(d));
        pending_ -= (uint32_t)done_.size();
        done_.clear();
    }

private:
    void loop() {
        for (;;) {
            size_t r;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                r = q_.front();
                q_.pop_front();
            }
            MineResult mr;
            int rc = pm_->mine(reqs_[r], mr);
            Done d{r, rc, mr.count, std::move(mr.graphBits), since(t0_)};
            std::lock_guard<std::mutex> l(m_);
            done_.push_back(std::move(d));
        }
    }
    PlainMiner* pm_;
    const std::vector<CompiledQuery>& reqs_;
    Clock::time_point t0_;
    std::thread th_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<size_t> q_;
    std::vector<Done> done_;
    std::atomic<uint32_t> pending_{0};
    bool stop_ = false;
};
}  // namespace

// A request whose estimated largest table exceeds a slot table's share
// 1 / kDirectShare goes straight to the plain path.  The estimate (from label
// statistics) runs low on the heaviest walk queries, which then overflow their
// slot after several iterations and are mined again on the plain path; on
// workload 5 at 10,000 offered q/s a share of 1/4 instead of the whole table
// cut such overflows from 75 to 21 in 15,000 requests and raised the batched
// server's rate by about 1.5 % (1/3 leaves 53 overflows and less of the
// gain; 1/6, 1/8 and 1/16 send so many requests to the one side worker that
// it becomes the bottleneck).
constexpr uint64_t kDirectShare = 4;

int serveBatched(BatchMiner& bm, const std::vector<CompiledQuery>& reqs,
                 const std::vector<double>& arrivalS, uint64_t memBudget, const GraphStats& gs,
                 ServeResult& out, PlainMiner* side) {
    int rc = bm.init();
    if (rc != GV_OK) return rc;
    const size_t n = reqs.size();
    out = ServeResult();
    out.latencyMs.assign(n, 0);
    out.counts.assign(n, 0);
    gv_batch_t b = bm.handle();
    const uint32_t slots = bm.options().slots;
    const uint64_t slotRowsBytes = bm.options().slotBytes;
    const uint32_t nwords = bm.graphWords();   // matched-graph bitset width
    out.nwords = nwords;
    out.graphBits.assign(n * (size_t)nwords, 0);
    std::vector<uint64_t> est(n);
    for (size_t i = 0; i < n; ++i) est[i] = estimatePeakBytes(reqs[i], gs);
    std::vector<size_t> slotReq(slots, SIZE_MAX);
    std::vector<gv_slot_info_t> info(slots);
    size_t overflowed = 0;
    const Clock::time_point t0 = Clock::now();
    SideWorker worker(side, reqs, t0);
    size_t finished = 0;
    auto finishPlain = [&](size_t r, uint64_t count, const std::vector<uint32_t>& bits, double atS) {
        out.counts[r] = count;
        out.latencyMs[r] = (atS - arrivalS[r]) * 1e3;
        std::copy(bits.begin(), bits.begin() + std::min<size_t>(bits.size(), nwords),
                  out.graphBits.begin() + r * (size_t)nwords);
        ++finished;
    };
    // a request that cannot stay on the batch engine goes to the plain path:
    // to the side worker, or inline (the loop waits for it)
    auto plain = [&](size_t r, bool direct) -> int {
        if (direct) ++out.direct;
        else ++overflowed;
        if (worker.enabled()) {
            worker.submit(r);
            return GV_OK;
        }
        PlainMiner* pm = bm.fallback();
        if (!pm) return GV_ERR_CAPACITY;
        MineResult mr;
        int rc2 = pm->mine(reqs[r], mr);
        if (rc2 != GV_OK) return rc2;
        finishPlain(r, mr.count, mr.graphBits, since(t0));
        return GV_OK;
    };
    std::vector<SideWorker::Done> sideDone;
    auto collectSide = [&]() -> int {
        sideDone.clear();
        worker.collect(sideDone);
        for (auto& d : sideDone) {
            if (d.status != GV_OK) return d.status;
            finishPlain(d.req, d.count, d.bits, d.finishS);
        }
        return GV_OK;
    };
    uint64_t inUse = 0;
    uint32_t running = 0;
    size_t next = 0;
    std::vector<uint32_t> doneSlots, overSlots, bits;
    while (finished < n) {
        rc = collectSide();
        if (rc != GV_OK) return rc;
        if (finished >= n) break;
        const double now = since(t0);
        // an idle engine and a lone arrived request: the plain path answers it
        // in one launch sequence instead of one iteration per vertex
        if (running == 0 && next < n && arrivalS[next] <= now && !reqs[next].empty &&
            (next + 1 >= n || arrivalS[next + 1] > now) && worker.idle() &&
            (worker.enabled() || bm.fallback())) {
            rc = plain(next, true);
            if (rc != GV_OK) return rc;
            ++next;
            continue;
        }
        // admit what has arrived, in arrival order, while a slot is free and
        // the memory bound holds (a lone request is always admitted; one whose
        // estimate exceeds a slot table's share goes straight to the plain
        // path)
        while (next < n && arrivalS[next] <= now && running < slots) {
            if (reqs[next].empty) {
                out.latencyMs[next] = (since(t0) - arrivalS[next]) * 1e3;
                ++finished;
                ++next;
                continue;
            }
            if (est[next] > slotRowsBytes / kDirectShare) {
                rc = plain(next, true);
                if (rc != GV_OK) return rc;
                ++next;
                continue;
            }
            if (running > 0 && inUse + est[next] > memBudget) break;
            uint32_t slot = 0;
            rc = bm.admit(reqs[next], next, &slot);
            if (rc == GV_ERR_CAPACITY) break;
            if (rc != GV_OK) return rc;
            slotReq[slot] = next;
            inUse += est[next];
            ++running;
            ++next;
        }
        out.maxConcurrent = std::max<uint64_t>(out.maxConcurrent, running + worker.pending());
        out.maxBatched = std::max<uint64_t>(out.maxBatched, running);
        if (running == 0) {
            // nothing on the batch engine: wait for the next arrival or for the
            // side worker, whichever comes first
            if (next < n) {
                // poll the worker until it has a result or the arrival is due
                // (results are collected at the top of the loop)
                const double until = arrivalS[next];
                const uint32_t before = worker.pending();
                while (before > 0 && worker.doneCount() == 0 && since(t0) < until)
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                if (before == 0) waitUntil(t0, until);
            } else if (worker.pending() > 0) {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            } else {
                break;
            }
            continue;
        }
        rc = gv_batch_run(b, 1);
        if (rc != GV_OK) return rc;
        ++out.iterations;
        rc = gv_batch_poll(b, info.data());
        if (rc != GV_OK) return rc;
        doneSlots.clear();
        overSlots.clear();
        for (uint32_t s = 0; s < slots; ++s) {
            if (slotReq[s] == SIZE_MAX) continue;
            if (info[s].state == GV_SLOT_DONE) doneSlots.push_back(s);
            else if (info[s].state == GV_SLOT_OVERFLOW) overSlots.push_back(s);
        }
        if (!doneSlots.empty()) {
            bits.assign(doneSlots.size() * (size_t)nwords, 0);
            rc = gv_batch_release_many(b, (uint32_t)doneSlots.size(), doneSlots.data(),
                                       nwords ? bits.data() : nullptr, nwords);
            if (rc != GV_OK) return rc;
            const double t = since(t0);
            for (size_t i = 0; i < doneSlots.size(); ++i) {
                const uint32_t s = doneSlots[i];
                const size_t r = slotReq[s];
                out.counts[r] = info[s].count;
                out.latencyMs[r] = (t - arrivalS[r]) * 1e3;
                std::copy(bits.begin() + i * nwords, bits.begin() + (i + 1) * nwords,
                          out.graphBits.begin() + r * (size_t)nwords);
                slotReq[s] = SIZE_MAX;
                inUse -= est[r];
                --running;
                ++finished;
            }
        }
        if (!overSlots.empty()) {
            rc = gv_batch_release_many(b, (uint32_t)overSlots.size(), overSlots.data(), nullptr, 0);
            if (rc != GV_OK) return rc;
            for (uint32_t s : overSlots) {
                const size_t r = slotReq[s];
                slotReq[s] = SIZE_MAX;
                inUse -= est[r];
                --running;
                rc = plain(r, false);
                if (rc != GV_OK) return rc;
            }
        }
    }
    out.overflowed = overflowed;
    out.wallS = since(t0) - (n ? arrivalS[0] : 0);
    return GV_OK;
}

int servePlain(PlainMiner& pm, const std::vector<CompiledQuery>& reqs,
               const std::vector<double>& arrivalS, ServeResult& out) {
    const size_t n = reqs.size();
    out = ServeResult();
    out.latencyMs.assign(n, 0);
    out.counts.assign(n, 0);
    const Clock::time_point t0 = Clock::now();
    for (size_t i = 0; i < n; ++i) {
        waitUntil(t0, arrivalS[i]);
        MineResult r;
        int rc = pm.mine(reqs[i], r);
        if (rc != GV_OK) return rc;
        if (out.nwords == 0 && !r.graphBits.empty()) {
            out.nwords = (uint32_t)r.graphBits.size();
            out.graphBits.assign(n * (size_t)out.nwords, 0);
        }
        out.counts[i] = r.count;
        if (out.nwords)
            std::copy(r.graphBits.begin(), r.graphBits.begin() + std::min<size_t>(r.graphBits.size(), out.nwords),
                      out.graphBits.begin() + i * (size_t)out.nwords);
        out.latencyMs[i] = (since(t0) - arrivalS[i]) * 1e3;
    }
    out.maxConcurrent = 1;
    out.wallS = since(t0) - (n ? arrivalS[0] : 0);
    return GV_OK;
}

}  // namespace geneva
