// Serving loop: continuous batching over the batch engine, and the plain
// one-query-at-a-time server it is compared with.
#include <algorithm>
#include <chrono>
#include <deque>
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
}  // namespace

int serveBatched(BatchMiner& bm, const std::vector<CompiledQuery>& reqs,
                 const std::vector<double>& arrivalS, uint64_t memBudget, const GraphStats& gs,
                 ServeResult& out) {
    int rc = bm.init();
    if (rc != GV_OK) return rc;
    const size_t n = reqs.size();
    out = ServeResult();
    out.latencyMs.assign(n, 0);
    out.counts.assign(n, 0);
    const uint32_t slots = bm.options().slots;
    const uint64_t slotRowsBytes = bm.options().slotBytes;
    std::vector<uint64_t> est(n);
    for (size_t i = 0; i < n; ++i) est[i] = estimatePeakBytes(reqs[i], gs);
    std::vector<size_t> slotReq(slots, SIZE_MAX);
    std::vector<gv_slot_info_t> info(slots);
    size_t overflowed = 0;
    const Clock::time_point t0 = Clock::now();
    // a request that cannot stay on the batch engine is mined on the plain
    // path at once (the loop waits for it, as a real server thread would)
    auto plain = [&](size_t r, bool direct) -> int {
        if (direct) ++out.direct;
        else ++overflowed;
        PlainMiner* pm = bm.fallback();
        if (!pm) return GV_ERR_CAPACITY;
        MineResult mr;
        int rc2 = pm->mine(reqs[r], mr);
        if (rc2 != GV_OK) return rc2;
        out.counts[r] = mr.count;
        out.latencyMs[r] = (since(t0) - arrivalS[r]) * 1e3;
        return GV_OK;
    };
    uint64_t inUse = 0;
    uint32_t running = 0;
    size_t next = 0, finished = 0;
    gv_batch_t b = bm.handle();
    while (finished < n) {
        const double now = since(t0);
        // an idle engine and a lone arrived request: the plain path answers it
        // in one launch sequence instead of one iteration per vertex
        if (running == 0 && next < n && arrivalS[next] <= now && !reqs[next].empty &&
            (next + 1 >= n || arrivalS[next + 1] > now) && bm.fallback()) {
            rc = plain(next, true);
            if (rc != GV_OK) return rc;
            ++finished;
            ++next;
            continue;
        }
        // admit what has arrived, in arrival order, while a slot is free and
        // the memory bound holds (a lone request is always admitted; one whose
        // estimate exceeds a slot table goes straight to the plain path)
        while (next < n && arrivalS[next] <= now && running < slots) {
            if (reqs[next].empty) {
                out.latencyMs[next] = (since(t0) - arrivalS[next]) * 1e3;
                ++finished;
                ++next;
                continue;
            }
            if (est[next] > slotRowsBytes) {
                rc = plain(next, false);
                if (rc != GV_OK) return rc;
                ++finished;
                ++next;
                continue;
            }
            if (running > 0 && inUse + est[next] > memBudget) break;
            uint32_t slot = 0;
            rc = gv_batch_admit_leaves(b, &reqs[next].q, reqs[next].hasLeaves ? &reqs[next].leaves : nullptr,
                                       next, &slot);
            if (rc == GV_ERR_CAPACITY) break;
            if (rc != GV_OK) return rc;
            slotReq[slot] = next;
            inUse += est[next];
            ++running;
            ++next;
        }
        out.maxConcurrent = std::max<uint64_t>(out.maxConcurrent, running);
        if (running == 0) {
            if (next < n) waitUntil(t0, arrivalS[next]);
            else break;
            continue;
        }
        rc = gv_batch_run(b, 1);
        if (rc != GV_OK) return rc;
        ++out.iterations;
        rc = gv_batch_poll(b, info.data());
        if (rc != GV_OK) return rc;
        const double t = since(t0);
        for (uint32_t s = 0; s < slots; ++s) {
            if (slotReq[s] == SIZE_MAX) continue;
            const size_t r = slotReq[s];
            if (info[s].state == GV_SLOT_DONE || info[s].state == GV_SLOT_OVERFLOW) {
                const bool done = info[s].state == GV_SLOT_DONE;
                if (done) {
                    out.counts[r] = info[s].count;
                    out.latencyMs[r] = (t - arrivalS[r]) * 1e3;
                }
                rc = gv_batch_release(b, s, nullptr, 0);
                if (rc != GV_OK) return rc;
                slotReq[s] = SIZE_MAX;
                inUse -= est[r];
                --running;
                if (!done) {
                    rc = plain(r, false);
                    if (rc != GV_OK) return rc;
                }
                ++finished;
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
        out.counts[i] = r.count;
        out.latencyMs[i] = (since(t0) - arrivalS[i]) * 1e3;
    }
    out.maxConcurrent = 1;
    out.wallS = since(t0) - (n ? arrivalS[0] : 0);
    return GV_OK;
}

}  // namespace geneva
