// The reference's phase-by-phase host driver (src/gpumine.cpp:866-1072 for DV
// plans, src/gpumine_singleV.cpp:707-928 for SV plans), restated over the
// reference-named launchers of include/geneva_launchers.hpp.
//
// It is written against the reference's interfaces only: loadDataBase for the
// files, genMatchOrder / genMatchOrderSingleV for the plan (whose upload hook
// puts the first partition block on the device, as the reference planner
// does), one device arena holding vLabel, the two counters and the tables,
// and the launchers initEmb_2V_NoHash ... reductionPhase / copyPartialEmb
// with the reference's arguments, including its auxArray encodings, its
// maxRowNum bound and the overflow protocol (counters read back after every
// launch; rows that do not fit are moved to host memory and the launch is
// repeated with isContinue).  A table that was moved to host memory in parts
// is processed part by part in the next phase.  The result is every final
// row, spilled parts included (the reference counts only the rows left on the
// device, src/gpumine.cpp:1068).
#include "phase_driver.h"

#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "../../../include/geneva_launchers.hpp"
#include "genMatchOrder.h"
#include "loadDataBase.h"

namespace geneva {

namespace {

thread_local gv_ctx_t t_uploadCtx = nullptr;
void uploadOnCtx(void* dst, const void* src, uint64_t n) {
    if (t_uploadCtx) gv_memcpy_h2d(t_uploadCtx, dst, src, n);
}

// A table of rows held in host memory after an overflow.
struct HostPart {
    std::vector<uint32_t> rows;
    uint64_t n = 0;
};

// the reference planner of each plan kind (file form, device upload through
// the planner's hook)
void planFile(const std::string& f, std::vector<Phase>& mo, std::map<uint64_t, uint>& rf,
              std::map<uint, uint>& rs, std::map<uint, std::map<uint, bool>>& ls, std::map<uint, uint>& ql,
              std::map<uint, uint*>& db, uint32_t* base, uint& stride) {
    genMatchOrder(f, mo, rf, rs, ls, ql, db, base, stride);
}
void planFile(const std::string& f, std::vector<PhaseSingleV>& mo, std::map<uint64_t, uint>& rf,
              std::map<uint, uint>& rs, std::map<uint, std::map<uint, bool>>& ls, std::map<uint, uint>& ql,
              std::map<uint, uint*>& db, uint32_t* base, uint& stride) {
    genMatchOrderSingleV(f, mo, rf, rs, ls, ql, db, base, stride);
}

template <class PhaseT>
class Driver {
public:
    Driver(gv_ctx_t ctx, std::map<uint, uint*>& db, const PhaseDriverOptions& opt)
        : ctx_(ctx), db_(db), opt_(opt) {}

    int run(const std::string& queryFile, bool singleV, PhaseRunResult& out);

private:
    int check() {
        const int e = geneva_launchers::take_error();
        if (e != GV_OK) err_ = e;
        return err_;
    }
    int readCounters() {
        int rc = gv_memcpy_d2h(ctx_, cnt_, counter_, 8);
        if (rc != GV_OK) return rc;
        ++out_->launches;
        return check();
    }
    int resetCounters() {
        cnt_[0] = cnt_[1] = 0;
        return gv_memcpy_h2d(ctx_, counter_, cnt_, 8);
    }
    // rows [0, n) of the table at dev (width w) to a new host part
    int spill(const uint32_t* dev, uint64_t n, uint32_t w, std::vector<HostPart>& to) {
        HostPart p;
        p.n = n;
        p.rows.resize(n * w);
        if (n) {
            int rc = gv_memcpy_d2h(ctx_, p.rows.data(), dev, 4 * n * w);
            if (rc != GV_OK) return rc;
        }
        to.push_back(std::move(p));
        ++out_->spills;
        return GV_OK;
    }
    uint32_t boundRows(uint64_t r) const {
        if (opt_.maxRows && r > opt_.maxRows) r = opt_.maxRows;
        return (uint32_t)std::min<uint64_t>(r, 0xffffffffull);
    }
    // the labels of the new columns and the same-label bookkeeping of the
    // reference driver (src/gpumine.cpp:508-549 and :367-387)
    void addCol(uint32_t qv) {
        const uint32_t col = (uint32_t)pos_.size();
        pos_[qv] = col;
        order_.push_back(qv);
    }
    int launchInit(const PhaseT& ph, uint32_t* newEmb, bool cont, uint32_t maxRow, uint32_t* rec1,
                   uint32_t* rec2);
    int launchExt(const PhaseT& ph, uint32_t* newEmb, uint32_t* partial, uint32_t rows, uint32_t embLen,
                  uint32_t* aux, bool last, bool cont, uint32_t maxRow, uint32_t* rec1, uint32_t* rec2);
    int uploadPartition(uint edgeLabel);

    gv_ctx_t ctx_;
    std::map<uint, uint*>& db_;
    PhaseDriverOptions opt_;
    PhaseRunResult* out_ = nullptr;
    int err_ = GV_OK;
    uint32_t* arena_ = nullptr;
    uint32_t* vlab_ = nullptr;
    uint32_t* counter_ = nullptr;
    uint32_t* base_ = nullptr;
    uint64_t avail_ = 0;           // uints from base_
    uint32_t* cnt_ = nullptr;      // host copy of the counters
    uint32_t cntStore_[4] = {0, 0, 0, 0};
    uint32_t* part_ = nullptr;     // partition block on the device (array0)
    uint32_t partLen_ = 0, intervals_ = 0, totSrc_ = 0, a3off_ = 0;
    uint edgeLabel_ = 0;
    std::map<uint, uint> pos_;                  // query vertex -> column
    std::vector<uint32_t> order_;
    std::map<uint, std::vector<uint32_t>> equalPos_;   // vertex label -> columns
    std::map<uint64_t, uint> repeatFlag_;
    std::map<uint, uint> restricts_;
    std::map<uint, std::map<uint, bool>> labelStat_;
    std::map<uint, uint> queryLabel_;
};

template <class PhaseT>
int Driver<PhaseT>::uploadPartition(uint edgeLabel) {
    auto it = db_.find(edgeLabel);
    if (it == db_.end()) return GV_ERR_ARG;
    const uint32_t* blk = it->second;
    partLen_ = blk[0] - 5;
    totSrc_ = blk[1];
    intervals_ = blk[2];
    a3off_ = blk[4];
    edgeLabel_ = edgeLabel;
    part_ = base_;
    return gv_memcpy_h2d(ctx_, part_, blk + 5, 4ull * partLen_);
}

template <class PhaseT>
int Driver<PhaseT>::launchInit(const PhaseT& ph, uint32_t* newEmb, bool cont, uint32_t maxRow, uint32_t* rec1,
                               uint32_t* rec2) {
    const auto& pr = ph.pairs;
    const uint32_t svid = pr[0];
    if (pr.size() == 2) {
        const uint32_t evid = pr[1];
        initEmb_2V_NoHash(vlab_, counter_, part_, part_ + a3off_, rec1, newEmb, intervals_, totSrc_,
                          queryLabel_[svid], queryLabel_[evid], restricts_[evid] == svid, rec1 != nullptr, cont,
                          maxRow);
    } else {
        const uint32_t e1 = pr[1], e2 = pr[3];
        const bool r12 = restricts_[e2] != 0xffffffffu && restricts_[e2] == e1;
        const bool r1 = restricts_[e1] != 0xffffffffu;
        const bool r2 = !r12 && restricts_[e2] != 0xffffffffu;
        const bool same = queryLabel_[e1] == queryLabel_[e2];
        initEmb_3V_1and2_NoHash(vlab_, counter_, part_, part_ + a3off_, rec1, rec2, newEmb, intervals_, totSrc_,
                                queryLabel_[svid], queryLabel_[e1], queryLabel_[e2], r1, r2, r12, same,
                                rec1 != nullptr, rec2 != nullptr, cont, maxRow);
    }
    return readCounters();
}

template <class PhaseT>
int Driver<PhaseT>::launchExt(const PhaseT& ph, uint32_t* newEmb, uint32_t* partial, uint32_t rows,
                              uint32_t embLen, uint32_t* aux, bool last, bool cont, uint32_t maxRow,
                              uint32_t* rec1, uint32_t* rec2) {
    const auto& pr = ph.pairs;
    if (pr.size() == 2) {
        const uint32_t evid = pr[1];
        extEmb_1V_NoHash(vlab_, counter_, part_, part_ + a3off_, rec1, aux, newEmb, partial, intervals_, rows,
                         embLen, restricts_[evid] != 0xffffffffu, rec1 != nullptr, false, last, cont, maxRow);
    } else {
        const uint32_t s1 = pr[0], e1 = pr[1], s2 = pr[2], e2 = pr[3];
        const bool r12 = restricts_[e2] != 0xffffffffu && restricts_[e2] == e1;
        const bool r1 = restricts_[e1] != 0xffffffffu;
        const bool r2 = !r12 && restricts_[e2] != 0xffffffffu;
        const bool same = queryLabel_[e1] == queryLabel_[e2];
        const bool oneSrc = s1 == s2;
        const uint32_t* a3 = part_ + a3off_;
        uint32_t* nd = const_cast<uint32_t*>(a3);
        if (oneSrc && same)
            extEmb_2V_1src_1and2_sameLabel_NoHash(vlab_, counter_, part_, nd, rec1, aux, newEmb, partial,
                                                  intervals_, rows, embLen, r1, r2, r12, rec1 != nullptr, false,
                                                  cont, maxRow);
        else if (oneSrc)
            extEmb_2V_1src_1and2_notSameLabel_NoHash(vlab_, counter_, part_, nd, rec1, rec2, aux, newEmb,
                                                     partial, intervals_, rows, embLen, r1, r2, rec1 != nullptr,
                                                     rec2 != nullptr, false, false, cont, maxRow);
        else if (same)
            extEmb_2V_2src_1and2_sameLabel_NoHash(vlab_, counter_, part_, nd, rec1, aux, newEmb, partial,
                                                  intervals_, rows, embLen, r1, r2, r12, rec1 != nullptr, false,
                                                  cont, maxRow);
        else
            extEmb_2V_2src_1and2_notSameLabel_NoHash(vlab_, counter_, part_, nd, rec1, rec2, aux, newEmb,
                                                     partial, intervals_, rows, embLen, r1, r2, rec1 != nullptr,
                                                     rec2 != nullptr, false, false, cont, maxRow);
    }
    return readCounters();
}

template <class PhaseT>
int Driver<PhaseT>::run(const std::string& queryFile, bool singleV, PhaseRunResult& out) {
    out = PhaseRunResult();
    out_ = &out;
    cnt_ = cntStore_;
    geneva_launchers::set_stream(ctx_);
    geneva_launchers::take_error();
    auto vit = db_.find(0xffffffffu);
    if (vit == db_.end()) return GV_ERR_ARG;
    const uint32_t* vLabel = vit->second;
    const uint64_t totV = vLabel[0];
    // the arena: vLabel, the two counters, then base_ (the partition block,
    // records, auxArray, the new table growing up, the input table at the
    // top), as src/gpumine.cpp:885-911 lays it out
    const uint64_t words = opt_.arenaBytes / 4;
    if (words < totV + 3 + 4096) return GV_ERR_ARG;
    int rc = gv_malloc(ctx_, 4 * words, (void**)&arena_);
    if (rc != GV_OK) return rc;
    struct Free {
        gv_ctx_t c;
        uint32_t*& p;
        ~Free() {
            if (p) gv_free(c, p);
        }
    } freeArena{ctx_, arena_};
    vlab_ = arena_;
    counter_ = arena_ + totV + 1;
    base_ = counter_ + 2;
    avail_ = words - (totV + 3);
    rc = gv_memcpy_h2d(ctx_, vlab_, vLabel, 4 * (totV + 1));
    if (rc != GV_OK) return rc;

    std::vector<PhaseT> mo;
    uint stride = 0;
    t_uploadCtx = ctx_;
    setPlannerUpload(uploadOnCtx);
    (void)singleV;
    try {
        planFile(queryFile, mo, repeatFlag_, restricts_, labelStat_, queryLabel_, db_, base_, stride);
    } catch (const PlanError& e) {
        setPlannerUpload(nullptr);
        out.error = e.what;
        return GV_ERR_ARG;
    }
    setPlannerUpload(nullptr);
    if (mo.empty()) return GV_ERR_ARG;
    // the planner put the first phase's partition at base_ (stride = its length)
    {
        auto it = db_.find(mo[0].edgeLabel);
        if (it == db_.end()) return GV_ERR_ARG;
        const uint32_t* blk = it->second;
        partLen_ = blk[0] - 5;
        totSrc_ = blk[1];
        intervals_ = blk[2];
        a3off_ = blk[4];
        edgeLabel_ = mo[0].edgeLabel;
        part_ = base_ + stride - partLen_;
    }
    rc = resetCounters();
    if (rc != GV_OK) return rc;

    // ---- first phase (src/gpumine.cpp:226-338) ----
    const PhaseT& p0 = mo[0];
    uint32_t embLen = p0.pairs.size() == 2 ? 2 : 3;
    {
        const auto& pr = p0.pairs;
        addCol(pr[0]);
        addCol(pr[1]);
        if (pr.size() == 4) addCol(pr[3]);
        for (uint32_t c = 0; c < order_.size(); ++c) equalPos_[queryLabel_[order_[c]]].push_back(c);
    }
    uint64_t st = stride;
    auto recordOf = [&](uint64_t key, uint64_t& st2) -> uint32_t* {
        const uint f = repeatFlag_[key];
        const bool isRecord = (f & 0xffff) > 1;
        if (!isRecord) return nullptr;
        uint32_t* r = base_ + st2;
        st2 += 2ull * totSrc_;
        std::vector<uint32_t> zero(2ull * totSrc_, 0);
        gv_memcpy_h2d(ctx_, r, zero.data(), 4 * zero.size());
        return r;
    };
    uint32_t *rec1 = recordOf(((uint64_t)p0.edgeLabel << 32) | queryLabel_[p0.pairs[1]], st), *rec2 = nullptr;
    if (p0.pairs.size() == 4 && queryLabel_[p0.pairs[1]] != queryLabel_[p0.pairs[3]])
        rec2 = recordOf(((uint64_t)p0.edgeLabel << 32) | queryLabel_[p0.pairs[3]], st);
    std::vector<HostPart> parts;   // spilled parts of the current table
    {
        const uint64_t maxRow = ((avail_ - st) / 2 - 38 * 1500) / embLen;
        const uint32_t mr = boundRows(maxRow);
        uint32_t* newEmb = base_ + st;
        rc = launchInit(p0, newEmb, false, mr, rec1, rec2);
        while (rc == GV_OK && cnt_[0] >= mr) {
            rc = spill(newEmb, cnt_[0] - cnt_[1], embLen, parts);
            if (rc == GV_OK) rc = resetCounters();
            if (rc == GV_OK) rc = launchInit(p0, newEmb, true, mr, rec1, rec2);
        }
        if (rc != GV_OK) return rc;
    }
    uint64_t rows = cnt_[0];
    uint32_t* partial = base_ + avail_ - rows * embLen;
    copyPartialEmb(base_ + st, partial, (uint32_t)(rows * embLen));
    rc = check();
    if (rc == GV_OK) rc = resetCounters();
    if (rc != GV_OK) return rc;

    // ---- the other phases (src/gpumine.cpp:993-1061) ----
    for (size_t i = 1; i < mo.size(); ++i) {
        const PhaseT& ph = mo[i];
        const bool last = i + 1 == mo.size();
        if (ph.edgeLabel != edgeLabel_) {
            rc = uploadPartition(ph.edgeLabel);
            if (rc != GV_OK) return rc;
        }
        st = partLen_;
        if (ph.phaseType == 'E') {
            const auto& pr = ph.pairs;
            const uint32_t pre = embLen;
            const bool two = pr.size() == 4;
            // auxArray (src/gpumine.cpp:40-61): labels, restriction columns,
            // source columns, then the same-label columns of each new vertex
            std::vector<uint32_t> aux(8, 0);
            const uint32_t e1 = pr[1], e2 = two ? pr[3] : pr[1];
            const uint32_t l1 = queryLabel_[e1], l2 = queryLabel_[e2];
            aux[0] = l1;
            aux[1] = l2;
            aux[4] = pos_[pr[0]];
            aux[5] = two ? pos_[pr[2]] : aux[4];
            const bool r12 = two && restricts_[e2] != 0xffffffffu && restricts_[e2] == e1;
            aux[2] = restricts_[e1] != 0xffffffffu ? pos_[restricts_[e1]] : 0;
            aux[3] = two ? (!r12 && restricts_[e2] != 0xffffffffu ? pos_[restricts_[e2]] : 0) : aux[2];
            std::vector<uint32_t> c1 = equalPos_[l1];
            aux[6] = (uint32_t)c1.size();
            aux.insert(aux.end(), c1.begin(), c1.end());
            if (two && l2 != l1) {
                std::vector<uint32_t> c2 = equalPos_[l2];
                aux[7] = aux[6] + (uint32_t)c2.size();
                aux.insert(aux.end(), c2.begin(), c2.end());
            } else {
                aux[7] = aux[6];
            }
            addCol(e1);
            equalPos_[l1].push_back(pos_[e1]);
            if (two) {
                addCol(e2);
                equalPos_[l2].push_back(pos_[e2]);
            }
            embLen = (uint32_t)order_.size();
            uint32_t* r1 = recordOf(((uint64_t)ph.edgeLabel << 32) | l1, st);
            uint32_t* r2 = two && l2 != l1 ? recordOf(((uint64_t)ph.edgeLabel << 32) | l2, st) : nullptr;
            uint32_t* auxDev = base_ + st;
            aux.resize(64, 0);   // the launchers read 64 uints
            rc = gv_memcpy_h2d(ctx_, auxDev, aux.data(), 4 * aux.size());
            if (rc != GV_OK) return rc;
            st += 64;
            uint32_t* newEmb = base_ + st;
            // the input parts: the one on the device, then those on the host
            std::vector<HostPart> outParts;
            const size_t nin = 1 + parts.size();
            for (size_t k = 0; k < nin; ++k) {
                uint32_t* in = partial;
                uint64_t inRows = rows;
                if (k > 0) {
                    // spill what the last part left, then bring the next part up
                    rc = spill(newEmb, cnt_[0], embLen, outParts);
                    if (rc == GV_OK) rc = resetCounters();
                    if (rc != GV_OK) return rc;
                    HostPart& hp = parts[k - 1];
                    inRows = hp.n;
                    in = base_ + avail_ - inRows * pre - 32;
                    if (inRows) rc = gv_memcpy_h2d(ctx_, in, hp.rows.data(), 4 * hp.rows.size());
                    if (rc != GV_OK) return rc;
                }
                const uint64_t maxRow = (avail_ - st - inRows * pre - 1500 * 38) / embLen;
                const uint32_t mr = boundRows(maxRow);
                rc = launchExt(ph, newEmb, in, (uint32_t)inRows, embLen, auxDev, last, false, mr, r1, r2);
                while (rc == GV_OK && cnt_[0] >= mr) {
                    rc = spill(newEmb, cnt_[0] - cnt_[1], embLen, outParts);
                    if (rc == GV_OK) rc = resetCounters();
                    if (rc == GV_OK)
                        rc = launchExt(ph, newEmb, in, (uint32_t)inRows, embLen, auxDev, last, true, mr, r1, r2);
                }
                if (rc != GV_OK) return rc;
            }
            parts = std::move(outParts);
            rows = cnt_[0];
            uint64_t total = rows;
            for (auto& p : parts) total += p.n;
            if (total == 0) {
                rows = 0;
                parts.clear();
                break;
            }
            partial = base_ + avail_ - rows * embLen;
            copyPartialEmb(newEmb, partial, (uint32_t)(rows * embLen));
            rc = check();
            if (rc == GV_OK) rc = resetCounters();
            if (rc != GV_OK) return rc;
        } else {
            // reduction (src/gpumine.cpp:741-862): pairs sorted, packed five
            // bits a column, three pairs a word
            std::vector<uint32_t> pr(ph.pairs.begin(), ph.pairs.end());
            if (pr.size() > 32) return GV_ERR_ARG;
            std::vector<uint32_t> aux(40, 0);
            const uint32_t np = (uint32_t)pr.size() / 2;
            aux[0] = (np + 2) / 3;
            for (uint32_t p = 0; p < np; ++p) {
                const uint32_t s = pos_[pr[2 * p]], e = pos_[pr[2 * p + 1]];
                aux[1 + p / 3] |= (e | (s << 5)) << (10 * (p % 3));
                aux[7 + p] = queryLabel_[pr[2 * p + 1]];
            }
            uint32_t* auxDev = base_ + st;
            rc = gv_memcpy_h2d(ctx_, auxDev, aux.data(), 4 * aux.size());
            if (rc != GV_OK) return rc;
            st += aux.size();
            std::vector<HostPart> outParts;
            const size_t nin = 1 + parts.size();
            uint32_t* keep = partial;
            for (size_t k = 0; k < nin; ++k) {
                uint32_t* in = partial;
                uint64_t inRows = rows;
                if (k > 0) {
                    rc = spill(keep, cnt_[0], embLen, outParts);
                    if (rc == GV_OK) rc = resetCounters();
                    if (rc != GV_OK) return rc;
                    HostPart& hp = parts[k - 1];
                    inRows = hp.n;
                    in = base_ + avail_ - inRows * embLen;
                    if (inRows) rc = gv_memcpy_h2d(ctx_, in, hp.rows.data(), 4 * hp.rows.size());
                    if (rc != GV_OK) return rc;
                }
                reductionPhase(vlab_, counter_, part_, part_ + a3off_, nullptr, auxDev, in, intervals_,
                               (uint32_t)inRows, embLen, last);
                rc = readCounters();
                if (rc != GV_OK) return rc;
                keep = in;
            }
            parts = std::move(outParts);
            rows = cnt_[0];
            uint64_t total = rows;
            for (auto& p : parts) total += p.n;
            if (total == 0) {
                rows = 0;
                parts.clear();
                break;
            }
            uint32_t* top = base_ + avail_ - rows * embLen;
            if (top != keep) copyPartialEmb(keep, top, (uint32_t)(rows * embLen));
            partial = top;
            rc = check();
            if (rc == GV_OK) rc = resetCounters();
            if (rc != GV_OK) return rc;
        }
    }
    // every final row: the spilled parts, then the rows on the device
    out.width = embLen;
    out.order = order_;
    uint64_t total = rows;
    for (auto& p : parts) total += p.n;
    out.count = total;
    if (opt_.keepRows) {
        out.rows.reserve(total * embLen);
        for (auto& p : parts) out.rows.insert(out.rows.end(), p.rows.begin(), p.rows.end());
        const size_t at = out.rows.size();
        out.rows.resize(at + rows * embLen);
        if (rows) {
            rc = gv_memcpy_d2h(ctx_, out.rows.data() + at, partial, 4 * rows * embLen);
            if (rc != GV_OK) return rc;
        }
    }
    return check();
}

}  // namespace

int choosePlanKind(const std::string& queryFile, bool& singleV, std::string* err) {
    // DV plans whenever the query has a two-vertex phase: the reference's own
    // timings (scripts/ReadData.py:254-262) show the DV driver at or below the
    // SV driver for every dataset and DV-phase count from one up; with none
    // the two plans are the same phase list
    std::ifstream f(queryFile);
    if (!f) {
        if (err) *err = "cannot open " + queryFile;
        return GV_ERR_ARG;
    }
    std::vector<Phase> mo;
    std::map<uint64_t, uint> rf;
    std::map<uint, uint> rs, ql;
    std::map<uint, std::map<uint, bool>> ls;
    try {
        planDV(f, true, mo, rf, rs, ls, ql, nullptr, nullptr, nullptr, nullptr);
    } catch (const PlanError& e) {
        if (err) *err = e.what;
        return GV_ERR_ARG;
    }
    singleV = calDVPhaseCount(mo) == 0;
    return GV_OK;
}

int runPhaseDriver(gv_ctx_t ctx, const std::string& mygraph, const std::string& metadata,
                   const std::string& queryFile, int planKind, const PhaseDriverOptions& opt,
                   PhaseRunResult& out) {
    out = PhaseRunResult();
    bool singleV = planKind == 1;
    if (planKind < 0) {
        int rc0 = choosePlanKind(queryFile, singleV, &out.error);
        if (rc0 != GV_OK) return rc0;
    }
    out.singleV = singleV;
    std::map<uint, uint*> db = loadDataBase(mygraph, metadata);
    if (db.find(0xffffffffu) == db.end()) {
        freeDataBase(db);
        out.error = "loadDataBase failed";
        return GV_ERR_FORMAT;
    }
    int rc;
    if (singleV) {
        Driver<PhaseSingleV> d(ctx, db, opt);
        rc = d.run(queryFile, true, out);
    } else {
        Driver<Phase> d(ctx, db, opt);
        rc = d.run(queryFile, false, out);
    }
    freeDataBase(db);
    return rc;
}

}  // namespace geneva
