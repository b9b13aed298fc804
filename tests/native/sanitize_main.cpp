// Host-side sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer) of
// the product planner, loader, format writers and renumbering, and of the
// oracle (its planner and its three mining modes).  Built and run by
// tests/test_cpu_sanitize.py with g++ -fsanitize=address,undefined; any
// report aborts the run.  Test infrastructure only.
//
//   sanitize_main <queries file> <Enron.mygraph> <Enron.metadata>
//
// The queries file holds .query texts separated by lines containing "%%".
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../geneva_amd/csrc/host/genMatchOrder.h"
#include "../../geneva_amd/csrc/host/loadDataBase.h"
#include "../../oracle/oracle_planner.h"

extern "C" {
void* oracle_graph_create(uint32_t nv, const uint32_t* vlabel, uint64_t m, const uint32_t* edges);
void oracle_graph_destroy(void* h);
int oracle_mine(void* h, const char* query, int mode, int threads, uint64_t startBegin, uint64_t startEnd,
                uint64_t stride, const uint32_t* graphOf, uint32_t ngraphs, uint8_t* bits, uint64_t* countOut,
                uint64_t* bytesOut, uint64_t* nstartOut, double* secs);
}

static std::vector<std::string> readQueries(const char* path) {
    std::ifstream f(path);
    std::vector<std::string> qs;
    std::string line, cur;
    while (std::getline(f, line)) {
        if (line == "%%") {
            if (!cur.empty()) qs.push_back(cur);
            cur.clear();
        } else {
            cur += line + "\n";
        }
    }
    if (!cur.empty()) qs.push_back(cur);
    return qs;
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const std::vector<std::string> qs = readQueries(argv[1]);
    // planners: the product's DV and SV planners and the oracle's
    size_t planned = 0;
    for (const std::string& q : qs) {
        std::vector<Phase> dv;
        std::vector<PhaseSingleV> sv;
        std::map<uint64_t, uint> rf;
        std::map<uint, uint> rs, ql;
        std::map<uint, std::map<uint, bool>> ls;
        try {
            std::istringstream a(q), b(q);
            geneva::planDV(a, true, dv, rf, rs, ls, ql, nullptr, nullptr, nullptr, nullptr);
            rf.clear(), rs.clear(), ql.clear(), ls.clear();
            geneva::planSV(b, false, sv, rf, rs, ls, ql, nullptr, nullptr, nullptr, nullptr);
            (void)calDVPhaseCount(dv);
        } catch (const geneva::PlanError&) {
        }
        oracle::OPlan p;
        std::string err;
        if (oracle::planQuery(q, p, err)) (void)oracle::dumpPlan(p);
        ++planned;
    }
    // loader, decoder and writers on the Enron fixture
    std::map<uint, uint*> db = loadDataBase(argv[2], argv[3]);
    if (db.find(0xffffffffu) == db.end()) return 3;
    geneva::DataGraph g;
    const uint* vLabel = db[0xffffffffu];
    g.nv = vLabel[0];
    g.vlabel.assign(vLabel, vLabel + g.nv + 1);
    g.vlabel[0] = 0;
    for (auto& kv : db) {
        if (kv.first == 0xffffffffu) continue;
        struct Ctx {
            geneva::DataGraph* g;
            uint label;
        } c{&g, kv.first};
        std::string err;
        geneva::forEachLabelGroup(
            kv.second, vLabel, g.nv,
            [](void* u, uint src, uint, const uint* nb, uint n) {
                Ctx* c = (Ctx*)u;
                for (uint i = 0; i < n; ++i)
                    if (src < nb[i]) c->g->edges.push_back({src, nb[i], c->label});
            },
            &c, &err);
    }
    geneva::freeDataBase(db);
    std::string err;
    std::map<uint, uint*> db2 = geneva::buildDataBase(g, nullptr, &err);
    geneva::freeDataBase(db2);
    std::vector<uint> newOf;
    (void)geneva::reorderForIntervals(g, newOf, &err);
    geneva::DataGraph h = geneva::relabel(g, newOf);
    (void)geneva::partitionRuns(h);
    // the oracle's modes on a small random labelled graph
    std::mt19937 rng(7);
    const uint32_t nv = 60;
    std::vector<uint32_t> vl(nv + 1, 0), edges;
    for (uint32_t v = 1; v <= nv; ++v) vl[v] = 1 + rng() % 3;
    for (int i = 0; i < 180; ++i) {
        edges.push_back(1 + rng() % nv);
        edges.push_back(1 + rng() % nv);
        edges.push_back(1 + rng() % 2);
    }
    void* og = oracle_graph_create(nv, vl.data(), edges.size() / 3, edges.data());
    size_t mined = 0;
    for (size_t i = 0; i < qs.size() && mined < 40; ++i) {
        uint64_t c[3] = {0, 0, 0};
        bool ok = true;
        for (int mode = 0; mode < 3 && ok; ++mode)
            ok = oracle_mine(og, qs[i].c_str(), mode, 1, 0, 0, 1, nullptr, 0, nullptr, &c[mode], nullptr,
                             nullptr, nullptr) == 0;
        if (!ok) continue;
        if (c[0] != c[1] || c[1] != c[2]) {
            std::fprintf(stderr, "oracle modes disagree on query %zu\n", i);
            return 4;
        }
        ++mined;
    }
    oracle_graph_destroy(og);
    std::printf("sanitize ok: %zu plans, %zu mined, %zu edges\n", planned, mined, g.edges.size());
    return 0;
}
