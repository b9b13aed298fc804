// ORACLE (test infrastructure only).  Restates src/genMatchOrderCPU.cpp.
#include "oracle_planner.h"

#include <algorithm>
#include <functional>
#include <list>
#include <sstream>

namespace oracle {

namespace {

struct NI {  // src/genMatchOrderCPU.cpp:41-49
    unsigned common = 0, s = 0, e = 0, el = 0, evl = 0, evn = 0;
};
struct KC {  // src/genMatchOrderCPU.cpp:50-57
    unsigned count = 0;
    std::vector<uint64_t> keys;
    std::vector<unsigned> where;
};
struct TI {  // src/genMatchOrderCPU.cpp:34-40
    unsigned idx = 0, cnt = 0, score = 0, third = 0;
};

}  // namespace

bool planQuery(const std::string& text, OPlan& P, std::string& err) {
    P = OPlan();
    std::istringstream in(text);
    std::map<unsigned, unsigned>& L = P.queryLabel;
    char c;
    // vertex section (src/genMatchOrderCPU.cpp:681-698; only '#' lines skipped)
    while (in.get(c)) {
        if (c == 'v') {
            unsigned id, lab;
            in >> id >> lab;
            if (!L.count(id)) L[id] = lab;
        } else if (c == 'e') {
            break;
        } else if (c == '#') {
            std::string skip;
            std::getline(in, skip);
        }
    }
    if (in) in.unget(); else in.clear();
    const unsigned n = (unsigned)L.size();
    P.vnum = n;
    if (n < 2 || n > 32) { err = "query size"; return false; }
    for (auto& kv : L) if (kv.first >= n) { err = "query ids"; return false; }
    std::vector<unsigned>& A = P.adj;
    A.assign(n * n, 0xffffffffu);
    std::vector<unsigned> nb(n, 0);                              // vertexNeig
    std::vector<unsigned> allNb(n, 0);                           // query[v].neigs
    std::vector<std::map<unsigned, std::pair<unsigned, unsigned>>> eln(n);  // label -> (count, mask)
    struct QE { unsigned s, e, l; };
    std::vector<QE> edges;
    while (in.get(c)) {
        if (c != 'e') continue;
        unsigned a, b, l;
        if (!(in >> a >> b >> l)) break;
        if (a >= n || b >= n || a == b) { err = "query edge"; return false; }
        for (int d = 0; d < 2; ++d) {
            unsigned x = d ? b : a, y = d ? a : b;
            allNb[x] |= 1u << y;
            auto& p = eln[x][l];
            p.first += 1;
            p.second |= 1u << y;
        }
        if (A[a * n + b] == 0xffffffffu) {
            edges.push_back({std::min(a, b), std::max(a, b), l});
            A[a * n + b] = A[b * n + a] = l;
        }
        nb[a] |= 1u << b;
        nb[b] |= 1u << a;
    }
    if (edges.empty()) { err = "no edges"; return false; }
    auto ql = [&](unsigned v) { return L[v]; };
    auto adj = [&](unsigned a, unsigned b) { return A[a * n + b]; };

    std::list<unsigned> order;
    unsigned inOrder = 0;
    std::map<uint64_t, unsigned> extKC, othKC;
    std::map<uint64_t, KC> redKC;
    std::vector<OPhase>& ph = P.phases;

    auto ext = [&](unsigned s1, unsigned x1, unsigned s2, unsigned x2, unsigned two) {
        // src/genMatchOrderCPU.cpp:288-334
        if (two == 2 && ql(x2) < ql(x1)) { std::swap(s1, s2); std::swap(x1, x2); }
        OPhase p{'E', adj(s1, x1), {s1, x1}};
        auto visit = [&](unsigned s, unsigned x) {
            extKC[((uint64_t)adj(s, x) << 32) | ql(x)] += 1;
            inOrder |= 1u << x;
            order.push_back(x);
            nb[s] &= ~(1u << x);
            nb[x] &= ~(1u << s);
        };
        visit(s1, x1);
        if (two == 2) {
            p.pairs.push_back(s2);
            p.pairs.push_back(x2);
            visit(s2, x2);
        }
        ph.push_back(p);
    };
    auto red = [&](const std::vector<unsigned>& pr) {
        // src/genMatchOrderCPU.cpp:335-395
        OPhase r{'R', 0, {}};
        for (unsigned i = 0; i < pr.size() / 2; ++i) {
            unsigned a = pr[2 * i], b = pr[2 * i + 1];
            r.pairs.push_back(a);
            r.pairs.push_back(b);
            nb[a] &= ~(1u << b);
            nb[b] &= ~(1u << a);
            r.edgeLabel = adj(a, b);
            uint64_t k = ((uint64_t)adj(a, b) << 32) | ql(b);
            uint64_t k1 = ((uint64_t)adj(a, b) << 32) | ql(a);
            if (k == k1) { othKC[k] += 1; continue; }
            for (int d = 0; d < 2; ++d) {
                KC& kc = redKC[d ? k1 : k];
                kc.count += 1;
                kc.keys.push_back(d ? k : k1);
                kc.where.push_back((unsigned)ph.size());
                kc.where.push_back(i);
            }
        }
        ph.push_back(r);
    };

    // start (src/genMatchOrderCPU.cpp:754-820)
    std::vector<TI> tri(edges.size());
    for (unsigned i = 0; i < edges.size(); ++i) {
        const QE& e = edges[i];
        unsigned t = nb[e.s] & nb[e.e];
        tri[i].idx = i;
        tri[i].cnt = (unsigned)__builtin_popcount(t);
        if (!t) continue;
        int best = -1, bi = -1;
        for (unsigned v = 0; v < 32; ++v) {
            if (!(t & (1u << v))) continue;
            unsigned a = adj(e.s, v), b = adj(e.e, v);
            int sc = 10 * ((a == b) + (a == e.l) + (b == e.l)) +
                     ((ql(e.s) == ql(e.e)) + (ql(e.s) == ql(v)) + (ql(e.e) == ql(v)));
            if (sc > best) { best = sc; bi = (int)v; }
        }
        tri[i].score = (unsigned)best;
        tri[i].third = (unsigned)bi;
    }
    std::sort(tri.begin(), tri.end(), [](const TI& a, const TI& b) { return a.cnt > b.cnt; });
    if (tri[0].cnt == 0) {
        // src/genMatchOrderCPU.cpp:177-286
        std::map<uint64_t, std::vector<unsigned>> m3, m2, m1;
        for (unsigned v = 0; v < n; ++v) {
            for (auto& kv : eln[v]) {
                uint64_t k2 = ((uint64_t)kv.first << 20) | v;
                unsigned msk = kv.second.second;
                if (__builtin_popcount(msk) >= 2) {
                    std::map<unsigned, std::vector<unsigned>> byl;
                    for (unsigned b = 0; b < 32; ++b) if (msk & (1u << b)) byl[ql(b)].push_back(b);
                    for (auto& g : byl) {
                        if (g.second.size() >= 2) m3.insert({((uint64_t)g.first << 40) | k2, g.second});
                        else { auto& t = m2[k2]; t.insert(t.end(), g.second.begin(), g.second.end()); }
                    }
                } else {
                    for (unsigned b = 0; b < 32; ++b) if (msk & (1u << b)) m1[v].push_back(b);
                }
            }
        }
        bool pm = !m3.empty() || !m2.empty();
        auto& m = !m3.empty() ? m3 : (!m2.empty() ? m2 : m1);
        unsigned gs = 0xffffffffu, g1 = 0, g2 = 0, ge1 = 0xffffffffu, ge2 = 0xffffffffu;
        for (auto& kv : m) {
            unsigned f1 = 0, f2 = 0, e1 = 0xffffffffu, e2 = 0xffffffffu;
            for (unsigned x : kv.second) {
                unsigned d = (unsigned)__builtin_popcount(allNb[x]);
                if (d > f1) { f2 = f1; e2 = e1; f1 = d; e1 = x; }
                else if (d > f2) { f2 = d; e2 = x; }
            }
            if (pm ? (f1 + f2 > g1 + g2) : (f1 > g1)) {
                g1 = f1; ge1 = e1; gs = (unsigned)(kv.first & 0xfffff);
                if (pm) { g2 = f2; ge2 = e2; }
            }
        }
        order.push_back(gs);
        inOrder |= 1u << gs;
        ext(gs, ge1, gs, ge2, ge2 == 0xffffffffu ? 1 : 2);
    } else {
        unsigned i = 1, pi = 0;
        while (i < tri.size() && tri[i].cnt == tri[0].cnt) {
            if (tri[i].score > tri[pi].score) pi = i;
            ++i;
        }
        unsigned sc = tri[pi].score;
        unsigned s = edges[tri[pi].idx].s, e = edges[tri[pi].idx].e, t = tri[pi].third;
        unsigned src = s, x1 = e, x2 = t;
        if (sc == 31) {
            if (ql(s) == ql(e)) { src = t; x1 = s; x2 = e; }
            else if (ql(s) == ql(t)) { src = e; x1 = s; x2 = t; }
            else if (ql(e) == ql(t)) { src = s; x1 = e; x2 = t; }
        } else if (sc >= 10 && sc <= 13) {
            unsigned a = adj(s, t), b = adj(e, t), el = edges[tri[pi].idx].l;
            if (a == b) { src = t; x1 = s; x2 = e; }
            else if (a == el) { src = s; x1 = e; x2 = t; }
            else if (b == el) { src = e; x1 = s; x2 = t; }
        }
        order.push_back(src);
        inOrder |= 1u << src;
        if (sc == 33 || sc == 31 || sc == 30 || (sc >= 10 && sc <= 13)) {
            ext(src, x1, src, x2, 2);
        } else {
            ext(src, x1, src, x2, 1);
            ext(src, x2, src, x2, 1);
        }
        red({x1, x2});
    }
    // greedy loop (src/genMatchOrderCPU.cpp:837-976)
    while (order.size() < n) {
        std::vector<NI> info;
        std::map<unsigned, std::vector<unsigned>> byEl;
        for (unsigned s : order) {
            for (unsigned x = 0; x < 32; ++x) {
                if (!(nb[s] & (1u << x))) continue;
                NI t;
                t.s = s; t.e = x; t.el = adj(s, x); t.evl = ql(x);
                t.common = (unsigned)__builtin_popcount(nb[x] & inOrder) - 1u;
                t.evn = (unsigned)__builtin_popcount(nb[s] & eln[s][t.el].second);
                info.push_back(t);
                if (!(inOrder & (1u << x))) { byEl[t.el].push_back(s); byEl[t.el].push_back(x); }
            }
        }
        if (info.empty()) { err = "disconnected query"; return false; }
        std::sort(info.begin(), info.end(), [](const NI& a, const NI& b) {
            if (a.common != b.common) return a.common > b.common;
            if (a.evn != b.evn) return a.evn > b.evn;
            if (a.el != b.el) return a.el < b.el;
            if (a.s != b.s) return a.s < b.s;
            return a.evl < b.evl;
        });
        if (info[0].common > 0) {
            unsigned x = info[0].e;
            ext(info[0].s, x, 0, 0, 1);
            std::vector<NI> back;
            for (unsigned v = 0; v < 32; ++v)
                if ((nb[x] & inOrder) & (1u << v)) {
                    NI t; t.s = x; t.e = v; t.el = adj(x, v); t.evl = ql(v);
                    back.push_back(t);
                }
            std::sort(back.begin(), back.end(), [](const NI& a, const NI& b) {
                return a.el != b.el ? a.el < b.el : a.evl < b.evl;
            });
            size_t g0 = 0;
            for (size_t k = 1; k <= back.size(); ++k) {
                if (k == back.size() || back[k].el != back[g0].el) {
                    std::vector<unsigned> pr;
                    for (size_t q = g0; q < k; ++q) { pr.push_back(back[q].s); pr.push_back(back[q].e); }
                    red(pr);
                    g0 = k;
                }
            }
        } else {
            bool found = false;
            for (auto& kv : byEl) {
                auto& v = kv.second;
                for (unsigned a = 0; a < v.size() / 2 && !found; ++a)
                    for (unsigned b = a + 1; b < v.size() / 2; ++b)
                        if (adj(v[2 * a + 1], v[2 * b + 1]) != 0xffffffffu) {
                            found = true;
                            ext(v[2 * a], v[2 * a + 1], v[2 * b], v[2 * b + 1], 2);
                            red({v[2 * a + 1], v[2 * b + 1]});
                            break;
                        }
                if (found) break;
            }
            if (!found) {
                if (info[0].evn > 1) {
                    ext(info[0].s, info[0].e, info[1].s, info[1].e, 2);
                } else {
                    bool pf = false;
                    for (size_t k = 0; k + 1 < info.size() && !pf; ++k)
                        if (info[k].el == info[k + 1].el) {
                            ext(info[k].s, info[k].e, info[k + 1].s, info[k + 1].e, 2);
                            pf = true;
                        }
                    if (!pf) ext(info[0].s, info[0].e, 0, 0, 1);
                }
            }
        }
    }
    // automorphisms (src/genMatchOrderCPU.cpp:397-463)
    auto equiv = [&](unsigned a, unsigned b) {
        if (a == b) return true;
        if (ql(a) != ql(b) || eln[a].size() != eln[b].size()) return false;
        for (auto ia = eln[a].begin(), ib = eln[b].begin(); ia != eln[a].end(); ++ia, ++ib)
            if (ia->first != ib->first || ia->second.first != ib->second.first) return false;
        return true;
    };
    std::vector<std::vector<unsigned>> auts;
    std::vector<unsigned> cand;
    std::vector<bool> used(n, false);
    std::vector<std::vector<unsigned>> stackAuts;
    std::function<void(bool)> rec;
    auto possible = [&](unsigned nid) {
        unsigned o = (unsigned)cand.size();
        if (!equiv(o, nid)) return false;
        for (unsigned i = 0; i < cand.size(); ++i)
            if (adj(i, o) != adj(cand[i], nid)) return false;
        return true;
    };
    rec = [&](bool ident) {
        if (cand.size() == n - 1) {
            for (unsigned i = 0; i < n; ++i) {
                if (used[i]) continue;
                if (ident && i == cand.size()) return;
                if (possible(i)) { cand.push_back(i); auts.push_back(cand); cand.pop_back(); }
                break;
            }
            return;
        }
        for (unsigned i = 0; i < n; ++i) {
            if (used[i] || !possible(i)) continue;
            cand.push_back(i);
            used[i] = true;
            rec(ident && i == cand.size() - 1);
            used[i] = false;
            cand.pop_back();
        }
    };
    rec(true);
    P.nAut = auts.size();
    // reduction-pair orientation (src/genMatchOrderCPU.cpp:509-557): the key
    // chosen each round is the last in map order; pairs whose first vertex
    // carries the key's vertex label are flipped.
    while (!redKC.empty()) {
        auto it = std::prev(redKC.end());
        uint64_t key = it->first;
        for (size_t i = 0; i + 1 < it->second.where.size(); i += 2) {
            std::vector<unsigned>& pr = ph[it->second.where[i]].pairs;
            unsigned j = it->second.where[i + 1];
            if ((uint64_t)ql(pr[2 * j]) == (key & 0xffffffffull)) std::swap(pr[2 * j], pr[2 * j + 1]);
        }
        redKC.erase(it);
    }
    // restrictions (src/genMatchOrderCPU.cpp:468-507)
    std::vector<bool> seen(n, false);
    for (auto& p : ph) {
        if (p.type == 'R') continue;
        for (unsigned v : p.pairs) {
            if (seen[v]) continue;
            seen[v] = true;
            std::vector<std::vector<unsigned>> keep;
            for (auto& a : auts) {
                if (a[v] == v) { keep.push_back(a); continue; }
                auto r = P.restricts.find(a[v]);
                if (r == P.restricts.end()) P.restricts[a[v]] = v;
                else if (r->second < v) r->second = v;
            }
            auts.swap(keep);
        }
    }
    for (unsigned i = 0; i <= n; ++i)
        if (!P.restricts.count(i)) P.restricts[i] = 0xffffffffu;
    return true;
}

std::string dumpPlan(const OPlan& P) {
    std::ostringstream o;
    for (auto& kv : P.restricts) o << "restrict: " << kv.first << "<" << kv.second << "\n";
    for (size_t i = 0; i < P.phases.size(); ++i) {
        const OPhase& p = P.phases[i];
        o << "phase " << i << " type " << p.type << " edge label:" << p.edgeLabel << "\n";
        for (size_t j = 0; j < p.pairs.size() / 2; ++j)
            o << "svid:" << p.pairs[2 * j] << " evid:" << p.pairs[2 * j + 1] << " ";
        o << "\n";
    }
    return o.str();
}

}  // namespace oracle
