// Deterministic synthetic graphs and queries.
#include "gen.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <map>
#include <set>
#include <sstream>

namespace geneva {

namespace {
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
}  // namespace

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix64(x);
}

uint64_t Rng::next() {  // xoshiro256**
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

uint64_t Rng::below(uint64_t n) {  // Lemire, with rejection of the biased low range
    if (n <= 1) return 0;
    uint64_t x = next();
    __uint128_t m = (__uint128_t)x * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
        uint64_t t = (0 - n) % n;
        while (l < t) {
            x = next();
            m = (__uint128_t)x * n;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

double Rng::unit() { return (double)(next() >> 11) * 0x1.0p-53; }
double Rng::unitOpen() { return ((double)(next() >> 11) + 1.0) * 0x1.0p-53; }

Zipf::Zipf(uint32_t n, double s) {
    cdf.resize(n);
    double acc = 0;
    for (uint32_t k = 1; k <= n; ++k) {
        acc += std::pow((double)k, -s);
        cdf[k - 1] = acc;
    }
    for (auto& c : cdf) c /= acc;
    cdf.back() = 1.0;
}

uint32_t Zipf::draw(Rng& r) const {
    double u = r.unit();
    return (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
}

namespace {

// Renumber: vertices with an edge get ids 1..S in the given order, isolated
// ones S+1..nv.  localLabel/localOwner are indexed by the original index.
void finish(Dataset& ds, uint32_t n, const std::vector<uint32_t>& label,
            const std::vector<uint32_t>& owner,
            const std::vector<std::pair<uint32_t, uint32_t>>& edges) {
    std::vector<uint32_t> deg(n, 0);
    for (auto& e : edges) {
        deg[e.first]++;
        deg[e.second]++;
    }
    std::vector<uint32_t> id(n, 0);
    uint32_t next = 1;
    for (uint32_t v = 0; v < n; ++v)
        if (deg[v]) id[v] = next++;
    for (uint32_t v = 0; v < n; ++v)
        if (!deg[v]) id[v] = next++;
    ds.g.nv = n;
    ds.g.vlabel.assign(n + 1, 0);
    ds.graphOf.assign(n + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
        ds.g.vlabel[id[v]] = label[v];
        ds.graphOf[id[v]] = owner.empty() ? 0 : owner[v];
    }
    ds.g.edges.clear();
    ds.g.edges.reserve(edges.size());
    for (auto& e : edges) ds.g.edges.push_back({id[e.first], id[e.second], 1u});
}

}  // namespace

Dataset genDatabase(uint32_t ngraphs, uint32_t nmin, uint32_t nmax, double avgDeg,
                    double baFraction, uint32_t baM, uint32_t nlabels, double zipfS,
                    uint64_t seed) {
    Rng r(seed);
    Zipf z(nlabels, zipfS);
    Dataset ds;
    ds.ngraphs = ngraphs;
    std::vector<uint32_t> label, owner;
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    uint32_t base = 0;
    for (uint32_t gi = 0; gi < ngraphs; ++gi) {
        const uint32_t n = nmin + (uint32_t)r.below(nmax - nmin + 1);
        const bool ba = r.unit() < baFraction;
        for (uint32_t v = 0; v < n; ++v) {
            label.push_back(zipfS == 0 ? 1 + (uint32_t)r.below(nlabels) : z.draw(r));
            owner.push_back(gi);
        }
        if (ba && n > baM + 1) {
            // grow from a clique on baM+1 vertices; each new vertex picks baM
            // distinct targets with probability proportional to degree
            std::vector<uint32_t> ends;
            for (uint32_t a = 0; a <= baM; ++a)
                for (uint32_t b = a + 1; b <= baM; ++b) {
                    edges.emplace_back(base + a, base + b);
                    ends.push_back(a);
                    ends.push_back(b);
                }
            for (uint32_t v = baM + 1; v < n; ++v) {
                std::vector<uint32_t> t;
                while (t.size() < baM) {
                    uint32_t c = ends[r.below(ends.size())];
                    if (std::find(t.begin(), t.end(), c) == t.end()) t.push_back(c);
                }
                for (uint32_t c : t) {
                    edges.emplace_back(base + c, base + v);
                    ends.push_back(c);
                    ends.push_back(v);
                }
            }
        } else if (n >= 2) {
            // G(n, p) by geometric skipping over the pair index space
            const double p = std::min(1.0, avgDeg / (double)(n - 1));
            const double lq = std::log(1.0 - p);
            const uint64_t npairs = (uint64_t)n * (n - 1) / 2;
            uint64_t k = 0;
            if (p >= 1.0) {
                for (uint32_t a = 0; a < n; ++a)
                    for (uint32_t b = a + 1; b < n; ++b) edges.emplace_back(base + a, base + b);
            } else if (p > 0) {
                uint32_t a = 0;
                uint64_t rowStart = 0;  // pair index of (a, a+1)
                while (true) {
                    k += (uint64_t)std::floor(std::log(r.unitOpen()) / lq);
                    if (k >= npairs) break;
                    while (k >= rowStart + (n - 1 - a)) {
                        rowStart += n - 1 - a;
                        ++a;
                    }
                    uint32_t b = a + 1 + (uint32_t)(k - rowStart);
                    edges.emplace_back(base + a, base + b);
                    ++k;
                }
            }
        }
        base += n;
    }
    finish(ds, base, label, owner, edges);
    return ds;
}

Dataset genChungLu(uint32_t nv, double gamma, double meanDeg, double wmax, uint32_t nlabels,
                   double zipfS, uint64_t seed) {
    Rng r(seed);
    Zipf z(nlabels, zipfS);
    const double alpha = 1.0 / (gamma - 1.0);
    if (wmax <= 0) wmax = std::sqrt(meanDeg * (double)nv);
    auto meanFor = [&](double i0) {
        double s = 0;
        for (uint32_t i = 0; i < nv; ++i) s += std::pow((i + i0) / i0, -alpha);
        return wmax * s / nv;
    };
    double lo = 1e-3, hi = (double)nv * 1e3;
    for (int it = 0; it < 60; ++it) {
        double mid = std::sqrt(lo * hi);
        if (meanFor(mid) < meanDeg) lo = mid;
        else hi = mid;
    }
    const double i0 = std::sqrt(lo * hi);
    std::vector<double> w(nv);
    double S = 0;
    for (uint32_t i = 0; i < nv; ++i) {
        w[i] = wmax * std::pow((i + i0) / i0, -alpha);
        S += w[i];
    }
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    edges.reserve((size_t)(S / 2 * 1.05));
    for (uint32_t u = 0; u + 1 < nv; ++u) {
        uint64_t v = u + 1;
        double p = std::min(w[u] * w[v] / S, 1.0);
        while (v < nv && p > 0) {
            if (p < 1.0) v += (uint64_t)std::floor(std::log(r.unitOpen()) / std::log(1.0 - p));
            if (v < nv) {
                double q = std::min(w[u] * w[v] / S, 1.0);
                if (r.unit() < q / p) edges.emplace_back(u, (uint32_t)v);
                p = q;
                ++v;
            }
        }
    }
    std::vector<uint32_t> label(nv);
    for (uint32_t i = 0; i < nv; ++i) label[i] = zipfS == 0 ? 1 + (uint32_t)r.below(nlabels) : z.draw(r);
    Dataset ds;
    ds.ngraphs = 1;
    finish(ds, nv, label, {}, edges);
    return ds;
}

std::string queryText(const std::vector<uint32_t>& labels,
                      const std::vector<std::pair<uint32_t, uint32_t>>& edges,
                      const std::vector<uint32_t>* edgeLabels) {
    std::ostringstream o;
    for (size_t i = 0; i < labels.size(); ++i) o << "v " << i << " " << labels[i] << "\n";
    for (size_t i = 0; i < edges.size(); ++i)
        o << "e " << edges[i].first << " " << edges[i].second << " "
          << (edgeLabels ? (*edgeLabels)[i] : 1u) << "\n";
    return o.str();
}

namespace {
struct Csr {
    std::vector<uint64_t> off;
    std::vector<uint64_t> nbr;  // (neighbour << 32) | edge label, sorted
};
Csr csrOf(const DataGraph& g) {
    Csr c;
    c.off.assign(g.nv + 2, 0);
    for (auto& e : g.edges) {
        if (e.a == e.b) continue;
        c.off[e.a + 1]++;
        c.off[e.b + 1]++;
    }
    for (uint32_t v = 1; v <= g.nv + 1; ++v) c.off[v] += c.off[v - 1];
    c.nbr.resize(c.off[g.nv + 1]);
    std::vector<uint64_t> p(c.off.begin(), c.off.end() - 1);
    for (auto& e : g.edges) {
        if (e.a == e.b) continue;
        c.nbr[p[e.a]++] = ((uint64_t)e.b << 32) | e.label;
        c.nbr[p[e.b]++] = ((uint64_t)e.a << 32) | e.label;
    }
    for (uint32_t v = 0; v <= g.nv; ++v)
        std::sort(c.nbr.begin() + c.off[v], c.nbr.begin() + c.off[v + 1]);
    return c;
}
// label of the first stored edge a-b (0 if none)
uint32_t edgeLabelOf(const Csr& c, uint32_t a, uint32_t b) {
    auto lo = std::lower_bound(c.nbr.begin() + c.off[a], c.nbr.begin() + c.off[a + 1],
                               (uint64_t)b << 32);
    if (lo != c.nbr.begin() + c.off[a + 1] && (uint32_t)(*lo >> 32) == b) return (uint32_t)*lo;
    return 0;
}
}  // namespace

std::vector<std::string> genWalkQueries(const Dataset& ds, uint32_t nq, uint32_t kmin,
                                        uint32_t kmax, bool induced, uint32_t minExtra,
                                        uint64_t seed, uint32_t minDeg) {
    Rng r(seed);
    Csr c = csrOf(ds.g);
    // start candidates: vertices with an edge, grouped by owner graph
    std::vector<uint32_t> starts;
    for (uint32_t v = 1; v <= ds.g.nv; ++v)
        if (c.off[v + 1] > c.off[v]) starts.push_back(v);
    std::vector<std::vector<uint32_t>> byGraph;
    if (ds.ngraphs > 1) {
        byGraph.resize(ds.ngraphs);
        for (uint32_t v : starts) byGraph[ds.graphOf[v]].push_back(v);
    }
    std::vector<std::string> out;
    while (out.size() < nq) {
        const uint32_t k = kmin + (uint32_t)r.below(kmax - kmin + 1);
        uint32_t start;
        if (ds.ngraphs > 1) {
            const auto& lst = byGraph[r.below(ds.ngraphs)];
            if (lst.empty()) continue;
            start = lst[r.below(lst.size())];
        } else {
            start = starts[r.below(starts.size())];
        }
        std::vector<uint32_t> chosen{start};
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> qe;  // query edge -> edge label
        uint32_t steps = 0;
        while (chosen.size() < k && steps < 200 * k) {
            ++steps;
            uint32_t si = (uint32_t)r.below(chosen.size());
            uint32_t s = chosen[si];
            uint64_t d = c.off[s + 1] - c.off[s];
            if (d == 0) continue;
            const uint64_t pick = c.nbr[c.off[s] + r.below(d)];
            const uint32_t x = (uint32_t)(pick >> 32);
            auto it = std::find(chosen.begin(), chosen.end(), x);
            uint32_t xi = (uint32_t)(it - chosen.begin());
            if (it == chosen.end()) chosen.push_back(x);
            qe.emplace(std::make_pair(std::min(si, xi), std::max(si, xi)), (uint32_t)pick);
        }
        if (chosen.size() < k) continue;
        if (induced) {
            for (uint32_t i = 0; i < k; ++i)
                for (uint32_t j = i + 1; j < k; ++j) {
                    uint32_t l = edgeLabelOf(c, chosen[i], chosen[j]);
                    if (l) qe.emplace(std::make_pair(i, j), l);
                }
        }
        if (qe.size() < (size_t)k - 1 + minExtra) continue;
        if (minDeg) {
            std::vector<uint32_t> qd(k, 0);
            for (auto& kv : qe) {
                qd[kv.first.first]++;
                qd[kv.first.second]++;
            }
            if (*std::min_element(qd.begin(), qd.end()) < minDeg) continue;
        }
        std::vector<uint32_t> labels(k);
        for (uint32_t i = 0; i < k; ++i) labels[i] = ds.g.vlabel[chosen[i]];
        std::vector<std::pair<uint32_t, uint32_t>> es;
        std::vector<uint32_t> els;
        for (auto& kv : qe) {
            es.push_back(kv.first);
            els.push_back(kv.second);
        }
        out.push_back(queryText(labels, es, &els));
    }
    return out;
}

std::vector<std::string> genShapeQueries(const Dataset& ds, uint64_t seed) {
    Rng r(seed);
    typedef std::vector<std::pair<uint32_t, uint32_t>> E;
    const std::vector<E> shapes = {
        {{0, 1}, {1, 2}, {2, 3}, {3, 4}},                          // path
        {{0, 1}, {0, 2}, {0, 3}, {0, 4}},                          // star
        {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {0, 4}},                  // 5-cycle
        {{0, 1}, {1, 2}, {0, 2}, {2, 3}, {3, 4}},                  // tailed triangle
        {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}, {3, 4}},  // 4-clique + pendant
    };
    std::vector<std::string> out;
    for (auto& e : shapes) {
        std::vector<uint32_t> labels(5);
        for (auto& l : labels) l = ds.g.vlabel[1 + r.below(ds.g.nv)];
        out.push_back(queryText(labels, e));
    }
    return out;
}

}  // namespace geneva
