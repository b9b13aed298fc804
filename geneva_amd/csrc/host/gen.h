// Deterministic synthetic inputs: labelled graphs, databases of graphs and
// random-walk query patterns.  Every draw comes from the project's own PRNG
// (SplitMix64-seeded xoshiro256**) and its own bounded-integer and Zipf
// samplers, never from std::*_distribution, so a seed fixes the bytes on
// every machine.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "loadDataBase.h"

namespace geneva {

class Rng {
public:
    explicit Rng(uint64_t seed);
    uint64_t next();
    uint64_t below(uint64_t n);          // uniform in [0, n), Lemire's method
    double unit();                       // uniform in [0, 1)
    double unitOpen();                   // uniform in (0, 1]
private:
    uint64_t s[4];
};

// Inverse-CDF sampler of P(k) ~ k^-s over k = 1..n.
class Zipf {
public:
    Zipf(uint32_t n, double s);
    uint32_t draw(Rng& r) const;         // 1..n
private:
    std::vector<double> cdf;
};

// A generated data set: the graph (ids 1..nv, edge label 1 everywhere),
// the owner graph of every vertex for databases (graphOf[0] = 0) and query
// texts in the .query format.
struct Dataset {
    DataGraph g;
    std::vector<uint32_t> graphOf;
    uint32_t ngraphs = 1;
    std::vector<std::string> queries;
};

// Database of `ngraphs` graphs, n ~ U[nmin, nmax]; a fraction baFraction are
// Barabasi-Albert (m = baM, grown from a triangle), the rest Erdos-Renyi with
// p = avgDeg/(n-1).  Labels Zipf(zipfS) over nlabels (zipfS = 0: uniform).
// Ids: every vertex with an edge first, graph by graph (so each graph owns a
// contiguous id range and the partition has one interval), isolated vertices
// after them.
Dataset genDatabase(uint32_t ngraphs, uint32_t nmin, uint32_t nmax, double avgDeg,
                    double baFraction, uint32_t baM, uint32_t nlabels, double zipfS,
                    uint64_t seed);

// Chung-Lu graph: weights w_i = wmax ((i+i0)/i0)^(-1/(gamma-1)) with i0 set so
// the mean weight is meanDeg, edges drawn with probability min(1, w_i w_j / S)
// by Miller-Hagberg edge skipping; wmax = 0 picks sqrt(S) (the largest weight
// for which no probability is capped).  Ids: vertices with edges first, in
// weight order, isolated ones after.
Dataset genChungLu(uint32_t nv, double gamma, double meanDeg, double wmax, uint32_t nlabels,
                   double zipfS, uint64_t seed);

// Random-walk query extraction (tools/genQuery.cpp:13-128): start at a random
// vertex with an edge (in a random graph of a database), repeatedly pick a
// random chosen vertex and a random neighbour, adding the vertex and the walk
// edge, until k distinct vertices.  induced: add every data edge among the
// chosen vertices.  minExtra: resample until |E| >= k - 1 + minExtra.
// minDeg: resample until every query vertex has at least minDeg query
// neighbours (2: every vertex lies on a cycle).
// k ~ U[kmin, kmax].  Query vertex 0 is the walk start.
std::vector<std::string> genWalkQueries(const Dataset& ds, uint32_t nq, uint32_t kmin,
                                        uint32_t kmax, bool induced, uint32_t minExtra,
                                        uint64_t seed, uint32_t minDeg = 0);

// Fixed shapes on 5 vertices: path, star, cycle, tailed triangle (triangle
// plus a 2-edge tail) and a 4-clique with a pendant vertex.  Vertex labels are
// drawn from the empirical label distribution (the label of a uniformly
// random vertex).
std::vector<std::string> genShapeQueries(const Dataset& ds, uint64_t seed);

// .query text of a pattern: vertex labels, undirected edges with the given
// edge labels (label 1 when edgeLabels is null).
std::string queryText(const std::vector<uint32_t>& labels,
                      const std::vector<std::pair<uint32_t, uint32_t>>& edges,
                      const std::vector<uint32_t>* edgeLabels = nullptr);

}  // namespace geneva
