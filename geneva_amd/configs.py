"""The five benchmark workloads and their scaled-down test forms.

Seeds are 0x47454E00 + workload number.  Every generator is the project's own
deterministic one, so a (workload, scale) pair fixes the bytes.

  1  database of 1,000 Erdos-Renyi graphs, n ~ U[10,30], average degree 3,
     8 uniform vertex labels; 10 random-walk queries of 4 vertices.
  2  Chung-Lu graph, exponent 2.1, 1M vertices, mean degree 20 (about 10M
     edges), largest weight sqrt(sum of weights); 16 labels Zipf(1.2);
     5 shapes of 5 vertices (path, star, cycle, tailed triangle, 4-clique +
     pendant), labels drawn from the vertex-label distribution.
  3  database of 100,000 graphs, n ~ U[20,200], half Erdos-Renyi (p = 4/(n-1)),
     half Barabasi-Albert (m = 2); 32 labels Zipf(1.1); 256 random-walk queries
     of 4-8 vertices.
  4  Chung-Lu graph, exponent 1.9, 4M vertices, mean degree 30, largest
     weight 1e5; 64 labels Zipf(1.0); 64 induced random-walk queries of 6-8
     vertices with at least two independent cycles and every query vertex on a
     cycle (a pendant vertex hung on a hub of degree ~5e4 gives a single query
     ~1e15 embeddings, beyond any enumerator).
  5  workload 3's database with a Poisson request stream of random-walk
     queries of 3-10 vertices.
"""
from . import Dataset

SEED = 0x47454E00


def seed(n):
    return SEED + n


def workload1(scale=1.0):
    ds = Dataset.database(max(1, int(1000 * scale)), 10, 30, 3.0, 0.0, 8, 0.0, seed(1))
    return ds.add_walk_queries(10, 4, 4, seed=seed(1) + 100)


def workload2(scale=1.0):
    nv = max(1000, int(1_000_000 * scale))
    ds = Dataset.chung_lu(nv, 2.1, 20.0, 0.0, 16, 1.2, seed(2))
    return ds.add_shape_queries(seed(2) + 100)


def workload3(scale=1.0, nqueries=256):
    ds = Dataset.database(max(1, int(100_000 * scale)), 20, 200, 4.0, 0.5, 32, 1.1, seed(3))
    return ds.add_walk_queries(nqueries, 4, 8, seed=seed(3) + 100)


def workload4(scale=1.0, nqueries=64, min_deg=2):
    """min_deg: every query vertex has at least this many query neighbours (2:
    every vertex on a cycle; 0: the configuration's walk queries as they come,
    pendant vertices included)."""
    nv = max(2000, int(4_000_000 * scale))
    wmax = max(50.0, 1e5 * scale)
    ds = Dataset.chung_lu(nv, 1.9, 30.0, wmax, 64, 1.0, seed(4))
    return ds.add_walk_queries(nqueries, 6, 8, induced=True, min_extra=2, seed=seed(4) + 100,
                               min_deg=min_deg)


def workload5_queries(ds, n, seed_offset=0):
    """Random-walk queries of 3-10 vertices over workload 3's database."""
    before = len(ds.queries)
    ds.add_walk_queries(n, 3, 10, seed=seed(5) + 100 + seed_offset)
    return ds.queries[before:]


WORKLOADS = {1: workload1, 2: workload2, 3: workload3, 4: workload4}
