// Deterministic synthetic molecule workloads.
//
// The reference converts RDKit SMARTS files (generate_data.py:22-92); those datasets and RDKit are
// not available, so workloads are generated here.  Labels and bond codes follow
// generate_data.py:8-17 (N=0 Cl=1 *=2 Br=3 I=4 P=5 H=6 O=7 C=8 S=9 F=10; single=1 double=2
// aromatic=4).  Every random draw comes from a project-owned xoshiro256** stream with Lemire's
// bounded integers, seeded per graph from (seed, stream, index), so results are bit-stable across
// machines and any prefix of a batch equals the smaller batch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "hostgraph.hpp"

namespace bms2 {
namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  Rng(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t x = seed * 0xd1342543de82ef95ull + stream;
    x = splitmix(x) ^ (index * 0x9e3779b97f4a7c15ull);
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Uniform integer in [0, n), Lemire's multiply-shift with rejection.
  uint32_t below(uint32_t n) {
    uint64_t m = uint64_t(uint32_t(next() >> 32)) * n;
    uint32_t l = uint32_t(m);
    if (l < n) {
      uint32_t t = uint32_t(-n) % n;
      while (l < t) {
        m = uint64_t(uint32_t(next() >> 32)) * n;
        l = uint32_t(m);
      }
    }
    return uint32_t(m >> 32);
  }
  uint32_t range(uint32_t lo, uint32_t hi) { return lo + below(hi - lo + 1); }
  double uniform01() { return double(next() >> 11) * 0x1.0p-53; }
  // Standard normal by Box-Muller (one value per call).
  double normal() {
    double u1 = uniform01(), u2 = uniform01();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

enum : uint8_t { kN = 0, kCl = 1, kBr = 3, kO = 7, kC = 8, kS = 9, kF = 10 };
enum : uint8_t { kSingle = 1, kDouble = 2, kAromatic = 4 };

// Element frequencies in permille: C .72, O .12, N .10, S .02, F .02, Cl .015, Br .005.
uint8_t draw_element(Rng& r) {
  uint32_t x = r.below(1000);
  if (x < 720) return kC;
  if (x < 840) return kO;
  if (x < 940) return kN;
  if (x < 960) return kS;
  if (x < 980) return kF;
  if (x < 995) return kCl;
  return kBr;
}

// Bond frequencies: single .60, double .15, aromatic .25.
uint8_t draw_bond(Rng& r) {
  uint32_t x = r.below(100);
  if (x < 60) return kSingle;
  if (x < 75) return kDouble;
  return kAromatic;
}

constexpr uint32_t kMaxDegree = 4;

struct Builder {
  EdgeListGraph g;
  std::vector<std::vector<uint32_t>> adj;
  uint32_t add_node(uint8_t label) {
    g.labels.push_back(label);
    adj.emplace_back();
    return g.num_nodes() - 1;
  }
  void add_edge(uint32_t u, uint32_t v, uint8_t label) {
    g.edges.push_back({u, v, label});
    adj[u].push_back(v);
    adj[v].push_back(u);
  }
  bool adjacent(uint32_t u, uint32_t v) const {
    return std::find(adj[u].begin(), adj[u].end(), v) != adj[u].end();
  }
  uint32_t deg(uint32_t u) const { return uint32_t(adj[u].size()); }
  // Attaches a new node to an existing node of degree < 4 picked uniformly (or, when `branchy`,
  // to the existing node of highest degree below 4 half of the time).
  bool attach(Rng& r, bool branchy) {
    std::vector<uint32_t> eligible;
    for (uint32_t u = 0; u < g.num_nodes(); ++u)
      if (deg(u) < kMaxDegree) eligible.push_back(u);
    if (eligible.empty()) return false;
    uint32_t p;
    if (branchy && r.below(2) == 0) {
      p = eligible[0];
      for (uint32_t u : eligible)
        if (deg(u) > deg(p)) p = u;
    } else {
      p = eligible[r.below(uint32_t(eligible.size()))];
    }
    uint32_t v = add_node(draw_element(r));
    add_edge(p, v, draw_bond(r));
    return true;
  }
  // Adds one edge closing a ring of `ring` nodes: picks uniformly among non-adjacent node pairs
  // of degree < 4 at shortest-path distance ring-1.
  bool close_ring(Rng& r, uint32_t ring) {
    const uint32_t n = g.num_nodes();
    std::vector<std::pair<uint32_t, uint32_t>> cand;
    std::vector<int> dist(n);
    std::vector<uint32_t> queue(n);
    for (uint32_t u = 0; u < n; ++u) {
      if (deg(u) >= kMaxDegree) continue;
      std::fill(dist.begin(), dist.end(), -1);
      uint32_t qh = 0, qt = 0;
      queue[qt++] = u;
      dist[u] = 0;
      while (qh < qt) {
        uint32_t x = queue[qh++];
        if (dist[x] >= int(ring) - 1) continue;
        for (uint32_t y : adj[x])
          if (dist[y] < 0) {
            dist[y] = dist[x] + 1;
            queue[qt++] = y;
          }
      }
      for (uint32_t v = u + 1; v < n; ++v)
        if (dist[v] == int(ring) - 1 && deg(v) < kMaxDegree) cand.push_back({u, v});
    }
    if (cand.empty()) return false;
    auto [a, b] = cand[r.below(uint32_t(cand.size()))];
    add_edge(a, b, draw_bond(r));
    return true;
  }
};

// Random tree with max degree 4 plus 0..max_closures ring-closing edges of 5- or 6-rings.
EdgeListGraph molecule(Rng& r, uint32_t n, uint32_t max_closures, bool branchy = false) {
  Builder b;
  if (n == 0) return b.g;
  b.add_node(draw_element(r));
  for (uint32_t i = 1; i < n; ++i) b.attach(r, branchy);
  uint32_t closures = r.below(max_closures + 1);
  for (uint32_t c = 0; c < closures; ++c) b.close_ring(r, 5 + r.below(2));
  return b.g;
}

// A fused ring system of `rings` 5/6-rings sharing edges, then tree atoms up to n nodes.
EdgeListGraph polycyclic(Rng& r, uint32_t n, uint32_t rings) {
  Builder b;
  uint32_t first = 5 + r.below(2);
  for (uint32_t i = 0; i < first; ++i) b.add_node(draw_element(r));
  for (uint32_t i = 0; i < first; ++i) b.add_edge(i, (i + 1) % first, kAromatic);
  std::vector<std::pair<uint32_t, uint32_t>> ring_edges;
  for (uint32_t i = 0; i < first; ++i) ring_edges.push_back({i, (i + 1) % first});
  for (uint32_t k = 1; k < rings; ++k) {
    std::vector<uint32_t> fusable;
    for (uint32_t e = 0; e < ring_edges.size(); ++e)
      if (b.deg(ring_edges[e].first) == 2 && b.deg(ring_edges[e].second) == 2) fusable.push_back(e);
    if (fusable.empty()) break;
    auto [u, v] = ring_edges[fusable[r.below(uint32_t(fusable.size()))]];
    uint32_t size = 5 + r.below(2);
    uint32_t prev = u;
    for (uint32_t i = 0; i + 2 < size; ++i) {
      uint32_t x = b.add_node(draw_element(r));
      b.add_edge(prev, x, kAromatic);
      ring_edges.push_back({prev, x});
      prev = x;
    }
    b.add_edge(prev, v, kAromatic);
    ring_edges.push_back({prev, v});
  }
  while (b.g.num_nodes() < n)
    if (!b.attach(r, false)) break;
  return b.g;
}

// A single 5/6-ring with tree atoms attached up to n nodes (n >= 5).
EdgeListGraph ring_query(Rng& r, uint32_t n) {
  Builder b;
  uint32_t ring = std::min<uint32_t>(n, 5 + r.below(2));
  for (uint32_t i = 0; i < ring; ++i) b.add_node(draw_element(r));
  for (uint32_t i = 0; i < ring; ++i) b.add_edge(i, (i + 1) % ring, draw_bond(r));
  while (b.g.num_nodes() < n)
    if (!b.attach(r, false)) break;
  return b.g;
}

// Connected induced subgraph of `src` with up to `s` nodes, grown from a random node by adding a
// uniformly chosen frontier node at each step; query node k is the k-th node added.
EdgeListGraph extract(Rng& r, const EdgeListGraph& src, uint32_t s) {
  const uint32_t n = src.num_nodes();
  EdgeListGraph q;
  if (n == 0) return q;
  std::vector<std::vector<std::pair<uint32_t, uint8_t>>> adj(n);
  for (const auto& e : src.edges) {
    adj[e.u].push_back({e.v, e.label});
    adj[e.v].push_back({e.u, e.label});
  }
  std::vector<int> qid(n, -1);
  std::vector<uint32_t> order, frontier;
  std::vector<char> in_frontier(n, 0);
  auto take = [&](uint32_t v) {
    qid[v] = int(order.size());
    order.push_back(v);
    for (auto [w, l] : adj[v])
      if (qid[w] < 0 && !in_frontier[w]) {
        in_frontier[w] = 1;
        frontier.push_back(w);
      }
  };
  take(r.below(n));
  while (order.size() < s && !frontier.empty()) {
    uint32_t i = r.below(uint32_t(frontier.size()));
    uint32_t v = frontier[i];
    frontier.erase(frontier.begin() + i);
    in_frontier[v] = 0;
    take(v);
  }
  for (uint32_t v : order) q.labels.push_back(src.labels[v]);
  for (uint32_t k = 0; k < order.size(); ++k)
    for (auto [w, l] : adj[order[k]])
      if (qid[w] > int(k)) q.edges.push_back({k, uint32_t(qid[w]), l});
  return q;
}

template <class F>
void parallel_range(size_t n, unsigned threads, F&& f) {
  if (threads <= 1 || n < 256) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (size_t i = t; i < n; i += threads) f(i);
    });
  for (auto& th : pool) th.join();
}

uint32_t data_size(Rng& r, const bms2_workload_spec& s) {
  if (s.data_size_dist == 1) {
    double x = std::exp(s.data_lognormal_mu + s.data_lognormal_sigma * r.normal());
    long v = std::lround(x);
    v = std::max<long>(v, s.data_min_nodes);
    v = std::min<long>(v, s.data_max_nodes);
    return uint32_t(v);
  }
  return r.range(s.data_min_nodes, s.data_max_nodes);
}

}  // namespace
}  // namespace bms2

extern "C" int bms2_generate_workload(const bms2_workload_spec* spec, bms2_host_csr* data_out,
                                      bms2_host_csr* query_out) {
  using namespace bms2;
  if (!spec || !data_out || !query_out) return set_error(BMS2_ERR_INVALID, "null argument");
  const bms2_workload_spec s = *spec;
  if (s.data_min_nodes == 0 || s.data_min_nodes > s.data_max_nodes ||
      s.data_max_nodes > BMS2_MAX_GRAPH_NODES)
    return set_error(BMS2_ERR_INVALID, "data node range [%u, %u] invalid", s.data_min_nodes,
                     s.data_max_nodes);
  if (s.query_min_nodes == 0 || s.query_min_nodes > s.query_max_nodes ||
      s.query_max_nodes > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_INVALID, "query node range [%u, %u] invalid", s.query_min_nodes,
                     s.query_max_nodes);
  if (s.num_query_graphs > 0 && s.extracted_permille > 0 && s.num_data_graphs == 0)
    return set_error(BMS2_ERR_INVALID, "extracted queries need data graphs");
  unsigned threads = s.num_threads ? s.num_threads : std::max(1u, std::thread::hardware_concurrency());

  std::vector<EdgeListGraph> data(s.num_data_graphs);
  parallel_range(s.num_data_graphs, threads, [&](size_t i) {
    Rng r(s.seed, 1, i);
    uint32_t n = data_size(r, s);
    bool poly = s.polycyclic_permille > 0 && r.below(1000) < s.polycyclic_permille;
    if (poly) {
      uint32_t rings = r.range(std::max(2u, s.poly_min_rings), std::max(2u, s.poly_max_rings));
      data[i] = polycyclic(r, n, rings);
    } else {
      data[i] = molecule(r, n, s.max_ring_closures);
    }
  });

  std::vector<EdgeListGraph> queries(s.num_query_graphs);
  parallel_range(s.num_query_graphs, threads, [&](size_t i) {
    Rng r(s.seed, 2, i);
    uint32_t n = r.range(s.query_min_nodes, s.query_max_nodes);
    bool extracted = r.below(1000) < s.extracted_permille;
    if (extracted) {
      // Retry a bounded number of graphs for one large enough; otherwise take what is there.
      EdgeListGraph best;
      for (int attempt = 0; attempt < 64; ++attempt) {
        const EdgeListGraph& src = data[r.below(s.num_data_graphs)];
        EdgeListGraph q = extract(r, src, n);
        if (q.num_nodes() > best.num_nodes()) best = std::move(q);
        if (best.num_nodes() >= n) break;
      }
      queries[i] = std::move(best);
    } else if (s.query_shapes == 1) {
      uint32_t shape = r.below(3);
      if (shape == 0 && n >= 5)
        queries[i] = ring_query(r, n);
      else if (shape == 1)
        queries[i] = molecule(r, n, 0, true);
      else
        queries[i] = molecule(r, n, s.max_ring_closures);
    } else {
      queries[i] = molecule(r, n, s.max_ring_closures);
    }
  });

  int rc = pack_graphs(data, data_out, threads);
  if (rc != BMS2_OK) return rc;
  rc = pack_graphs(queries, query_out, threads);
  if (rc != BMS2_OK) {
    bms2_host_csr_free(data_out);
    return rc;
  }
  return BMS2_OK;
}
