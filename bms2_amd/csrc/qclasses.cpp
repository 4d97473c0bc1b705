// Isomorphism classes of a query batch, on the host.
//
// Every field the filter and the refinement test for a query node -- its label, its degree, the
// path Bloom filter of the label sequences leaving it, its ring counts and its shell histograms
// (isomorphism.hpp:71-97,142-196) -- is computed from labels and adjacency alone, so a
// label-preserving isomorphism q -> q' carries each node's domain to its image's domain in every
// data graph.  Two isomorphic queries therefore have the same GMCR membership and the same
// induced-embedding count against every data graph (an embedding of q composed with the inverse
// isomorphism is one of q').  The pipeline screens and joins one query per class and copies its
// outputs to the others.  Edge labels take no part in matching (the reference never reads them).
//
// Classes are proven, never guessed: queries are bucketed by colour refinement (a necessary
// condition for isomorphism) and a query joins a class only after an explicit isomorphism to the
// class's first query has been found.  A search that exceeds its step budget, a query with
// parallel edges or a self-loop (its degree then differs from its neighbour count) and a query
// above 32 nodes stay classes of their own.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "common.hpp"

namespace {

constexpr uint64_t kSearchBudget = 1u << 16;  // backtracking steps per isomorphism test
// a query is compared with at most kBucketCompares classes of its colour-refinement bucket, and
// the orbit searches of one query share kOrbitBudget steps: beyond them (inputs colour refinement
// cannot tell apart, such as many non-isomorphic regular graphs of one label) queries and nodes
// stay alone, which keeps every result exact
constexpr uint32_t kBucketCompares = 32;
constexpr uint64_t kOrbitBudget = 8 * kSearchBudget;

uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

struct QueryGraph {
  uint32_t n = 0, m = 0;
  bool simple = true;                 // no parallel edge, no self-loop, at most 32 nodes
  uint32_t adj[32] = {};
  uint8_t label[32] = {};
  uint64_t colour[32] = {};           // refined colours (isomorphism invariant)
  std::vector<uint64_t> key;          // n, m, sorted colours
};

QueryGraph load(const bms2_host_csr* q, uint32_t g) {
  QueryGraph G;
  const uint32_t b = q->graph_offsets[g];
  G.n = q->graph_offsets[g + 1] - b;
  if (G.n > 32) {
    G.simple = false;
    return G;
  }
  uint64_t entries = 0;
  for (uint32_t u = 0; u < G.n; ++u) {
    const uint32_t r0 = q->row_offsets[b + u], r1 = q->row_offsets[b + u + 1];
    entries += r1 - r0;
    for (uint32_t k = r0; k < r1; ++k) {
      const uint32_t v = q->column_indices[k] - b;
      if (v >= G.n || v == u || ((G.adj[u] >> v) & 1u)) {
        G.simple = false;
        continue;
      }
      G.adj[u] |= 1u << v;
    }
  }
  G.m = uint32_t(entries / 2);
  // colour refinement: start from (label, degree), then n rounds of (own colour, multiset of
  // neighbour colours) hashed with two independent commutative sums
  for (uint32_t u = 0; u < G.n; ++u) G.label[u] = q->node_labels[b + u];
  // initial colour: label, degree and the label counts at each BFS distance (an isomorphism
  // invariant that splits most ties colour refinement leaves, e.g. between regular graphs)
  for (uint32_t u = 0; u < G.n; ++u) {
    uint64_t c = mix64((uint64_t(q->node_labels[b + u]) << 8) | __builtin_popcount(G.adj[u]));
    uint32_t seen = 1u << u, front = seen;
    for (uint32_t dist = 1; front; ++dist) {
      uint32_t next = 0;
      for (uint32_t f = front; f; f &= f - 1) next |= G.adj[__builtin_ctz(f)];
      front = next & ~seen;
      seen |= front;
      uint64_t shell = 0;  // commutative over the shell's labels
      for (uint32_t f = front; f; f &= f - 1) shell += mix64(0x100u + G.label[__builtin_ctz(f)]);
      c = mix64(c ^ mix64(shell + dist));
    }
    G.colour[u] = c;
  }
  uint64_t next[32];
  for (uint32_t round = 0; round < G.n; ++round) {
    for (uint32_t u = 0; u < G.n; ++u) {
      uint64_t s1 = 0, s2 = 0;
      for (uint32_t a = G.adj[u]; a; a &= a - 1) {
        const uint64_t c = G.colour[__builtin_ctz(a)];
        s1 += mix64(c);
        s2 += mix64(c ^ 0x5bd1e9955bd1e995ull);
      }
      next[u] = mix64(G.colour[u] ^ mix64(s1 + 0x632be59bd9b4e019ull * s2));
    }
    std::memcpy(G.colour, next, sizeof(uint64_t) * G.n);
  }
  G.key.reserve(G.n + 2);
  G.key.push_back(G.n);
  G.key.push_back(G.m);
  for (uint32_t u = 0; u < G.n; ++u) G.key.push_back(G.colour[u]);
  std::sort(G.key.begin() + 2, G.key.end());
  return G;
}

// Backtracking search for a label-, colour- and adjacency-preserving bijection A -> B (edges and
// non-edges both preserved, as induced matching needs).  With fix_a >= 0 the bijection must map
// node fix_a of A to node fix_b of B.  On success map[u] is the image of A's node u.  False when
// none exists or the budget runs out.
bool find_isomorphism(const QueryGraph& A, const QueryGraph& B, int fix_a, int fix_b,
                      uint32_t* map, uint64_t* spent = nullptr) {
  const uint32_t n = A.n;
  if (n != B.n || A.m != B.m || A.key != B.key) return false;
  if (n == 0) return true;
  // A's nodes in search order: fix_a (else the rarest colour) first, then always a node with the
  // most neighbours already placed (ties: rarer colour, lower id)
  auto multiplicity = [&](uint64_t c) {
    uint32_t k = 0;
    for (uint32_t u = 0; u < n; ++u) k += A.colour[u] == c;
    return k;
  };
  uint32_t order[32], placed = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t best = 0, best_links = 0, best_mult = ~0u;
    bool have = false;
    if (i == 0 && fix_a >= 0) {
      best = uint32_t(fix_a);
      have = true;
    }
    for (uint32_t u = 0; u < n && !(i == 0 && fix_a >= 0); ++u) {
      if ((placed >> u) & 1u) continue;
      const uint32_t links = __builtin_popcount(A.adj[u] & placed), mult = multiplicity(A.colour[u]);
      if (!have || links > best_links || (links == best_links && mult < best_mult)) {
        best = u;
        best_links = links;
        best_mult = mult;
        have = true;
      }
    }
    order[i] = best;
    placed |= 1u << best;
  }
  uint32_t image[32];  // image[i] = node of B matched to order[i]
  uint32_t next_try[32];
  uint32_t used = 0;
  uint64_t steps = 0;
  struct Tally {  // adds the steps taken to *spent on every exit
    uint64_t* spent;
    const uint64_t& steps;
    ~Tally() {
      if (spent) *spent += steps;
    }
  } tally{spent, steps};
  int i = 0;
  next_try[0] = fix_a >= 0 ? uint32_t(fix_b) : 0;
  while (i >= 0) {
    if (i == int(n)) {
      for (uint32_t j = 0; j < n; ++j) map[order[j]] = image[j];
      return true;
    }
    const uint32_t a = order[i];
    const uint32_t last = (i == 0 && fix_a >= 0) ? uint32_t(fix_b) + 1 : n;
    bool advanced = false;
    for (uint32_t b = next_try[i]; b < last; ++b) {
      if (++steps > kSearchBudget) return false;
      if (((used >> b) & 1u) || B.label[b] != A.label[a] || B.colour[b] != A.colour[a]) continue;
      bool ok = true;
      for (int j = 0; j < i && ok; ++j)
        ok = (((A.adj[a] >> order[j]) & 1u) == ((B.adj[b] >> image[j]) & 1u));
      if (!ok) continue;
      image[i] = b;
      used |= 1u << b;
      next_try[i] = b + 1;
      ++i;
      if (i < int(n)) next_try[i] = 0;
      advanced = true;
      break;
    }
    if (!advanced) {
      --i;
      if (i >= 0) used &= ~(1u << image[i]);
    }
  }
  return false;
}

bool isomorphic(const QueryGraph& A, const QueryGraph& B) {
  uint32_t map[32];
  return find_isomorphism(A, B, -1, -1, map);
}

}  // namespace

namespace bms2 {

// rep[g] = the first query (lowest index) of g's isomorphism class; returns the number of
// queries with rep[g] != g.
uint32_t query_classes(const bms2_host_csr* q, uint32_t* rep) {
  std::vector<QueryGraph> G(q->num_graphs);
  std::map<std::vector<uint64_t>, std::vector<uint32_t>> buckets;  // key -> first query per class
  uint32_t dups = 0;
  for (uint32_t g = 0; g < q->num_graphs; ++g) {
    G[g] = load(q, g);
    rep[g] = g;
    if (!G[g].simple || G[g].n < 2) continue;
    std::vector<uint32_t>& reps = buckets[G[g].key];
    for (size_t i = 0; i < reps.size() && i < kBucketCompares; ++i)
      if (isomorphic(G[reps[i]], G[g])) {
        rep[g] = reps[i];
        ++dups;
        break;
      }
    if (rep[g] == g) reps.push_back(g);
  }
  return dups;
}

// orbit[u] = the lowest local node of u's orbit under the label-preserving automorphisms of query
// g (every node its own orbit for a query that is not simple or has more than 32 nodes, or when a
// search runs over budget).  Automorphic nodes have equal filter and refinement fields, hence
// equal domains in every data graph.
void query_orbits(const bms2_host_csr* q, uint32_t g, uint8_t* orbit) {
  const QueryGraph G = load(q, g);
  const uint32_t n = q->graph_offsets[g + 1] - q->graph_offsets[g];
  for (uint32_t u = 0; u < n; ++u) orbit[u] = uint8_t(u);
  if (!G.simple || n < 2) return;
  uint32_t parent[32];
  for (uint32_t u = 0; u < n; ++u) parent[u] = u;
  auto root = [&](uint32_t u) {
    while (parent[u] != u) u = parent[u] = parent[parent[u]];
    return u;
  };
  uint32_t map[32];
  uint64_t spent = 0;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t v = u + 1; v < n; ++v) {
      if (G.colour[u] != G.colour[v] || G.label[u] != G.label[v] || root(u) == root(v)) continue;
      if (spent >= kOrbitBudget) break;
      if (!find_isomorphism(G, G, int(u), int(v), map, &spent)) continue;
      for (uint32_t w = 0; w < n; ++w) {  // every pair an automorphism relates shares an orbit
        const uint32_t a = root(w), b = root(map[w]);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
      }
    }
  for (uint32_t u = 0; u < n; ++u) {
    uint32_t r = root(u);
    orbit[u] = uint8_t(r);
  }
}

// Symmetry breaking for query g under the matching order `order` (order[d] = local node matched
// at depth d): among the automorphism orbits of two or more nodes, the one whose second-matched
// member comes earliest; du = the depth of its first-matched member, vmask = the depths of the
// others.  Every induced embedding f of g (domains of automorphic nodes are equal) has exactly
// one image f o a under the automorphisms a whose value at du is the orbit's minimum (the group
// maps u_du onto each orbit member equally often), so the embeddings with f(d) > f(du) for every
// d in vmask number exactly 1 / (|orbit|) of all: the join counts those and multiplies by
// popcount(vmask) + 1.  du = 0xff, vmask = 0 without such an orbit.
void query_symmetry(const bms2_host_csr* q, uint32_t g, const uint8_t* order, const uint8_t* orbit,
                    uint8_t* du, uint32_t* vmask) {
  const uint32_t n = q->graph_offsets[g + 1] - q->graph_offsets[g];
  *du = 0xff;
  *vmask = 0;
  if (n < 2 || n > 32) return;
  uint32_t depth_of[32];
  for (uint32_t d = 0; d < n; ++d) depth_of[order ? order[d] : d] = d;
  uint32_t best_second = ~0u;
  for (uint32_t r = 0; r < n; ++r) {
    if (orbit[r] != r) continue;  // r is the lowest node of its orbit
    uint32_t depths = 0;
    for (uint32_t u = 0; u < n; ++u)
      if (orbit[u] == r) depths |= 1u << depth_of[u];
    if (__builtin_popcount(depths) < 2) continue;
    const uint32_t first = __builtin_ctz(depths), rest = depths & (depths - 1);
    const uint32_t second = __builtin_ctz(rest);
    if (second < best_second) {
      best_second = second;
      *du = uint8_t(first);
      *vmask = rest;
    }
  }
}

}  // namespace bms2

extern "C" int bms2_query_symmetry(const bms2_host_csr* q, const uint8_t* order, uint8_t* du,
                                   uint32_t* vmask) {
  if (!q || !du || !vmask ||
      (q->num_graphs && (!q->graph_offsets || !q->row_offsets || !q->node_labels)))
    return bms2::set_error(BMS2_ERR_INVALID, "null query batch or output");
  std::vector<uint8_t> orbit(std::max<uint64_t>(q->total_nodes, 1));
  for (uint32_t g = 0; g < q->num_graphs; ++g) {
    const uint32_t b = q->graph_offsets[g], n = q->graph_offsets[g + 1] - b;
    if (order) {
      uint32_t seen = 0;
      for (uint32_t d = 0; d < n && n <= 32; ++d) {
        const uint32_t u = order[size_t(g) * 32 + d];
        if (u >= n || ((seen >> u) & 1))
          return bms2::set_error(BMS2_ERR_INVALID, "order of query %u is not a permutation", g);
        seen |= 1u << u;
      }
    }
    bms2::query_orbits(q, g, orbit.data() + b);
    bms2::query_symmetry(q, g, order ? order + size_t(g) * 32 : nullptr, orbit.data() + b, du + g,
                         vmask + g);
  }
  return BMS2_OK;
}

extern "C" int bms2_query_classes(const bms2_host_csr* q, uint32_t* rep, uint32_t* num_dups) {
  if (!q || !rep || (q->num_graphs && (!q->graph_offsets || !q->row_offsets || !q->node_labels)))
    return bms2::set_error(BMS2_ERR_INVALID, "null query batch or output");
  const uint32_t d = bms2::query_classes(q, rep);
  if (num_dups) *num_dups = d;
  return BMS2_OK;
}

extern "C" int bms2_query_orbits(const bms2_host_csr* q, uint8_t* orbit) {
  if (!q || !orbit || (q->num_graphs && (!q->graph_offsets || !q->row_offsets || !q->node_labels)))
    return bms2::set_error(BMS2_ERR_INVALID, "null query batch or output");
  for (uint32_t g = 0; g < q->num_graphs; ++g)
    bms2::query_orbits(q, g, orbit + q->graph_offsets[g]);
  return BMS2_OK;
}
