// Query-side inputs and outputs of the matching-order network, on the host: the feature columns
// that depend on a query's structure only, and connected matching orders from per-node scores.
// Queries have at most 32 nodes, so a node's neighbours fit one u32 mask.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace {

constexpr int kFeatures = 32;
constexpr uint32_t kFar = 99;  // distance of an unreachable node
// element codes of the project generator (workload.cpp): N, Cl, Br, O, C, S, F
constexpr int kElements[7] = {0, 1, 3, 7, 8, 9, 10};

int check_queries(const bms2_host_csr* q) {
  if (!q || (q->num_graphs && (!q->graph_offsets || !q->row_offsets || !q->node_labels)))
    return bms2::set_error(BMS2_ERR_INVALID, "null query batch");
  for (uint32_t g = 0; g < q->num_graphs; ++g)
    if (q->graph_offsets[g + 1] - q->graph_offsets[g] > 32)
      return bms2::set_error(BMS2_ERR_UNSUPPORTED, "query %u has more than 32 nodes", g);
  return BMS2_OK;
}

// Neighbour masks of query g (local ids; parallel edges collapse, a self-loop sets the node's
// own bit).
int neighbour_masks(const bms2_host_csr* q, uint32_t g, uint32_t* adj) {
  const uint32_t b = q->graph_offsets[g], n = q->graph_offsets[g + 1] - b;
  for (uint32_t u = 0; u < n; ++u) {
    adj[u] = 0;
    for (uint32_t k = q->row_offsets[b + u]; k < q->row_offsets[b + u + 1]; ++k) {
      const uint32_t v = q->column_indices[k] - b;
      if (v >= n) return bms2::set_error(BMS2_ERR_INVALID, "query %u: neighbour outside its graph", g);
      adj[u] |= 1u << v;
    }
  }
  return BMS2_OK;
}

// Nodes reachable from s over adj (s included).
uint32_t closure(const uint32_t* adj, uint32_t s) {
  uint32_t seen = 1u << s, front = seen;
  while (front) {
    uint32_t next = 0;
    for (uint32_t f = front; f; f &= f - 1) next |= adj[__builtin_ctz(f)];
    front = next & ~seen;
    seen |= next;
  }
  return seen;
}

}  // namespace

extern "C" {

// Columns 0-5, 8-10, 13-25, 28 and 30 of every query node's row (rows in global node order,
// kFeatures floats each; the other columns are set to 0): query size, edges, cycle rank,
// largest degree and diameter, distinct labels; per node degree, eccentricity, closeness, leaf
// neighbours, summed neighbour degrees, ring membership (an edge that is not a bridge), wildcard
// and element one-hot, position, degree rank (descending, ties by id) and the largest-degree
// flag.  Values are computed in double and rounded once, as the numpy statement of the same
// columns does.  No counterpart in the reference.
int bms2_query_static_features(const bms2_host_csr* q, float* out) {
  if (int rc = check_queries(q)) return rc;
  if (!out) return bms2::set_error(BMS2_ERR_INVALID, "null output");
  std::memset(out, 0, sizeof(float) * kFeatures * q->total_nodes);
  uint32_t adj[32];
  uint32_t dist[32][32];
  for (uint32_t g = 0; g < q->num_graphs; ++g) {
    const uint32_t b = q->graph_offsets[g], n = q->graph_offsets[g + 1] - b;
    if (!n) continue;
    if (int rc = neighbour_masks(q, g, adj)) return rc;
    uint32_t deg[32], degsum = 0, maxdeg = 0, argmax = 0;
    for (uint32_t u = 0; u < n; ++u) {
      deg[u] = uint32_t(__builtin_popcount(adj[u]));
      degsum += deg[u];
      if (deg[u] > maxdeg) {
        maxdeg = deg[u];
        argmax = u;
      }
    }
    // breadth-first distances from every node
    for (uint32_t s = 0; s < n; ++s) {
      for (uint32_t v = 0; v < n; ++v) dist[s][v] = kFar;
      dist[s][s] = 0;
      uint32_t seen = 1u << s, front = seen;
      for (uint32_t d = 1; front; ++d) {
        uint32_t next = 0;
        for (uint32_t f = front; f; f &= f - 1) next |= adj[__builtin_ctz(f)];
        front = next & ~seen;
        seen |= next;
        for (uint32_t f = front; f; f &= f - 1) dist[s][__builtin_ctz(f)] = d;
      }
    }
    const uint32_t m = degsum / 2;
    uint32_t ecc[32], diam = 0;
    double close[32];
    for (uint32_t u = 0; u < n; ++u) {
      uint32_t e = 0;
      uint64_t sum = 0;
      for (uint32_t v = 0; v < n; ++v) {
        const uint32_t d = dist[u][v] < kFar ? dist[u][v] : 0;
        e = std::max(e, d);
        sum += d;
      }
      ecc[u] = e;
      diam = std::max(diam, e);
      close[u] = double(sum) / double(std::max<uint32_t>(n - 1, 1)) / 4.0;
    }
    uint32_t labmask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t u = 0; u < n; ++u) {
      const uint8_t l = q->node_labels[b + u];
      labmask[l >> 5] |= 1u << (l & 31);
    }
    uint32_t distinct = 0;
    for (uint32_t w = 0; w < 8; ++w) distinct += uint32_t(__builtin_popcount(labmask[w]));
    // rank of each node in a stable sort by descending degree
    uint32_t rank[32];
    for (uint32_t u = 0; u < n; ++u) {
      uint32_t r = 0;
      for (uint32_t v = 0; v < n; ++v) r += deg[v] > deg[u] || (deg[v] == deg[u] && v < u);
      rank[u] = r;
    }
    const double span = double(std::max<uint32_t>(n - 1, 1));
    for (uint32_t u = 0; u < n; ++u) {
      float* x = out + uint64_t(b + u) * kFeatures;
      x[0] = float(double(n) / 16.0);
      x[1] = float(double(m) / 16.0);
      x[2] = float((double(m) - double(n) + 1.0) / 4.0);
      x[3] = float(double(maxdeg) / 4.0);
      x[4] = float(double(diam) / 8.0);
      x[5] = float(double(distinct) / 8.0);
      x[8] = float(double(deg[u]) / 4.0);
      x[9] = float(double(ecc[u]) / 8.0);
      x[10] = float(close[u]);
      uint32_t leaves = 0, nsum = 0;
      for (uint32_t f = adj[u]; f; f &= f - 1) {
        const uint32_t v = uint32_t(__builtin_ctz(f));
        leaves += deg[v] == 1;
        nsum += deg[v];
      }
      x[13] = float(double(leaves) / 4.0);
      x[14] = float(double(nsum) / 16.0);
      // ring membership: some edge u-v leaves v reachable from u once it is removed
      bool ring = false;
      for (uint32_t f = adj[u]; f && !ring; f &= f - 1) {
        const uint32_t v = uint32_t(__builtin_ctz(f));
        uint32_t cut[32];
        std::memcpy(cut, adj, sizeof(uint32_t) * n);
        cut[u] &= ~(1u << v);
        cut[v] &= ~(1u << u);
        ring = (closure(cut, u) >> v) & 1;
      }
      x[15] = ring ? 1.0f : 0.0f;
      const uint8_t lab = q->node_labels[b + u];
      x[16] = lab == 255 ? 1.0f : 0.0f;
      int elem = -1;
      for (int k = 0; k < 7; ++k)
        if (kElements[k] == lab) elem = k;
      if (elem >= 0)
        x[17 + elem] = 1.0f;
      else if (lab != 255)
        x[24] = 1.0f;
      x[25] = float(double(u) / span);
      x[28] = float(double(rank[u]) / span);
      x[30] = u == argmax ? 1.0f : 0.0f;
    }
  }
  return BMS2_OK;
}

// Matching order of every query (order[32 g + d] = local node matched at depth d; positions past
// the query's nodes are left as they are): for a query with use[g] != 0 the connected order by
// descending score -- the best-scored node first, then always the best-scored node adjacent to
// the matched set (any remaining node when none is), ties to the lower id; identity otherwise.
// scores are per query node in global node order.  No counterpart in the reference.
int bms2_score_orders(const bms2_host_csr* q, const float* scores, const uint8_t* use,
                      uint8_t* order) {
  if (int rc = check_queries(q)) return rc;
  if (!order || (q->total_nodes && !scores)) return bms2::set_error(BMS2_ERR_INVALID, "null argument");
  uint32_t adj[32];
  for (uint32_t g = 0; g < q->num_graphs; ++g) {
    const uint32_t b = q->graph_offsets[g], n = q->graph_offsets[g + 1] - b;
    uint8_t* o = order + uint64_t(g) * 32;
    if (!n) continue;
    if (!use || !use[g]) {
      for (uint32_t d = 0; d < n; ++d) o[d] = uint8_t(d);
      continue;
    }
    if (int rc = neighbour_masks(q, g, adj)) return rc;
    const float* sc = scores + b;
    const uint32_t all = n == 32 ? ~0u : (1u << n) - 1;
    uint32_t S = 0, near = 0;
    for (uint32_t d = 0; d < n; ++d) {
      uint32_t pool = near & ~S;
      if (!pool) pool = all & ~S;
      uint32_t best = 32;
      for (uint32_t f = pool; f; f &= f - 1) {
        const uint32_t v = uint32_t(__builtin_ctz(f));
        if (best == 32 || sc[v] > sc[best]) best = v;  // ascending ids: ties keep the lower
      }
      o[d] = uint8_t(best);
      S |= 1u << best;
      near |= adj[best];
    }
  }
  return BMS2_OK;
}

}  // extern "C"
