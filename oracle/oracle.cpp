// CPU restatement of the reference's batched subgraph-isomorphism path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU baseline.  The
// product never links or calls this file.  See oracle.h for the decisions taken where the
// reference is undefined; each function below cites the reference lines it follows.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint8_t kWildcard = 255;  // types.hpp:18

// SignatureDevice::getLabelCount / incrementLabelCount (signature.hpp:38-50): 16 labels of 4 bits,
// labels >= 16 ignored, counts saturate at 15.
inline unsigned label_count(uint64_t sig, unsigned l) { return unsigned(sig >> (4 * l)) & 15u; }
inline void inc_label(uint64_t& sig, uint8_t label) {
  if (label < 16) {
    uint64_t c = (sig >> (4 * label)) & 15u;
    if (c < 15) {
      sig &= ~(uint64_t(15) << (4 * label));
      sig |= (c + 1) << (4 * label);
    }
  }
}

// Exact bitset over the local node ids of one graph (up to 256 nodes).
struct Bits256 {
  uint64_t w[4] = {0, 0, 0, 0};
  void set(unsigned i) { w[i >> 6] |= 1ull << (i & 63); }
  void unset(unsigned i) { w[i >> 6] &= ~(1ull << (i & 63)); }
  bool get(unsigned i) const { return (w[i >> 6] >> (i & 63)) & 1; }
  bool empty() const { return !(w[0] | w[1] | w[2] | w[3]); }
};

struct GraphRange {
  uint32_t begin, end;  // global node ids [begin, end)
};

inline GraphRange graph_range(const orc_csr* g, uint32_t gi) {
  return {g->graph_offsets[gi], g->graph_offsets[gi + 1]};
}

// Path Bloom insert (PathSignatureDevice::insert, signature.hpp:418-430): bits h, fmix(h),
// fmix(fmix(h)), each mod 256.
inline void bloom_insert(uint64_t* bits, uint32_t h1) {
  uint32_t h2 = orc_hash32(h1);
  uint32_t h3 = orc_hash32(h2);
  uint32_t i1 = h1 % 256, i2 = h2 % 256, i3 = h3 % 256;
  bits[i1 / 64] |= 1ull << (i1 % 64);
  bits[i2 / 64] |= 1ull << (i2 % 64);
  bits[i3 / 64] |= 1ull << (i3 % 64);
}

// Path signature of one node (signature.hpp:482-514).
void path_signature_node(const orc_csr* g, uint32_t u, uint64_t* bits) {
  bits[0] = bits[1] = bits[2] = bits[3] = 0;
  const uint32_t* ro = g->row_offsets;
  const uint32_t* ci = g->column_indices;
  const uint8_t* lab = g->node_labels;
  const uint32_t lu = lab[u];
  for (uint32_t i = ro[u]; i < ro[u + 1]; ++i) {
    const uint32_t v = ci[i];
    const uint32_t lv = lab[v];
    for (uint32_t j = ro[v]; j < ro[v + 1]; ++j) {
      const uint32_t w = ci[j];
      if (w == u) continue;
      const uint32_t lw = lab[w];
      const uint32_t p2 = orc_combine_hash(lu, orc_combine_hash(lv, lw));
      bloom_insert(bits, orc_hash32(p2));
      for (uint32_t k = ro[w]; k < ro[w + 1]; ++k) {
        const uint32_t x = ci[k];
        if (x == v || x == u) continue;
        const uint32_t p3 = orc_combine_hash(p2, lab[x]);
        bloom_insert(bits, orc_hash32(p3));
      }
    }
  }
}

// Cycle signature of one node: the literal iterative DFS of signature.hpp:602-715, including the
// 128-entry stack whose overflowing pushes are dropped (:638, :697), the u16 counters and the
// skip of graphs above 64 nodes (:614, where the value stays at its zero fill).
uint64_t cycle_signature_node(const orc_csr* g, uint32_t u, GraphRange r) {
  const uint32_t n = r.end - r.begin;
  if (n > 64) return 0;
  const uint32_t* ro = g->row_offsets;
  const uint32_t* ci = g->column_indices;
  const uint8_t* lab = g->node_labels;
  const uint8_t lu = lab[u];
  const uint32_t ul = u - r.begin;
  struct Frame {
    uint32_t node, depth, all_same;
  };
  Frame stack[128];
  uint64_t masks[128];
  int top = 0;
  uint16_t c5s = 0, c5d = 0, c6s = 0, c6d = 0;
  for (uint32_t i = ro[u]; i < ro[u + 1]; ++i) {
    const uint32_t nb = ci[i];
    const uint32_t nl = nb - r.begin;
    if (top < 128) {
      masks[top] = (1ull << ul) | (1ull << nl);
      stack[top] = {nl, 1, lab[nb] == lu ? 1u : 0u};
      ++top;
    }
  }
  while (top > 0) {
    --top;
    const Frame f = stack[top];
    const uint64_t mask = masks[top];
    const uint32_t cur = f.node + r.begin;
    if (f.depth == 4 || f.depth == 5) {
      for (uint32_t i = ro[cur]; i < ro[cur + 1]; ++i)
        if (ci[i] == u) {
          if (f.depth == 4)
            f.all_same ? ++c5s : ++c5d;
          else
            f.all_same ? ++c6s : ++c6d;
        }
    }
    if (f.depth < 5) {
      for (uint32_t i = ro[cur]; i < ro[cur + 1]; ++i) {
        const uint32_t nb = ci[i];
        if (nb == u) continue;
        const uint32_t nl = nb - r.begin;
        if (!(mask & (1ull << nl))) {
          if (top < 128) {
            masks[top] = mask | (1ull << nl);
            stack[top] = {nl, f.depth + 1, (f.all_same && lab[nb] == lu) ? 1u : 0u};
            ++top;
          }
        }
      }
    }
  }
  return uint64_t(uint16_t(c5s / 2)) | (uint64_t(uint16_t(c5d / 2)) << 16) |
         (uint64_t(uint16_t(c6s / 2)) << 32) | (uint64_t(uint16_t(c6d / 2)) << 48);
}

// CycleSignatureDevice::contains (signature.hpp:551-554): data >= query in every field.
inline bool cycle_contains(uint64_t d, uint64_t q) {
  for (int f = 0; f < 4; ++f)
    if (((d >> (16 * f)) & 0xffff) < ((q >> (16 * f)) & 0xffff)) return false;
  return true;
}

// PathSignatureDevice::contains (signature.hpp:432-437).
inline bool path_contains(const uint64_t* d, const uint64_t* q) {
  for (int i = 0; i < 4; ++i)
    if ((d[i] & q[i]) != q[i]) return false;
  return true;
}

// Label-count dominance of refineCandidates (isomorphism.hpp:172-179).
inline bool hist_dominates(uint64_t d, uint64_t q) {
  for (unsigned l = 0; l < 16; ++l)
    if (label_count(q, l) > label_count(d, l)) return false;
  return true;
}

inline uint32_t graph_of_node(const orc_csr* g, uint32_t node) {
  // getGraphID (graph.hpp:96-107)
  uint32_t lo = 0, hi = g->num_graphs;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (node >= g->graph_offsets[mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

inline bool is_neighbor(const orc_csr* g, uint32_t a, uint32_t b) {
  // DeviceBatchedCSRGraph::isNeighbor (graph.hpp:68-73)
  for (uint32_t i = g->row_offsets[a]; i < g->row_offsets[a + 1]; ++i)
    if (g->column_indices[i] == b) return true;
  return false;
}

}  // namespace

extern "C" {

uint32_t orc_hash32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

uint32_t orc_combine_hash(uint32_t h1, uint32_t h2) {
  return h1 ^ (h2 + 0x9e3779b9u + (h1 << 6) + (h1 >> 2));
}

void orc_label_signatures(const orc_csr* g, uint64_t* sig) {
  for (uint64_t u = 0; u < g->total_nodes; ++u) {
    uint64_t s = sig[u];  // increments on the existing value (signature.hpp:142-145)
    for (uint32_t i = g->row_offsets[u]; i < g->row_offsets[u + 1]; ++i)
      inc_label(s, g->node_labels[g->column_indices[i]]);
    sig[u] = s;
  }
}

void orc_path_signatures(const orc_csr* g, uint64_t* out) {
  for (uint64_t u = 0; u < g->total_nodes; ++u) path_signature_node(g, uint32_t(u), out + 4 * u);
}

void orc_cycle_signatures(const orc_csr* g, uint64_t* out) {
  for (uint32_t gi = 0; gi < g->num_graphs; ++gi) {
    GraphRange r = graph_range(g, gi);
    for (uint32_t u = r.begin; u < r.end; ++u) out[u] = cycle_signature_node(g, u, r);
  }
}

void orc_refine_signatures(const orc_csr* g, uint32_t view_size, uint64_t* sig, uint64_t* reach) {
  // refineCSRSignatures<PowerGraph> (signature.hpp:347-385) with exact bitsets.
  for (uint32_t gi = 0; gi < g->num_graphs; ++gi) {
    GraphRange r = graph_range(g, gi);
    for (uint32_t u = r.begin; u < r.end; ++u) {
      const uint32_t ul = u - r.begin;
      Bits256 reachable, frontier, old;
      std::memcpy(old.w, reach + 4 * uint64_t(u), sizeof(old.w));
      frontier.set(ul);
      reachable.set(ul);
      for (uint32_t it = 0; it < view_size && !frontier.empty(); ++it) {
        Bits256 next;
        for (uint32_t x = 0; x < r.end - r.begin; ++x) {
          if (!frontier.get(x)) continue;
          const uint32_t xg = x + r.begin;
          for (uint32_t i = g->row_offsets[xg]; i < g->row_offsets[xg + 1]; ++i) {
            const uint32_t y = g->column_indices[i] - r.begin;
            if (!reachable.get(y)) {
              reachable.set(y);
              next.set(y);
            }
          }
        }
        frontier = next;
      }
      reachable.unset(ul);
      for (int w = 0; w < 4; ++w) reachable.w[w] &= ~old.w[w];
      uint64_t s = 0;
      for (uint32_t x = 0; x < r.end - r.begin; ++x)
        if (reachable.get(x)) inc_label(s, g->node_labels[x + r.begin]);
      sig[u] = s;
      for (int w = 0; w < 4; ++w) reach[4 * uint64_t(u) + w] = reachable.w[w] | old.w[w];
    }
  }
}

void orc_filter(const orc_csr* q, const orc_csr* d, const uint64_t* qpath, const uint64_t* dpath,
                const uint64_t* qcycle, const uint64_t* dcycle, uint64_t* cand, uint64_t W) {
  // FilterCandidatesKernel_v2 (isomorphism.hpp:59-97)
  for (uint64_t dn = 0; dn < d->total_nodes; ++dn) {
    const uint32_t ddeg = d->row_offsets[dn + 1] - d->row_offsets[dn];
    for (uint64_t qn = 0; qn < q->total_nodes; ++qn) {
      if (q->node_labels[qn] != d->node_labels[dn] && q->node_labels[qn] != kWildcard) continue;
      const uint32_t qdeg = q->row_offsets[qn + 1] - q->row_offsets[qn];
      if (qdeg > ddeg) continue;
      if (!path_contains(dpath + 4 * dn, qpath + 4 * qn)) continue;
      if (!cycle_contains(dcycle[dn], qcycle[qn])) continue;
      cand[qn * W + dn / 64] |= 1ull << (dn % 64);
    }
  }
}

void orc_refine_candidates(const orc_csr* q, const orc_csr* d, const uint64_t* qsig,
                           const uint64_t* dsig, uint64_t* cand, uint64_t W) {
  // RefineCandidatesKernel_v3 (isomorphism.hpp:142-196)
  for (uint64_t qn = 0; qn < q->total_nodes; ++qn) {
    for (uint64_t w = 0; w < W; ++w) {
      uint64_t block = cand[qn * W + w];
      uint64_t bits = block;
      while (bits) {
        const unsigned b = unsigned(__builtin_ctzll(bits));
        bits &= bits - 1;
        const uint64_t dn = w * 64 + b;
        if (dn >= d->total_nodes || !hist_dominates(dsig[dn], qsig[qn])) block &= ~(1ull << b);
      }
      cand[qn * W + w] = block;
    }
  }
}

uint64_t orc_gmcr(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                  uint32_t* offsets, uint32_t* indices, uint64_t cap) {
  // generateGMCR (gmcr.hpp:54-127): valid iff n_q > 1 and every query node keeps a candidate
  // inside the data graph's node range; segments in ascending query id.
  uint64_t total = 0;
  offsets[0] = 0;
  for (uint32_t g = 0; g < d->num_graphs; ++g) {
    GraphRange r = graph_range(d, g);
    for (uint32_t qi = 0; qi < q->num_graphs; ++qi) {
      GraphRange qr = graph_range(q, qi);
      if (qr.end - qr.begin <= 1) continue;
      bool ok = true;
      for (uint32_t qn = qr.begin; qn < qr.end && ok; ++qn) {
        bool any = false;
        for (uint32_t dn = r.begin; dn < r.end && !any; ++dn)
          any = (cand[qn * W + dn / 64] >> (dn % 64)) & 1;
        ok = any;
      }
      if (ok) {
        if (total < cap) indices[total] = qi;
        ++total;
      }
    }
    offsets[g + 1] = uint32_t(total);
  }
  return total;
}

uint64_t orc_join(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                  const uint32_t* offsets, const uint32_t* indices, int find_first,
                  const uint32_t* order, uint64_t* counts, uint64_t* dfs_nodes) {
  return orc_join_sym(q, d, cand, W, offsets, indices, find_first, order, nullptr, nullptr, counts,
                      dfs_nodes);
}

uint64_t orc_join_sym(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                      const uint32_t* offsets, const uint32_t* indices, int find_first,
                      const uint32_t* order, const uint8_t* sym_du, const uint32_t* sym_vmask,
                      uint64_t* counts, uint64_t* dfs_nodes) {
  // JoinCandidates3Optimized (isomorphism.hpp:309-468) with standard backtracking: candidates of
  // each depth in ascending data node id (getCandidateAt, candidates.hpp:128-158), injectivity by
  // a visited set, induced adjacency test of isValidMapping_fast_bitwise (:268-288).  With
  // sym_du / sym_vmask (per query; not a reference behaviour) a candidate at a depth of
  // sym_vmask must exceed the data node matched at depth sym_du, and the count is NOT multiplied.
  uint64_t total = 0;
  for (uint32_t g = 0; g < d->num_graphs; ++g) {
    GraphRange r = graph_range(d, g);
    for (uint32_t t = offsets[g]; t < offsets[g + 1]; ++t) {
      const uint32_t qi = indices[t];
      GraphRange qr = graph_range(q, qi);
      const uint32_t nq = qr.end - qr.begin;
      std::vector<uint32_t> ord(nq);
      for (uint32_t i = 0; i < nq; ++i) ord[i] = order ? order[qr.begin + i] : i;
      std::vector<std::vector<uint32_t>> lists(nq);
      for (uint32_t i = 0; i < nq; ++i)
        for (uint32_t dn = r.begin; dn < r.end; ++dn)
          if ((cand[uint64_t(qr.begin + ord[i]) * W + dn / 64] >> (dn % 64)) & 1)
            lists[i].push_back(dn);
      // qadj[i][j]: query nodes at depths i and j adjacent
      std::vector<std::vector<char>> qadj(nq, std::vector<char>(nq, 0));
      for (uint32_t i = 0; i < nq; ++i)
        for (uint32_t j = 0; j < nq; ++j)
          qadj[i][j] = is_neighbor(q, qr.begin + ord[i], qr.begin + ord[j]);
      std::vector<uint32_t> mapping(nq);
      std::vector<char> visited(r.end - r.begin, 0);
      uint64_t count = 0, nodes = 0;
      bool stop = false;
      // explicit stack of candidate cursors (one per depth)
      std::vector<uint32_t> cursor(nq, 0);
      int depth = 0;
      cursor[0] = 0;
      while (depth >= 0 && !stop) {
        if (cursor[depth] >= lists[depth].size()) {
          --depth;
          if (depth >= 0) visited[mapping[depth] - r.begin] = 0;
          continue;
        }
        const uint32_t c = lists[depth][cursor[depth]++];
        if (visited[c - r.begin]) continue;
        bool ok = true;
        for (int i = 0; i < depth && ok; ++i) ok = (qadj[i][depth] != 0) == is_neighbor(d, mapping[i], c);
        if (ok && sym_du && sym_du[qi] != 0xff && ((sym_vmask[qi] >> depth) & 1))
          ok = c > mapping[sym_du[qi]];
        if (!ok) continue;
        if (depth == int(nq) - 1) {
          ++count;
          if (find_first) stop = true;
          continue;
        }
        ++nodes;
        mapping[depth] = c;
        visited[c - r.begin] = 1;
        ++depth;
        cursor[depth] = 0;
      }
      if (find_first) count = count ? 1 : 0;
      counts[t] = count;
      if (dfs_nodes) dfs_nodes[t] = nodes;
      total += count;
    }
  }
  return total;
}

uint64_t orc_join_partial(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                          uint32_t end, int32_t* out, uint64_t cap) {
  // joinPartialCandidates2 (isomorphism.hpp:1021-1178): in every data graph, every injective map
  // of query 0's nodes 0..end-1 into the graph with each node inside its domain and the induced
  // adjacency test at every depth (isValidMapping_fast_bitwise), standard backtracking.  Record =
  // the end mapped global data node ids, then the data graph id (:983-988).  Query 0 with fewer
  // than end nodes contributes nothing (:1087).  Records in data graph order, then DFS order.
  if (q->num_graphs == 0 || end == 0) return 0;
  const GraphRange qr = graph_range(q, 0);
  if (end > qr.end - qr.begin) return 0;
  uint64_t n = 0;
  std::vector<uint32_t> mapping(end);
  for (uint32_t g = 0; g < d->num_graphs; ++g) {
    const GraphRange r = graph_range(d, g);
    std::vector<std::vector<uint32_t>> lists(end);
    for (uint32_t i = 0; i < end; ++i)
      for (uint32_t dn = r.begin; dn < r.end; ++dn)
        if ((cand[uint64_t(qr.begin + i) * W + dn / 64] >> (dn % 64)) & 1) lists[i].push_back(dn);
    std::vector<char> visited(r.end - r.begin, 0);
    std::vector<uint32_t> cursor(end, 0);
    int depth = 0;
    while (depth >= 0) {
      if (cursor[depth] >= lists[depth].size()) {
        --depth;
        if (depth >= 0) visited[mapping[depth] - r.begin] = 0;
        continue;
      }
      const uint32_t c = lists[depth][cursor[depth]++];
      if (visited[c - r.begin]) continue;
      bool ok = true;
      for (int i = 0; i < depth && ok; ++i)
        ok = is_neighbor(q, qr.begin + i, qr.begin + depth) == is_neighbor(d, mapping[i], c);
      if (!ok) continue;
      mapping[depth] = c;
      if (depth == int(end) - 1) {
        if (n < cap && out)
          for (uint32_t i = 0; i <= end; ++i)
            out[n * (end + 1) + i] = int32_t(i < end ? mapping[i] : g);
        ++n;
        continue;
      }
      visited[c - r.begin] = 1;
      ++depth;
      cursor[depth] = 0;
    }
  }
  return n;
}

uint64_t orc_join_with_partial(const orc_csr* q, const orc_csr* d, const uint64_t* cand,
                               uint64_t W, const uint32_t* offsets, const uint32_t* indices,
                               uint32_t end, const int32_t* partial, uint64_t num_partial,
                               int find_first, uint64_t* counts) {
  // joinWithPartialMatches (isomorphism.hpp:756-977): for every partial match (mapping of depths
  // 0..end-1 in data graph g) and every GMCR query q of g: n_q == end counts the partial match
  // itself (:857-868), n_q > end counts its extensions to depths end..n_q-1 with q's domains, q's
  // induced adjacency against every earlier depth (the prototype part included) and
  // injectivity (:871-948), n_q < end nothing.  Standard backtracking.  counts[t] per GMCR entry
  // t; find_first keeps 1 per entry with any match (the pair_done flags, :836-866, :952-960).
  const uint64_t T = offsets[d->num_graphs];
  for (uint64_t t = 0; t < T; ++t) counts[t] = 0;
  uint64_t total = 0;
  for (uint64_t m = 0; m < num_partial; ++m) {
    const int32_t* rec = partial + m * (end + 1);
    const uint32_t g = uint32_t(rec[end]);
    const GraphRange r = graph_range(d, g);
    for (uint32_t t = offsets[g]; t < offsets[g + 1]; ++t) {
      if (find_first && counts[t]) continue;
      const uint32_t qi = indices[t];
      const GraphRange qr = graph_range(q, qi);
      const uint32_t nq = qr.end - qr.begin;
      if (nq < end) continue;
      uint64_t c = 0;
      if (nq == end) {
        c = 1;
      } else {
        std::vector<uint32_t> mapping(nq);
        std::vector<char> visited(r.end - r.begin, 0);
        for (uint32_t i = 0; i < end; ++i) {
          mapping[i] = uint32_t(rec[i]);
          visited[mapping[i] - r.begin] = 1;
        }
        std::vector<std::vector<uint32_t>> lists(nq);
        for (uint32_t i = end; i < nq; ++i)
          for (uint32_t dn = r.begin; dn < r.end; ++dn)
            if ((cand[uint64_t(qr.begin + i) * W + dn / 64] >> (dn % 64)) & 1)
              lists[i].push_back(dn);
        std::vector<uint32_t> cursor(nq, 0);
        int depth = int(end);
        bool stop = false;
        while (depth >= int(end) && !stop) {
          if (cursor[depth] >= lists[depth].size()) {
            --depth;
            if (depth >= int(end)) visited[mapping[depth] - r.begin] = 0;
            continue;
          }
          const uint32_t cn = lists[depth][cursor[depth]++];
          if (visited[cn - r.begin]) continue;
          bool ok = true;
          for (int i = 0; i < depth && ok; ++i)
            ok = is_neighbor(q, qr.begin + i, qr.begin + depth) == is_neighbor(d, mapping[i], cn);
          if (!ok) continue;
          if (depth == int(nq) - 1) {
            ++c;
            if (find_first) stop = true;
            continue;
          }
          mapping[depth] = cn;
          visited[cn - r.begin] = 1;
          ++depth;
          cursor[depth] = 0;
        }
      }
      if (find_first) c = c ? 1 : 0;
      counts[t] += c;
      total += c;
    }
  }
  return total;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Fused per-data-graph pipeline (the timed CPU baseline).

namespace {

struct NodeTuple {
  uint32_t label;
  uint32_t deg;
  uint64_t path[4];
  uint64_t cycle;
  uint64_t hist[8];  // shells 1..K (K <= 8)
};

// Exact BFS shells of every node of one graph: shell_k histograms for k = 1..K.  Equal to
// running refineCSRSignatures with view sizes 1, 2, ..., K in order from zeroed state.
template <int W>
void shell_histograms(const orc_csr* g, GraphRange r, uint32_t K, NodeTuple* tuples) {
  const uint32_t n = r.end - r.begin;
  std::vector<std::array<uint64_t, W>> adj(n);
  for (uint32_t x = 0; x < n; ++x) {
    adj[x].fill(0);
    const uint32_t xg = x + r.begin;
    for (uint32_t i = g->row_offsets[xg]; i < g->row_offsets[xg + 1]; ++i) {
      uint32_t y = g->column_indices[i] - r.begin;
      adj[x][y >> 6] |= 1ull << (y & 63);
    }
  }
  std::array<std::array<uint64_t, W>, 16> labmask{};
  for (uint32_t x = 0; x < n; ++x) {
    uint8_t l = g->node_labels[x + r.begin];
    if (l < 16) labmask[l][x >> 6] |= 1ull << (x & 63);
  }
  for (uint32_t u = 0; u < n; ++u) {
    std::array<uint64_t, W> seen{}, frontier{};
    seen[u >> 6] |= 1ull << (u & 63);
    frontier = seen;
    for (uint32_t k = 1; k <= K; ++k) {
      std::array<uint64_t, W> next{};
      for (int w = 0; w < W; ++w) {
        uint64_t f = frontier[w];
        while (f) {
          unsigned b = unsigned(__builtin_ctzll(f));
          f &= f - 1;
          const auto& a = adj[w * 64 + b];
          for (int v = 0; v < W; ++v) next[v] |= a[v];
        }
      }
      for (int w = 0; w < W; ++w) {
        next[w] &= ~seen[w];
        seen[w] |= next[w];
      }
      uint64_t h = 0;
      for (unsigned l = 0; l < 16; ++l) {
        unsigned c = 0;
        for (int w = 0; w < W; ++w) c += unsigned(__builtin_popcountll(next[w] & labmask[l][w]));
        h |= uint64_t(c > 15 ? 15 : c) << (4 * l);
      }
      tuples[u].hist[k - 1] = h;
      frontier = next;
    }
  }
}

inline bool tuple_test(const NodeTuple& qt, const NodeTuple& dt, uint32_t K) {
  if (qt.label != dt.label && qt.label != kWildcard) return false;
  if (qt.deg > dt.deg) return false;
  if (!path_contains(dt.path, qt.path)) return false;
  if (!cycle_contains(dt.cycle, qt.cycle)) return false;
  for (uint32_t k = 0; k < K; ++k)
    if (!hist_dominates(dt.hist[k], qt.hist[k])) return false;
  return true;
}

template <int W>
void node_tuples(const orc_csr* g, uint32_t gi, uint32_t K, NodeTuple* tuples) {
  GraphRange r = graph_range(g, gi);
  for (uint32_t u = r.begin; u < r.end; ++u) {
    NodeTuple& t = tuples[u - r.begin];
    t.label = g->node_labels[u];
    t.deg = g->row_offsets[u + 1] - g->row_offsets[u];
    path_signature_node(g, u, t.path);
    t.cycle = cycle_signature_node(g, u, r);
  }
  shell_histograms<W>(g, r, K, tuples);
}

struct QueryInfo {
  uint32_t begin, n;
  std::vector<uint32_t> adjmask;  // adjmask[i] bit j: query nodes i, j adjacent
};

template <int W>
struct JoinState {
  const std::array<uint64_t, W>* dom;
  const std::array<uint64_t, W>* nadj;  // data adjacency rows
  const uint32_t* qadj;
  uint32_t nq;
  bool find_first;
  uint64_t count = 0, nodes = 0;
  bool stop = false;
  uint32_t f[32];

  void rec(uint32_t depth, const std::array<uint64_t, W>& used) {
    std::array<uint64_t, W> R;
    for (int w = 0; w < W; ++w) R[w] = dom[depth][w] & ~used[w];
    const uint32_t am = qadj[depth];
    for (uint32_t i = 0; i < depth; ++i) {
      const auto& a = nadj[f[i]];
      if ((am >> i) & 1)
        for (int w = 0; w < W; ++w) R[w] &= a[w];
      else
        for (int w = 0; w < W; ++w) R[w] &= ~a[w];
    }
    if (depth == nq - 1) {
      uint64_t c = 0;
      for (int w = 0; w < W; ++w) c += uint64_t(__builtin_popcountll(R[w]));
      count += c;
      if (find_first && c) stop = true;
      return;
    }
    for (int w = 0; w < W && !stop; ++w) {
      uint64_t bits = R[w];
      while (bits && !stop) {
        unsigned b = unsigned(__builtin_ctzll(bits));
        bits &= bits - 1;
        const uint32_t v = uint32_t(w * 64 + b);
        ++nodes;
        f[depth] = v;
        std::array<uint64_t, W> u2 = used;
        u2[w] |= 1ull << b;
        rec(depth + 1, u2);
      }
    }
  }
};

template <int W>
void pipeline_graph(const orc_csr* q, const orc_csr* d, uint32_t K, bool find_first,
                    const std::vector<NodeTuple>& qt, const std::vector<QueryInfo>& qinfo,
                    uint32_t gi, uint64_t* counts_row, uint64_t* valid_row, uint64_t* stats,
                    std::vector<NodeTuple>& dt) {
  GraphRange r = graph_range(d, gi);
  const uint32_t n = r.end - r.begin;
  dt.resize(n);
  node_tuples<W>(d, gi, K, dt.data());
  std::vector<std::array<uint64_t, W>> nadj(n);
  for (uint32_t x = 0; x < n; ++x) {
    nadj[x].fill(0);
    for (uint32_t i = d->row_offsets[x + r.begin]; i < d->row_offsets[x + r.begin + 1]; ++i) {
      uint32_t y = d->column_indices[i] - r.begin;
      nadj[x][y >> 6] |= 1ull << (y & 63);
    }
  }
  std::array<std::array<uint64_t, W>, 256> labmask{};
  for (uint32_t x = 0; x < n; ++x) labmask[dt[x].label][x >> 6] |= 1ull << (x & 63);
  std::array<uint64_t, W> all{};
  for (uint32_t x = 0; x < n; ++x) all[x >> 6] |= 1ull << (x & 63);
  std::array<uint64_t, W> dom[32];
  const uint32_t Q = q->num_graphs;
  for (uint32_t qi = 0; qi < Q; ++qi) {
    counts_row[qi] = 0;
    const QueryInfo& info = qinfo[qi];
    if (info.n <= 1) continue;
    bool ok = true;
    for (uint32_t i = 0; i < info.n && ok; ++i) {
      const NodeTuple& t = qt[info.begin + i];
      const std::array<uint64_t, W>& cand = t.label == kWildcard ? all : labmask[t.label];
      std::array<uint64_t, W> m{};
      bool any = false;
      for (int w = 0; w < W; ++w) {
        uint64_t bits = cand[w];
        while (bits) {
          unsigned b = unsigned(__builtin_ctzll(bits));
          bits &= bits - 1;
          if (tuple_test(t, dt[w * 64 + b], K)) {
            m[w] |= 1ull << b;
            any = true;
          }
        }
      }
      dom[i] = m;
      ok = any;
    }
    if (!ok) continue;
    valid_row[qi / 64] |= 1ull << (qi % 64);
    JoinState<W> js;
    js.dom = dom;
    js.nadj = nadj.data();
    js.qadj = info.adjmask.data();
    js.nq = info.n;
    js.find_first = find_first;
    std::array<uint64_t, W> used{};
    js.rec(0, used);
    uint64_t c = find_first ? (js.count ? 1 : 0) : js.count;
    counts_row[qi] = c;
    stats[0] += 1;
    stats[1] += c ? 1 : 0;
    stats[2] += c;
    stats[3] += js.nodes;
  }
}

}  // namespace

extern "C" int orc_pipeline(const orc_csr* q, const orc_csr* d, uint32_t K, int find_first,
                            uint32_t g_begin, uint32_t g_end, uint64_t* counts, uint64_t* valid,
                            uint64_t* stats, int threads) {
  if (K > 8 || g_end > d->num_graphs || g_begin > g_end) return 1;
  const uint32_t Q = q->num_graphs;
  const uint64_t VW = (Q + 63) / 64;
  // Query-side tuples (the reference regenerates them per data batch, BMS2_batch.cpp:162-211).
  std::vector<NodeTuple> qt(q->total_nodes);
  std::vector<QueryInfo> qinfo(Q);
  for (uint32_t qi = 0; qi < Q; ++qi) {
    GraphRange r = graph_range(q, qi);
    if (r.end - r.begin > 32) return 2;
    node_tuples<4>(q, qi, K, qt.data() + r.begin);
    QueryInfo& info = qinfo[qi];
    info.begin = r.begin;
    info.n = r.end - r.begin;
    info.adjmask.assign(32, 0);
    for (uint32_t i = 0; i < info.n; ++i)
      for (uint32_t k = q->row_offsets[r.begin + i]; k < q->row_offsets[r.begin + i + 1]; ++k)
        info.adjmask[i] |= 1u << (q->column_indices[k] - r.begin);
  }
  for (uint32_t g = g_begin; g < g_end; ++g) {
    uint32_t n = d->graph_offsets[g + 1] - d->graph_offsets[g];
    if (n > 256) return 3;
  }
  const uint64_t G = g_end - g_begin;
  std::memset(valid, 0, G * VW * sizeof(uint64_t));
  uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel reduction(+ : s0, s1, s2, s3)
  {
    std::vector<NodeTuple> dt;
    uint64_t st[4] = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 16)
    for (int64_t gg = 0; gg < int64_t(G); ++gg) {
      const uint32_t g = g_begin + uint32_t(gg);
      const uint32_t n = d->graph_offsets[g + 1] - d->graph_offsets[g];
      uint64_t* crow = counts + uint64_t(gg) * Q;
      uint64_t* vrow = valid + uint64_t(gg) * VW;
      if (n <= 64)
        pipeline_graph<1>(q, d, K, find_first != 0, qt, qinfo, g, crow, vrow, st, dt);
      else
        pipeline_graph<4>(q, d, K, find_first != 0, qt, qinfo, g, crow, vrow, st, dt);
    }
    s0 += st[0];
    s1 += st[1];
    s2 += st[2];
    s3 += st[3];
  }
  if (stats) {
    stats[0] = s0;
    stats[1] = s1;
    stats[2] = s2;
    stats[3] = s3;
  }
  return 0;
}
