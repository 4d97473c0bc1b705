// Packing of edge-list graphs into the batched CSR layout.
#include "hostgraph.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace bms2 {

namespace {
template <class F>
void parallel_for_chunks(size_t n, unsigned threads, F&& f) {
  if (threads <= 1 || n < 1024) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  size_t chunk = (n + threads - 1) / threads;
  for (unsigned t = 0; t < threads; ++t) {
    size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}
}  // namespace

int pack_graphs(const std::vector<EdgeListGraph>& graphs, bms2_host_csr* out, unsigned threads) {
  std::memset(out, 0, sizeof(*out));
  const size_t G = graphs.size();
  if (G >= (1ull << 32)) return set_error(BMS2_ERR_UNSUPPORTED, "too many graphs");
  std::vector<uint64_t> node_off(G + 1, 0), edge_off(G + 1, 0);
  for (size_t g = 0; g < G; ++g) {
    for (const auto& e : graphs[g].edges) {
      if (e.u == e.v)
        return set_error(BMS2_ERR_UNSUPPORTED,
                         "graph %zu has a self-loop on node %u; self-loops are not supported", g,
                         e.u);
      if (e.u >= graphs[g].num_nodes() || e.v >= graphs[g].num_nodes())
        return set_error(BMS2_ERR_FORMAT, "graph %zu has an edge to a node outside the graph", g);
    }
    node_off[g + 1] = node_off[g] + graphs[g].num_nodes();
    edge_off[g + 1] = edge_off[g] + 2ull * graphs[g].edges.size();
  }
  const uint64_t N = node_off[G], E = edge_off[G];
  if (N >= (1ull << 32) || E >= (1ull << 32))
    return set_error(BMS2_ERR_UNSUPPORTED, "batch exceeds 2^32-1 nodes or CSR entries");
  out->num_graphs = static_cast<uint32_t>(G);
  out->total_nodes = N;
  out->total_edges = E;
  out->graph_offsets = static_cast<uint32_t*>(std::malloc((G + 1) * sizeof(uint32_t)));
  out->row_offsets = static_cast<uint32_t*>(std::malloc((N + 1) * sizeof(uint32_t)));
  out->column_indices = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(E, 1) * 4));
  out->node_labels = static_cast<uint8_t*>(std::malloc(std::max<uint64_t>(N, 1)));
  out->edge_labels = static_cast<uint8_t*>(std::malloc(std::max<uint64_t>(E, 1)));
  if (!out->graph_offsets || !out->row_offsets || !out->column_indices || !out->node_labels ||
      !out->edge_labels) {
    bms2_host_csr_free(out);
    return set_error(BMS2_ERR_NOMEM, "host allocation of %llu nodes failed",
                     (unsigned long long)N);
  }
  for (size_t g = 0; g <= G; ++g) out->graph_offsets[g] = static_cast<uint32_t>(node_off[g]);
  out->row_offsets[N] = static_cast<uint32_t>(E);
  parallel_for_chunks(G, threads, [&](size_t gb, size_t ge) {
    std::vector<uint32_t> deg, pos;
    for (size_t g = gb; g < ge; ++g) {
      const EdgeListGraph& gr = graphs[g];
      const uint32_t n = gr.num_nodes();
      const uint32_t base = static_cast<uint32_t>(node_off[g]);
      const uint32_t ebase = static_cast<uint32_t>(edge_off[g]);
      deg.assign(n + 1, 0);
      for (const auto& e : gr.edges) {
        deg[e.u + 1]++;
        deg[e.v + 1]++;
      }
      for (uint32_t i = 1; i <= n; ++i) deg[i] += deg[i - 1];
      for (uint32_t i = 0; i < n; ++i) {
        out->row_offsets[base + i] = ebase + deg[i];
        out->node_labels[base + i] = gr.labels[i];
      }
      pos.assign(n, 0);
      for (const auto& e : gr.edges) {
        uint32_t a = e.u, b = e.v;
        uint32_t ia = ebase + deg[a] + pos[a];
        out->column_indices[ia] = base + b;
        out->edge_labels[ia] = e.label;
        uint32_t ib = ebase + deg[b] + pos[b];
        out->column_indices[ib] = base + a;
        out->edge_labels[ib] = e.label;
        pos[a]++;
        pos[b]++;
      }
    }
  });
  return BMS2_OK;
}

}  // namespace bms2
