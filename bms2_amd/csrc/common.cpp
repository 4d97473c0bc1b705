// Error reporting and host CSR validation for the bms2 C ABI.
#include "common.hpp"

#include <cstdlib>
#include <cstring>

namespace bms2 {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

int validate_host_csr(const bms2_host_csr* g, uint32_t max_graph_nodes, uint32_t* max_nodes_out) {
  if (!g) return set_error(BMS2_ERR_INVALID, "graph batch is null");
  if (g->num_graphs > 0 && (!g->graph_offsets || !g->row_offsets))
    return set_error(BMS2_ERR_INVALID, "graph batch has null offset arrays");
  if (g->total_nodes > 0 && !g->node_labels)
    return set_error(BMS2_ERR_INVALID, "graph batch has null node labels");
  if (g->total_edges > 0 && !g->column_indices)
    return set_error(BMS2_ERR_INVALID, "graph batch has null column indices");
  if (g->total_nodes >= (1ull << 32) || g->total_edges >= (1ull << 32))
    return set_error(BMS2_ERR_UNSUPPORTED, "batch exceeds 2^32-1 nodes or CSR entries (u32 ids)");
  if (g->num_graphs == 0) {
    if (max_nodes_out) *max_nodes_out = 0;
    return g->total_nodes == 0 ? BMS2_OK : set_error(BMS2_ERR_INVALID, "nodes without graphs");
  }
  if (g->graph_offsets[0] != 0 || g->graph_offsets[g->num_graphs] != g->total_nodes)
    return set_error(BMS2_ERR_INVALID, "graph_offsets must start at 0 and end at total_nodes");
  if (g->row_offsets[0] != 0 || g->row_offsets[g->total_nodes] != g->total_edges)
    return set_error(BMS2_ERR_INVALID, "row_offsets must start at 0 and end at total_edges");
  uint32_t max_nodes = 0;
  for (uint32_t gi = 0; gi < g->num_graphs; ++gi) {
    uint32_t s = g->graph_offsets[gi], e = g->graph_offsets[gi + 1];
    if (e < s) return set_error(BMS2_ERR_INVALID, "graph_offsets not monotone at graph %u", gi);
    uint32_t n = e - s;
    if (n > max_nodes) max_nodes = n;
    if (n > max_graph_nodes)
      return set_error(BMS2_ERR_UNSUPPORTED, "graph %u has %u nodes; the limit is %u", gi, n,
                       max_graph_nodes);
    for (uint32_t u = s; u < e; ++u) {
      uint32_t rb = g->row_offsets[u], re = g->row_offsets[u + 1];
      if (re < rb) return set_error(BMS2_ERR_INVALID, "row_offsets not monotone at node %u", u);
      for (uint32_t k = rb; k < re; ++k) {
        uint32_t v = g->column_indices[k];
        if (v < s || v >= e)
          return set_error(BMS2_ERR_INVALID,
                           "CSR entry %u of node %u points to node %u outside graph %u", k, u, v,
                           gi);
      }
    }
  }
  if (max_nodes_out) *max_nodes_out = max_nodes;
  return BMS2_OK;
}

}  // namespace bms2

extern "C" const char* bms2_last_error(void) { return bms2::g_last_error.c_str(); }
extern "C" int bms2_version(void) { return 100; }

extern "C" void bms2_host_csr_free(bms2_host_csr* g) {
  if (!g) return;
  std::free(g->graph_offsets);
  std::free(g->row_offsets);
  std::free(g->column_indices);
  std::free(g->node_labels);
  std::free(g->edge_labels);
  std::memset(g, 0, sizeof(*g));
}
