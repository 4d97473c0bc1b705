// Host-side single-graph edge lists and their packing into a batched CSR (bms2_host_csr).
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace bms2 {

struct EdgeListGraph {
  std::vector<uint8_t> labels;  // node labels, node id = index
  struct Edge {
    uint32_t u, v;
    uint8_t label;
  };
  std::vector<Edge> edges;  // undirected, in file / generation order
  uint32_t num_nodes() const { return static_cast<uint32_t>(labels.size()); }
};

// Packs graphs into one batched CSR.  Neighbour order per node follows the edge order, each edge
// appended under both endpoints (toCSRGraph, graph.hpp:212-243; createDeviceCSRGraph,
// graph.hpp:251-290).  Arrays are malloc-ed; release with bms2_host_csr_free.
int pack_graphs(const std::vector<EdgeListGraph>& graphs, bms2_host_csr* out, unsigned threads);

}  // namespace bms2
