// Shared host-side helpers of the bms2 library: error reporting and small utilities.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/bms2.h"

namespace bms2 {

// Records a formatted message for bms2_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Host-side view of a packed batch (host or device pointers, same layout).
struct CsrView {
  uint32_t num_graphs;
  uint64_t total_nodes;
  uint64_t total_edges;
  const uint32_t* graph_offsets;
  const uint32_t* row_offsets;
  const uint32_t* column_indices;
  const uint8_t* node_labels;
  const uint8_t* edge_labels;
};

// Checks the structural invariants every kernel relies on: monotone offsets, column ids inside
// their own graph, sizes within BMS2_MAX_GRAPH_NODES.  Returns BMS2_OK or an error code with the
// message set.  `max_nodes_out` receives the largest graph size.
int validate_host_csr(const bms2_host_csr* g, uint32_t max_graph_nodes, uint32_t* max_nodes_out);

// Isomorphism classes of a query batch (labels and adjacency; edge labels ignored): rep[g] = the
// lowest-indexed query isomorphic to query g, proven by an explicit isomorphism.  Returns the
// number of queries with rep[g] != g.
uint32_t query_classes(const bms2_host_csr* q, uint32_t* rep);

// orbit[u] (u local to query g) = lowest local node of u's orbit under g's label-preserving
// automorphisms, each proven by an explicit automorphism (u alone when unproven).
void query_orbits(const bms2_host_csr* q, uint32_t g, uint8_t* orbit);

// Symmetry breaking of query g under a matching order (order[d] = local node at depth d, null =
// identity) from its orbits: the candidates at the depths of vmask must exceed the node matched
// at depth du, and the count is multiplied by popcount(vmask) + 1 (du = 0xff: none).
void query_symmetry(const bms2_host_csr* q, uint32_t g, const uint8_t* order, const uint8_t* orbit,
                    uint8_t* du, uint32_t* vmask);

}  // namespace bms2
