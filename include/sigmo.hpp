// sigmo.hpp -- header-only C++ facade with the reference's library/include API (namespace sigmo,
// sigmo.hpp:1-14) over the bms2 C ABI.  Host code written against the reference compiles against
// this header with two changes: sigmo::Queue replaces sycl::queue, and join totals are returned
// through a host pointer the call adds to after it completes.
//
// Errors: every failing C call throws std::runtime_error with bms2_last_error()'s message.
#pragma once

#include <chrono>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <cstring>
#include <vector>

#include "bms2.h"

namespace sigmo {

// types.hpp:10-24
namespace types {
using row_offset_t = uint32_t;
using col_index_t = uint32_t;
using label_t = uint8_t;
using node_t = uint32_t;
using candidates_t = uint64_t;  // u64 words; the reference's u32 words hold the same bits
constexpr label_t WILDCARD_NODE = 255;
constexpr node_t NULL_NODE = 0xffffffffu;
}  // namespace types

inline void check(int rc) {
  if (rc != BMS2_OK) throw std::runtime_error(std::string("bms2: ") + bms2_last_error());
}

// Stands in for sycl::queue (BMS2.cpp:14): one in-order HIP stream on one device.
class Queue {
 public:
  explicit Queue(int device = 0) { check(bms2_queue_create(device, &q_)); }
  Queue(const Queue&) = delete;
  Queue& operator=(const Queue&) = delete;
  ~Queue() { bms2_queue_destroy(q_); }
  void wait() { check(bms2_queue_wait(q_)); }
  void wait_and_throw() { wait(); }
  bms2_queue handle() const { return q_; }

 private:
  bms2_queue q_ = nullptr;
};

namespace utils {
// BatchedEvent (utils.hpp:123-157): events of one call; profiling info = summed kernel time.
class BatchedEvent {
 public:
  BatchedEvent() = default;
  void add(bms2_event e) {
    if (e) events_.push_back(std::shared_ptr<bms2_event_s>(e, bms2_event_destroy));
  }
  void wait() {
    float ms;
    for (auto& e : events_) check(bms2_event_elapsed_ms(e.get(), &ms));
  }
  std::chrono::duration<double> getProfilingInfo() {
    double total = 0;
    for (auto& e : events_) {
      float ms = 0;
      check(bms2_event_elapsed_ms(e.get(), &ms));
      total += ms / 1000.0;
    }
    return std::chrono::duration<double>(total);
  }
  size_t numEvents() const { return events_.size(); }

 private:
  std::vector<std::shared_ptr<bms2_event_s>> events_;
};
}  // namespace utils

// CSRGraph (graph.hpp:33-56)
class CSRGraph {
 public:
  CSRGraph(std::vector<types::row_offset_t> row_offsets, std::vector<types::col_index_t> column_indices,
           std::vector<types::label_t> node_labels, std::vector<types::label_t> edge_labels,
           size_t num_nodes)
      : row_offsets_(std::move(row_offsets)), column_indices_(std::move(column_indices)),
        node_labels_(std::move(node_labels)), edge_labels_(std::move(edge_labels)), num_nodes_(num_nodes) {}
  const types::row_offset_t* getRowOffsets() const { return row_offsets_.data(); }
  const types::col_index_t* getColumnIndices() const { return column_indices_.data(); }
  const types::label_t* getNodeLabels() const { return node_labels_.data(); }
  const types::label_t* getEdgeLabels() const { return edge_labels_.data(); }
  size_t getNumNodes() const { return num_nodes_; }

 private:
  std::vector<types::row_offset_t> row_offsets_;
  std::vector<types::col_index_t> column_indices_;
  std::vector<types::label_t> node_labels_;
  std::vector<types::label_t> edge_labels_;
  size_t num_nodes_;
};

// DeviceBatchedCSRGraph (graph.hpp:58-108): same fields, device pointers.
struct DeviceBatchedCSRGraph : bms2_device_csr {
  DeviceBatchedCSRGraph() : bms2_device_csr{} {}
};

// createDeviceCSRGraph (graph.hpp:251-310) with the packing on the GPU (graph.hpp:249 "TODO
// offload on GPU"): the host only lays the graphs' own arrays end to end; the device adds the
// node and entry offsets and validates (bms2_device_csr_pack).
inline DeviceBatchedCSRGraph createDeviceCSRGraph(Queue& queue, std::vector<CSRGraph>& graphs) {
  size_t N = 0, E = 0;
  for (const CSRGraph& gr : graphs) {
    N += gr.getNumNodes();
    E += gr.getRowOffsets()[gr.getNumNodes()];
  }
  std::vector<uint32_t> nodes(graphs.size()), ro(N), ci(E);
  std::vector<uint8_t> nl(N), el(E);
  size_t n0 = 0, e0 = 0;
  for (size_t g = 0; g < graphs.size(); ++g) {
    const CSRGraph& gr = graphs[g];
    const size_t n = gr.getNumNodes(), m = gr.getRowOffsets()[n];
    nodes[g] = uint32_t(n);
    if (n) {
      std::memcpy(ro.data() + n0, gr.getRowOffsets() + 1, 4 * n);
      std::memcpy(nl.data() + n0, gr.getNodeLabels(), n);
    }
    if (m) {
      std::memcpy(ci.data() + e0, gr.getColumnIndices(), 4 * m);
      std::memcpy(el.data() + e0, gr.getEdgeLabels(), m);
    }
    n0 += n;
    e0 += m;
  }
  DeviceBatchedCSRGraph d;
  check(bms2_device_csr_pack(queue.handle(), uint32_t(graphs.size()), nodes.data(), ro.data(),
                             ci.data(), nl.data(), el.data(), N, E, &d));
  return d;
}

inline void destroyDeviceCSRGraph(DeviceBatchedCSRGraph& g, Queue& queue) {
  check(bms2_device_csr_destroy(queue.handle(), &g));
}

inline size_t getDeviceGraphAllocSize(const DeviceBatchedCSRGraph& g) {
  return size_t(bms2_device_csr_alloc_size(&g));
}

namespace io {
namespace detail {
// Splits a packed host batch into per-graph CSRGraphs (local ids, row offsets from 0).
inline std::vector<CSRGraph> split(bms2_host_csr& h) {
  std::vector<CSRGraph> out;
  out.reserve(h.num_graphs);
  for (uint32_t g = 0; g < h.num_graphs; ++g) {
    const uint32_t s = h.graph_offsets[g], e = h.graph_offsets[g + 1];
    const uint32_t eb = h.row_offsets[s], ee = h.row_offsets[e];
    std::vector<types::row_offset_t> ro;
    for (uint32_t u = s; u <= e; ++u) ro.push_back(h.row_offsets[u] - eb);
    std::vector<types::col_index_t> ci;
    for (uint32_t k = eb; k < ee; ++k) ci.push_back(h.column_indices[k] - s);
    out.emplace_back(std::move(ro), std::move(ci),
                     std::vector<types::label_t>(h.node_labels + s, h.node_labels + e),
                     std::vector<types::label_t>(h.edge_labels + eb, h.edge_labels + ee), e - s);
  }
  bms2_host_csr_free(&h);
  return out;
}
}  // namespace detail

// loadCSRGraphsFromFile (io.hpp:43-52) over the library's .dat reader.
inline std::vector<CSRGraph> loadCSRGraphsFromFile(const std::string& filename) {
  bms2_host_csr h{};
  check(bms2_load_dat(filename.c_str(), 0, 0, &h));
  return detail::split(h);
}

// loadCSRGraphsFromLines (io.hpp:34-41); skip_malformed drops lines the reader rejects (the
// batch driver's try/catch around each line, BMS2_batch.cpp:119-122).
inline std::vector<CSRGraph> loadCSRGraphsFromLines(const std::vector<std::string>& lines,
                                                    bool skip_malformed = false) {
  std::vector<const char*> ptrs(lines.size());
  for (size_t i = 0; i < lines.size(); ++i) ptrs[i] = lines[i].c_str();
  bms2_host_csr h{};
  check(bms2_parse_dat_lines(ptrs.data(), ptrs.size(), skip_malformed ? 1 : 0, &h));
  return detail::split(h);
}
}  // namespace io

namespace signature {
enum class Algorithm { ViewBased, PowerGraph };

// Signature<PowerGraph, 4> (signature.hpp:18-390); only the PowerGraph algorithm is provided.
template <Algorithm A = Algorithm::PowerGraph, size_t Bits = 4>
class Signature {
  static_assert(A == Algorithm::PowerGraph && Bits == 4, "only Signature<PowerGraph, 4>");

 public:
  Signature(Queue& queue, size_t data_nodes, size_t query_nodes) {
    check(bms2_signature_create(queue.handle(), data_nodes, query_nodes, &s_));
  }
  ~Signature() { bms2_signature_destroy(s_); }
  Signature(const Signature&) = delete;
  utils::BatchedEvent generateDataSignatures(DeviceBatchedCSRGraph& g) { return gen(g, BMS2_SCOPE_DATA); }
  utils::BatchedEvent generateQuerySignatures(DeviceBatchedCSRGraph& g) { return gen(g, BMS2_SCOPE_QUERY); }
  utils::BatchedEvent refineDataSignatures(DeviceBatchedCSRGraph& g, size_t view_size = 1) {
    return ref(g, view_size, BMS2_SCOPE_DATA);
  }
  utils::BatchedEvent refineQuerySignatures(DeviceBatchedCSRGraph& g, size_t view_size = 1) {
    return ref(g, view_size, BMS2_SCOPE_QUERY);
  }
  static constexpr size_t getMaxLabels() { return 16; }
  bms2_signature handle() const { return s_; }

 private:
  utils::BatchedEvent gen(DeviceBatchedCSRGraph& g, int scope) {
    bms2_event e = nullptr;
    check(bms2_signature_generate(s_, &g, scope, &e));
    utils::BatchedEvent be;
    be.add(e);
    return be;
  }
  utils::BatchedEvent ref(DeviceBatchedCSRGraph& g, size_t view, int scope) {
    bms2_event e = nullptr;
    check(bms2_signature_refine(s_, &g, uint32_t(view), scope, &e));
    utils::BatchedEvent be;
    be.add(e);
    return be;
  }
  bms2_signature s_ = nullptr;
};

// PathSignature (signature.hpp:406-532)
class PathSignature {
 public:
  PathSignature(Queue& queue, size_t data_nodes, size_t query_nodes) {
    check(bms2_path_signature_create(queue.handle(), data_nodes, query_nodes, &s_));
  }
  ~PathSignature() { bms2_path_signature_destroy(s_); }
  PathSignature(const PathSignature&) = delete;
  utils::BatchedEvent generateDataPathSignatures(DeviceBatchedCSRGraph& g) { return gen(g, BMS2_SCOPE_DATA); }
  utils::BatchedEvent generateQueryPathSignatures(DeviceBatchedCSRGraph& g) { return gen(g, BMS2_SCOPE_QUERY); }
  bms2_path_signature handle() const { return s_; }

 private:
  utils::BatchedEvent gen(DeviceBatchedCSRGraph& g, int scope) {
    bms2_event e = nullptr;
    check(bms2_path_signature_generate(s_, &g, scope, &e));
    utils::BatchedEvent be;
    be.add(e);
    return be;
  }
  bms2_path_signature s_ = nullptr;
};

// CycleSignature (signature.hpp:535-732)
class CycleSignature {
 public:
  CycleSignature(Queue& queue, size_t data_nodes, size_t query_nodes) {
    check(bms2_cycle_signature_create(queue.handle(), data_nodes, query_nodes, &s_));
  }
  ~CycleSignature() { bms2_cycle_signature_destroy(s_); }
  CycleSignature(const CycleSignature&) = delete;
  utils::BatchedEvent generateDataCycleSignatures(DeviceBatchedCSRGraph& g) { return gen(g, BMS2_SCOPE_DATA); }
  utils::BatchedEvent generateQueryCycleSignatures(DeviceBatchedCSRGraph& g) { return gen(g, BMS2_SCOPE_QUERY); }
  bms2_cycle_signature handle() const { return s_; }

 private:
  utils::BatchedEvent gen(DeviceBatchedCSRGraph& g, int scope) {
    bms2_event e = nullptr;
    check(bms2_cycle_signature_generate(s_, &g, scope, &e));
    utils::BatchedEvent be;
    be.add(e);
    return be;
  }
  bms2_cycle_signature s_ = nullptr;
};
}  // namespace signature

namespace candidates {
// Candidates (candidates.hpp:10-200)
class Candidates {
 public:
  Candidates(Queue& queue, size_t source_nodes, size_t target_nodes) : rows_(source_nodes) {
    check(bms2_candidates_create(queue.handle(), source_nodes, target_nodes, &c_));
  }
  ~Candidates() { bms2_candidates_destroy(c_); }
  Candidates(const Candidates&) = delete;
  // Host copy of the bit-matrix: rows x getWordsPerRow() u64 words.
  std::vector<uint64_t> getHostCandidates() {
    std::vector<uint64_t> out(rows_ * getWordsPerRow());
    check(bms2_candidates_copy_to_host(c_, out.data()));
    return out;
  }
  // getCandidatesCount(source_node) for every row (candidates.hpp:73-80).
  std::vector<uint32_t> getCandidatesCounts() {
    std::vector<uint32_t> out(rows_);
    check(bms2_candidates_row_counts(c_, out.data()));
    return out;
  }
  size_t getWordsPerRow() const { return size_t(bms2_candidates_words_per_row(c_)); }
  size_t getAllocationSize() const { return rows_ * getWordsPerRow() * sizeof(uint64_t); }
  bms2_candidates handle() const { return c_; }

 private:
  bms2_candidates c_ = nullptr;
  size_t rows_;
};
}  // namespace candidates

namespace isomorphism {
namespace filter {
// filterCandidates (isomorphism.hpp:30-104)
inline utils::BatchedEvent filterCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graph,
                                            DeviceBatchedCSRGraph& data_graph,
                                            signature::Signature<>& label_signatures,
                                            signature::PathSignature& path_signatures,
                                            signature::CycleSignature& cycle_signatures,
                                            candidates::Candidates& candidates) {
  bms2_event e = nullptr;
  check(bms2_filter_candidates(queue.handle(), &query_graph, &data_graph, label_signatures.handle(),
                               path_signatures.handle(), cycle_signatures.handle(),
                               candidates.handle(), &e));
  utils::BatchedEvent be;
  be.add(e);
  return be;
}

// refineCandidates (isomorphism.hpp:106-203)
inline utils::BatchedEvent refineCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graph,
                                            DeviceBatchedCSRGraph& data_graph,
                                            signature::Signature<>& signatures,
                                            candidates::Candidates& candidates) {
  bms2_event e = nullptr;
  check(bms2_refine_candidates(queue.handle(), &query_graph, &data_graph, signatures.handle(),
                               candidates.handle(), &e));
  utils::BatchedEvent be;
  be.add(e);
  return be;
}
}  // namespace filter

namespace mapping {
// GMCR (gmcr.hpp:16-145)
class GMCR {
 public:
  explicit GMCR(Queue& queue) { check(bms2_gmcr_create(queue.handle(), &g_)); }
  ~GMCR() { bms2_gmcr_destroy(g_); }
  GMCR(const GMCR&) = delete;
  utils::BatchedEvent generateGMCR(DeviceBatchedCSRGraph& query_graphs, DeviceBatchedCSRGraph& data_graphs,
                                   candidates::Candidates& candidates) {
    bms2_event e = nullptr;
    check(bms2_gmcr_generate(g_, &query_graphs, &data_graphs, candidates.handle(), &e));
    utils::BatchedEvent be;
    be.add(e);
    return be;
  }
  size_t getTotalQueryIndices() const { return size_t(bms2_gmcr_total(g_)); }
  bms2_gmcr handle() const { return g_; }

 private:
  bms2_gmcr g_ = nullptr;
};
}  // namespace mapping

namespace join {
// joinCandidates (isomorphism.hpp:291-472).  The total is added to *num_matches when the call
// returns (the reference adds it on the device into host-visible memory); per_pair (optional)
// receives one count per GMCR entry.
inline utils::BatchedEvent joinCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graphs,
                                          DeviceBatchedCSRGraph& data_graphs,
                                          candidates::Candidates& candidates, mapping::GMCR& gmcr,
                                          size_t* num_matches, bool find_first = true,
                                          std::vector<uint64_t>* per_pair = nullptr) {
  void* d_total = nullptr;
  void* d_pairs = nullptr;
  const size_t T = gmcr.getTotalQueryIndices();
  check(bms2_device_malloc(queue.handle(), 8, &d_total));
  check(bms2_memset_zero(queue.handle(), d_total, 8));
  if (per_pair) check(bms2_device_malloc(queue.handle(), 8 * (T ? T : 1), &d_pairs));
  bms2_event e = nullptr;
  int rc = bms2_join_candidates(queue.handle(), &query_graphs, &data_graphs, candidates.handle(),
                                gmcr.handle(), static_cast<uint64_t*>(d_total), find_first ? 1 : 0,
                                static_cast<uint64_t*>(d_pairs), &e);
  uint64_t total = 0;
  if (rc == BMS2_OK) rc = bms2_memcpy_d2h(queue.handle(), &total, d_total, 8);
  if (rc == BMS2_OK && per_pair) {
    per_pair->assign(T, 0);
    if (T) rc = bms2_memcpy_d2h(queue.handle(), per_pair->data(), d_pairs, 8 * T);
  }
  bms2_device_free(queue.handle(), d_total);
  if (d_pairs) bms2_device_free(queue.handle(), d_pairs);
  check(rc);
  if (num_matches) *num_matches += size_t(total);
  utils::BatchedEvent be;
  be.add(e);
  return be;
}

// joinPartialCandidates2 (isomorphism.hpp:1021-1178), the first step of BMS2 --use-cs: every
// embedding of query 0's first `end` nodes, written to device memory `partial_matches` (records
// of end + 1 ints {data node ids, data graph id}; `capacity` records, e.g. from
// sigmo::malloc_device<int>(capacity * (end + 1), queue)).  *num_partial_matches receives the
// number found.  Throws when it exceeds capacity (the reference writes past its buffer).
inline utils::BatchedEvent joinPartialCandidates2(Queue& queue, DeviceBatchedCSRGraph& query_graphs,
                                                  DeviceBatchedCSRGraph& data_graphs,
                                                  candidates::Candidates& candidates, size_t end,
                                                  int* partial_matches, size_t* num_partial_matches,
                                                  size_t capacity) {
  uint64_t n = 0;
  bms2_event e = nullptr;
  const int rc = bms2_join_partial_candidates(queue.handle(), &query_graphs, &data_graphs,
                                              candidates.handle(), uint32_t(end),
                                              reinterpret_cast<int32_t*>(partial_matches),
                                              capacity, &n, &e);
  if (num_partial_matches) *num_partial_matches = size_t(n);
  check(rc);
  utils::BatchedEvent be;
  be.add(e);
  return be;
}

// joinWithPartialMatches (isomorphism.hpp:756-977), the second step of BMS2 --use-cs: extends
// the partial matches to every GMCR query of their data graph; the total is added to
// *total_full_matches when the call returns.  The reference's pair_done flags ([Q * D]) are kept
// per GMCR entry inside the call, so pair_done is not read; per_pair (optional) receives one
// count (find_first: 0/1) per GMCR entry.
inline utils::BatchedEvent joinWithPartialMatches(Queue& queue, DeviceBatchedCSRGraph& query_graphs,
                                                  DeviceBatchedCSRGraph& data_graphs,
                                                  candidates::Candidates& candidates,
                                                  mapping::GMCR& gmcr, size_t end,
                                                  const int* partial_matches,
                                                  size_t num_partial_matches,
                                                  size_t* total_full_matches, bool find_first,
                                                  uint32_t* pair_done = nullptr,
                                                  std::vector<uint64_t>* per_pair = nullptr) {
  (void)pair_done;
  void* d_total = nullptr;
  void* d_pairs = nullptr;
  const size_t T = gmcr.getTotalQueryIndices();
  check(bms2_device_malloc(queue.handle(), 8, &d_total));
  check(bms2_memset_zero(queue.handle(), d_total, 8));
  check(bms2_device_malloc(queue.handle(), 8 * (T ? T : 1), &d_pairs));
  bms2_event e = nullptr;
  int rc = bms2_join_with_partial_matches(
      queue.handle(), &query_graphs, &data_graphs, candidates.handle(), gmcr.handle(),
      uint32_t(end), reinterpret_cast<const int32_t*>(partial_matches), num_partial_matches,
      static_cast<uint64_t*>(d_total), find_first ? 1 : 0, static_cast<uint64_t*>(d_pairs), &e);
  uint64_t total = 0;
  if (rc == BMS2_OK) rc = bms2_memcpy_d2h(queue.handle(), &total, d_total, 8);
  if (rc == BMS2_OK && per_pair) {
    per_pair->assign(T, 0);
    if (T) rc = bms2_memcpy_d2h(queue.handle(), per_pair->data(), d_pairs, 8 * T);
  }
  bms2_device_free(queue.handle(), d_total);
  bms2_device_free(queue.handle(), d_pairs);
  check(rc);
  if (total_full_matches) *total_full_matches += size_t(total);
  utils::BatchedEvent be;
  be.add(e);
  return be;
}
}  // namespace join
}  // namespace isomorphism

// sycl::malloc_device / sycl::free for the buffers the drivers pass between calls.
template <typename T>
T* malloc_device(size_t count, Queue& queue) {
  void* p = nullptr;
  check(bms2_device_malloc(queue.handle(), sizeof(T) * (count ? count : 1), &p));
  return static_cast<T*>(p);
}
inline void free(void* p, Queue& queue) { bms2_device_free(queue.handle(), p); }
}  // namespace sigmo
