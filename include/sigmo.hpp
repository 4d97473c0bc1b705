// sigmo.hpp -- header-only C++ facade with the reference's library/include API (namespace sigmo,
// sigmo.hpp:1-14) over the bms2 C ABI.  A driver written against the reference (BMS2.cpp,
// BMS2_batch.cpp) compiles against this header once `#include <sycl/sycl.hpp>` is dropped and its
// `sycl::` names are spelt `sigmo::`: sigmo::queue (the stream), gpu_selector_v and
// property::queue::enable_profiling, queue.get_device().get_info<info::device::...>(),
// queue.memset(...).wait(), malloc_shared / malloc_device / free.  Join totals are added to the
// caller's host-visible counter when the call returns (the reference adds them on the device).
//
// Errors: every failing C call throws std::runtime_error with bms2_last_error()'s message.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
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

// device.hpp:9: the filter's candidate layout; only the query-major layout (the default of every
// stage call and the only one the drivers use) is provided.
enum class CandidatesDomain { Query, Data };

inline void check(int rc) {
  if (rc != BMS2_OK) throw std::runtime_error(std::string("bms2: ") + bms2_last_error());
}

// Tags standing in for sycl::gpu_selector_v, sycl::property::queue::enable_profiling and
// sycl::info::device::{global_mem_size, name, max_compute_units} (BMS2.cpp:14-16).
struct gpu_selector_t {};
inline constexpr gpu_selector_t gpu_selector_v{};
namespace property {
namespace queue {
struct enable_profiling {};
struct in_order {};
}  // namespace queue
}  // namespace property
namespace info {
namespace device {
struct global_mem_size {
  using return_type = size_t;
};
struct name {
  using return_type = std::string;
};
struct max_compute_units {
  using return_type = uint32_t;
};
}  // namespace device
}  // namespace info

// A device of a queue: get_info<info::device::...>().
class Device {
 public:
  explicit Device(bms2_queue q) : q_(q) {}
  template <typename Param>
  typename Param::return_type get_info() const {
    uint64_t mem = 0;
    uint32_t cus = 0;
    char name[256] = {0};
    check(bms2_queue_device_info(q_, &mem, &cus, name, sizeof(name)));
    if constexpr (std::is_same_v<Param, info::device::global_mem_size>) {
      return size_t(mem);
    } else if constexpr (std::is_same_v<Param, info::device::name>) {
      return std::string(name);
    } else {
      static_assert(std::is_same_v<Param, info::device::max_compute_units>, "unknown device info");
      return cus;
    }
  }

 private:
  bms2_queue q_;
};

namespace detail {
// Allocations made through malloc_shared / malloc_host / malloc_device, so that free() releases
// each the right way and the CS join can size its output from the caller's buffer (the
// reference's call passes no capacity, BMS2.cpp:250-252).
enum class Kind { Device, Shared };
struct Allocation {
  size_t bytes;
  Kind kind;
};
inline std::mutex& alloc_mutex() {
  static std::mutex m;
  return m;
}
inline std::map<uintptr_t, Allocation>& allocations() {
  static std::map<uintptr_t, Allocation> a;
  return a;
}
inline void remember(void* p, size_t bytes, Kind k) {
  std::lock_guard<std::mutex> g(alloc_mutex());
  allocations()[reinterpret_cast<uintptr_t>(p)] = Allocation{bytes, k};
}
// The allocation containing p: {bytes from p to its end, kind}; bytes == 0 when p is unknown.
inline Allocation lookup(const void* p) {
  std::lock_guard<std::mutex> g(alloc_mutex());
  const auto& a = allocations();
  const uintptr_t x = reinterpret_cast<uintptr_t>(p);
  auto it = a.upper_bound(x);
  if (!p || it == a.begin()) return Allocation{0, Kind::Device};
  --it;
  if (x >= it->first + it->second.bytes) return Allocation{0, Kind::Device};
  return Allocation{it->first + it->second.bytes - x, it->second.kind};
}
inline bool forget(void* p, Kind* kind) {
  std::lock_guard<std::mutex> g(alloc_mutex());
  auto it = allocations().find(reinterpret_cast<uintptr_t>(p));
  if (it == allocations().end()) return false;
  *kind = it->second.kind;
  allocations().erase(it);
  return true;
}
}  // namespace detail

// Stands in for sycl::queue (BMS2.cpp:14): one in-order HIP stream on one device.
class Queue {
 public:
  // The value returned by queue.memset(...) (a sycl::event): wait() waits for the queue.
  class Event {
   public:
    explicit Event(bms2_queue q) : q_(q) {}
    void wait() { check(bms2_queue_wait(q_)); }
    void wait_and_throw() { wait(); }

   private:
    bms2_queue q_;
  };

  explicit Queue(int device = 0) { check(bms2_queue_create(device, &q_)); }
  // sycl::queue{sycl::gpu_selector_v, sycl::property::queue::enable_profiling{}}: device 0;
  // every stage call is timed with its own event pair, so profiling is always on.
  template <typename... Props>
  explicit Queue(const gpu_selector_t&, const Props&...) : Queue(0) {}
  Queue(const Queue&) = delete;
  Queue& operator=(const Queue&) = delete;
  ~Queue() { bms2_queue_destroy(q_); }
  void wait() { check(bms2_queue_wait(q_)); }
  void wait_and_throw() { wait(); }
  Device get_device() const { return Device(q_); }
  // queue.memset(ptr, value, bytes): host-visible memory from malloc_shared is set on the host
  // after the queue's earlier work; device memory takes value 0 only.
  Event memset(void* ptr, int value, size_t bytes) {
    const detail::Allocation a = detail::lookup(ptr);
    if (a.bytes && a.kind == detail::Kind::Shared) {
      wait();
      std::memset(ptr, value, bytes);
    } else {
      if (value != 0) throw std::runtime_error("bms2: memset of device memory takes value 0 only");
      check(bms2_memset_zero(q_, ptr, bytes));
    }
    return Event(q_);
  }
  bms2_queue handle() const { return q_; }

 private:
  bms2_queue q_ = nullptr;
};
using queue = Queue;  // the reference's spelling, sycl::queue

namespace device {
// DeviceOptions (device.hpp:13-17): the reference's work-group sizes of its filter and join
// kernels, set from the command line (src/utils.hpp:57-58).  This library's kernels fix their
// own launch shapes (one wave per data-node word for the filter, persistent waves for the join;
// the fused pipeline's shapes come from bms2_pipeline_autotune), so the values are kept for the
// drivers' reports and not read by the calls.
struct DeviceOptions {
  size_t join_work_group_size = 128;
  size_t filter_work_group_size = 512;
};
inline DeviceOptions deviceOptions{};

// device.hpp:19-23
inline size_t getDeviceMemorySize(Queue& queue) {
  return queue.get_device().get_info<info::device::global_mem_size>();
}
inline size_t getPreferredWorkGroupSize(Queue&) { return 1024; }
}  // namespace device

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

// IntermediateGraph (graph.hpp:147-243): one .dat line; the constructor throws
// std::runtime_error on a malformed line (as BMS2_batch.cpp:122 expects), toCSRGraph() gives the
// graph with each edge under both endpoints in file order.
class IntermediateGraph {
 public:
  IntermediateGraph() = default;
  explicit IntermediateGraph(const std::string& graph_str) {
    const char* line = graph_str.c_str();
    bms2_host_csr h{};
    check(bms2_parse_dat_lines(&line, 1, 0, &h));
    auto graphs = io::detail::split(h);
    if (graphs.size() != 1) throw std::runtime_error("bms2: expected one graph per line");
    graph_ = std::make_shared<CSRGraph>(std::move(graphs[0]));
  }
  CSRGraph toCSRGraph() const {
    if (!graph_) throw std::runtime_error("bms2: empty IntermediateGraph");
    return *graph_;
  }
  size_t getNumNodes() const { return graph_ ? graph_->getNumNodes() : 0; }

 private:
  std::shared_ptr<CSRGraph> graph_;
};

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
  // signature.hpp:180-193: device bytes of each scope (histograms + counted-node planes)
  size_t getDataSignatureAllocationSize() const {
    return size_t(bms2_signature_alloc_size(s_, BMS2_SCOPE_DATA));
  }
  size_t getQuerySignatureAllocationSize() const {
    return size_t(bms2_signature_alloc_size(s_, BMS2_SCOPE_QUERY));
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
  // signature.hpp:521-522
  size_t getDataSignatureAllocationSize() const {
    return size_t(bms2_path_signature_alloc_size(s_, BMS2_SCOPE_DATA));
  }
  size_t getQuerySignatureAllocationSize() const {
    return size_t(bms2_path_signature_alloc_size(s_, BMS2_SCOPE_QUERY));
  }
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
  // signature.hpp:721-722
  size_t getDataSignatureAllocationSize() const {
    return size_t(bms2_cycle_signature_alloc_size(s_, BMS2_SCOPE_DATA));
  }
  size_t getQuerySignatureAllocationSize() const {
    return size_t(bms2_cycle_signature_alloc_size(s_, BMS2_SCOPE_QUERY));
  }
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
namespace detail {
// popcount and the k-th set bit of the bits [start, end) of a row of u64 words (the range
// queries of candidates.hpp:82-158, over exact ranges: the reference's mask shifts by 32 when
// end is a multiple of 32).
inline uint32_t range_count(const uint64_t* row, uint64_t start, uint64_t end) {
  uint32_t n = 0;
  for (uint64_t w = start / 64; start < end && w <= (end - 1) / 64; ++w) {
    uint64_t m = row[w];
    if (w == start / 64) m &= ~uint64_t(0) << (start % 64);
    if (w == (end - 1) / 64 && end % 64) m &= ~uint64_t(0) >> (64 - end % 64);
    n += uint32_t(__builtin_popcountll(m));
  }
  return n;
}
inline types::node_t range_select(const uint64_t* row, uint32_t idx, uint64_t start, uint64_t end) {
  for (uint64_t w = start / 64; start < end && w <= (end - 1) / 64; ++w) {
    uint64_t m = row[w];
    if (w == start / 64) m &= ~uint64_t(0) << (start % 64);
    if (w == (end - 1) / 64 && end % 64) m &= ~uint64_t(0) >> (64 - end % 64);
    const uint32_t c = uint32_t(__builtin_popcountll(m));
    if (idx < c) {
      for (; idx; --idx) m &= m - 1;
      return types::node_t(w * 64 + uint64_t(__builtin_ctzll(m)));
    }
    idx -= c;
  }
  return types::NULL_NODE;
}
}  // namespace detail

// The host copy returned by Candidates::getHostCandidates (the reference's host-side
// CandidatesDevice, candidates.hpp:181-191): rows of u64 words, bit d % 64 of word d / 64 set
// when data node d is a candidate of the row's query node.
class HostCandidates {
 public:
  HostCandidates(std::vector<uint64_t> words, size_t rows, size_t words_per_row, size_t targets)
      : words_(std::move(words)), rows_(rows), W_(words_per_row), targets_(targets) {}
  // getCandidatesCount(source_node) (candidates.hpp:73-80)
  uint32_t getCandidatesCount(types::node_t source_node) const {
    return detail::range_count(row(source_node), 0, targets_);
  }
  // getCandidatesCount(source_node, graph_start, graph_end) (candidates.hpp:82-100)
  uint32_t getCandidatesCount(types::node_t source_node, uint32_t graph_start, uint32_t graph_end) const {
    return detail::range_count(row(source_node), graph_start, graph_end);
  }
  // getCandidateAt(source_node, idx[, graph_start, graph_end]) (candidates.hpp:105-158):
  // the idx-th candidate in ascending node id, NULL_NODE when there are fewer
  types::node_t getCandidateAt(types::node_t source_node, uint32_t idx) const {
    return detail::range_select(row(source_node), idx, 0, targets_);
  }
  types::node_t getCandidateAt(types::node_t source_node, uint32_t idx, uint32_t graph_start,
                               uint32_t graph_end) const {
    return detail::range_select(row(source_node), idx, graph_start, graph_end);
  }
  bool contains(types::node_t source_node, types::node_t candidate) const {
    return (row(source_node)[candidate / 64] >> (candidate % 64)) & 1;
  }
  size_t getWordsPerRow() const { return W_; }
  size_t numRows() const { return rows_; }
  const std::vector<uint64_t>& words() const { return words_; }
  const uint64_t* data() const { return words_.data(); }
  size_t size() const { return words_.size(); }

 private:
  const uint64_t* row(types::node_t r) const {
    if (r >= rows_) throw std::out_of_range("bms2: candidate row out of range");
    return words_.data() + size_t(r) * W_;
  }
  std::vector<uint64_t> words_;
  size_t rows_, W_, targets_;
};

// Candidates (candidates.hpp:10-200)
class Candidates {
 public:
  Candidates(Queue& queue, size_t source_nodes, size_t target_nodes)
      : rows_(source_nodes), targets_(target_nodes) {
    check(bms2_candidates_create(queue.handle(), source_nodes, target_nodes, &c_));
  }
  ~Candidates() { bms2_candidates_destroy(c_); }
  Candidates(const Candidates&) = delete;
  // getHostCandidates (candidates.hpp:181-191): a host copy of the whole matrix.
  HostCandidates getHostCandidates() {
    std::vector<uint64_t> out(rows_ * getWordsPerRow());
    check(bms2_candidates_copy_to_host(c_, out.data()));
    return HostCandidates(std::move(out), rows_, getWordsPerRow(), targets_);
  }
  // getCandidatesCount(source_node[, start, end]) (candidates.hpp:193-196), reading the row
  // from the device.
  size_t getCandidatesCount(types::node_t source_node) const {
    const std::vector<uint64_t> r = hostRow(source_node);
    return detail::range_count(r.data(), 0, targets_);
  }
  size_t getCandidatesCount(types::node_t source_node, uint32_t graph_start, uint32_t graph_end) const {
    const std::vector<uint64_t> r = hostRow(source_node);
    return detail::range_count(r.data(), graph_start, graph_end);
  }
  // getCandidatesCount for every row at once, counted on the device.
  std::vector<uint32_t> getCandidatesCounts() {
    std::vector<uint32_t> out(rows_);
    check(bms2_candidates_row_counts(c_, out.data()));
    return out;
  }
  size_t getWordsPerRow() const { return size_t(bms2_candidates_words_per_row(c_)); }
  size_t getAllocationSize() const { return rows_ * getWordsPerRow() * sizeof(uint64_t); }
  bms2_candidates handle() const { return c_; }

 private:
  std::vector<uint64_t> hostRow(types::node_t r) const {
    std::vector<uint64_t> out(getWordsPerRow());
    check(bms2_candidates_copy_rows(c_, r, 1, out.data()));
    return out;
  }
  bms2_candidates c_ = nullptr;
  size_t rows_, targets_;
};
}  // namespace candidates

namespace isomorphism {
namespace filter {
// filterCandidates (isomorphism.hpp:30-104)
template <CandidatesDomain D = CandidatesDomain::Query>
inline utils::BatchedEvent filterCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graph,
                                            DeviceBatchedCSRGraph& data_graph,
                                            signature::Signature<>& label_signatures,
                                            signature::PathSignature& path_signatures,
                                            signature::CycleSignature& cycle_signatures,
                                            candidates::Candidates& candidates) {
  static_assert(D == CandidatesDomain::Query, "only the query-major candidate layout");
  bms2_event e = nullptr;
  check(bms2_filter_candidates(queue.handle(), &query_graph, &data_graph, label_signatures.handle(),
                               path_signatures.handle(), cycle_signatures.handle(),
                               candidates.handle(), &e));
  utils::BatchedEvent be;
  be.add(e);
  return be;
}

// refineCandidates (isomorphism.hpp:106-203)
template <CandidatesDomain D = CandidatesDomain::Query>
inline utils::BatchedEvent refineCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graph,
                                            DeviceBatchedCSRGraph& data_graph,
                                            signature::Signature<>& signatures,
                                            candidates::Candidates& candidates) {
  static_assert(D == CandidatesDomain::Query, "only the query-major candidate layout");
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
    num_data_ = data_graphs.num_graphs;
    utils::BatchedEvent be;
    be.add(e);
    return be;
  }
  size_t getTotalQueryIndices() const { return size_t(bms2_gmcr_total(g_)); }
  // Host copies: data_graph_offsets [num data graphs + 1], query_graph_indices [total].
  void copyToHost(std::vector<uint32_t>& offsets, std::vector<uint32_t>& indices) const {
    offsets.assign(size_t(num_data_) + 1, 0);
    indices.assign(getTotalQueryIndices(), 0);
    check(bms2_gmcr_copy_to_host(g_, offsets.data(), indices.data()));
  }
  bms2_gmcr handle() const { return g_; }

 private:
  bms2_gmcr g_ = nullptr;
  uint32_t num_data_ = 0;
};
}  // namespace mapping

namespace join {
namespace detail {
// Runs one of the per-pair join entry points with a device total and a device count buffer
// sized to the GMCR, then adds the total to *total_out and copies the counts out when asked.
template <typename Call>
inline utils::BatchedEvent with_pair_buffers(Queue& queue, mapping::GMCR& gmcr, size_t* total_out,
                                             std::vector<uint64_t>* per_pair, bool need_pairs,
                                             Call&& call) {
  void* d_total = nullptr;
  void* d_pairs = nullptr;
  const size_t T = gmcr.getTotalQueryIndices();
  check(bms2_device_malloc(queue.handle(), 8, &d_total));
  int rc = bms2_memset_zero(queue.handle(), d_total, 8);
  if (rc == BMS2_OK && (per_pair || need_pairs))
    rc = bms2_device_malloc(queue.handle(), 8 * (T ? T : 1), &d_pairs);
  bms2_event e = nullptr;
  if (rc == BMS2_OK)
    rc = call(static_cast<uint64_t*>(d_total), static_cast<uint64_t*>(d_pairs), uint64_t(d_pairs ? T : 0), &e);
  uint64_t total = 0;
  if (rc == BMS2_OK) rc = bms2_memcpy_d2h(queue.handle(), &total, d_total, 8);
  if (rc == BMS2_OK && per_pair) {
    per_pair->assign(T, 0);
    if (T) rc = bms2_memcpy_d2h(queue.handle(), per_pair->data(), d_pairs, 8 * T);
  }
  bms2_device_free(queue.handle(), d_total);
  if (d_pairs) bms2_device_free(queue.handle(), d_pairs);
  if (rc != BMS2_OK && e) bms2_event_destroy(e);
  check(rc);
  if (total_out) *total_out += size_t(total);
  utils::BatchedEvent be;
  be.add(e);
  return be;
}
}  // namespace detail

// joinCandidates (isomorphism.hpp:291-472).  The total is added to *num_matches when the call
// returns (the reference adds it on the device into host-visible memory); per_pair (optional)
// receives one count per GMCR entry.  SG_SIZE (the reference's sub-group width) has no effect:
// the join runs on 64-wide wavefronts.
template <int SG_SIZE = 32>
inline utils::BatchedEvent joinCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graphs,
                                          DeviceBatchedCSRGraph& data_graphs,
                                          candidates::Candidates& candidates, mapping::GMCR& gmcr,
                                          size_t* num_matches, bool find_first = true,
                                          std::vector<uint64_t>* per_pair = nullptr) {
  static_assert(SG_SIZE > 0, "sub-group size");
  return detail::with_pair_buffers(
      queue, gmcr, num_matches, per_pair, false,
      [&](uint64_t* d_total, uint64_t* d_pairs, uint64_t cap, bms2_event* e) {
        return bms2_join_candidates(queue.handle(), &query_graphs, &data_graphs,
                                    candidates.handle(), gmcr.handle(), d_total,
                                    find_first ? 1 : 0, d_pairs, cap, e);
      });
}

// joinPartialCandidates2 (isomorphism.hpp:1021-1178), the first step of BMS2 --use-cs: every
// embedding of query 0's first `end` nodes, written to `partial_matches` (records of end + 1
// ints {data node ids, data graph id}).  *num_partial_matches receives the number found.  The
// record capacity is `capacity` when given, else the size of the malloc_shared / malloc_device
// allocation holding partial_matches (the reference's call passes none, BMS2.cpp:250-252);
// throws when more records exist than fit (the reference writes past its buffer), after setting
// *num_partial_matches, so a caller can size the buffer exactly.
template <int SG_SIZE = 32>
inline utils::BatchedEvent joinPartialCandidates2(Queue& queue, DeviceBatchedCSRGraph& query_graphs,
                                                  DeviceBatchedCSRGraph& data_graphs,
                                                  candidates::Candidates& candidates, size_t end,
                                                  int* partial_matches, size_t* num_partial_matches,
                                                  size_t capacity = 0) {
  static_assert(SG_SIZE > 0, "sub-group size");
  if (capacity == 0 && partial_matches)
    capacity = sigmo::detail::lookup(partial_matches).bytes / (sizeof(int) * (end + 1));
  uint64_t n = 0;
  bms2_event e = nullptr;
  const int rc = bms2_join_partial_candidates(queue.handle(), &query_graphs, &data_graphs,
                                              candidates.handle(), uint32_t(end),
                                              reinterpret_cast<int32_t*>(partial_matches),
                                              capacity, &n, &e);
  if (num_partial_matches) *num_partial_matches = size_t(n);
  if (rc != BMS2_OK && e) bms2_event_destroy(e);
  check(rc);
  utils::BatchedEvent be;
  be.add(e);
  return be;
}

// joinWithPartialMatches (isomorphism.hpp:756-977), the second step of BMS2 --use-cs: extends
// the partial matches to every GMCR query of their data graph; the total is added to
// *total_full_matches when the call returns.  find_first keeps its flags per GMCR entry inside
// the call; when pair_done is a malloc_shared allocation of Q x D flags it also receives them
// (pair_done[q * D + g] = 1 for every pair with a match, isomorphism.hpp:838-851).  per_pair
// (optional) receives one count (find_first: 0/1) per GMCR entry.
template <int SG_SIZE = 32>
inline utils::BatchedEvent joinWithPartialMatches(Queue& queue, DeviceBatchedCSRGraph& query_graphs,
                                                  DeviceBatchedCSRGraph& data_graphs,
                                                  candidates::Candidates& candidates,
                                                  mapping::GMCR& gmcr, size_t end,
                                                  const int* partial_matches,
                                                  size_t num_partial_matches,
                                                  size_t* total_full_matches, bool find_first,
                                                  uint32_t* pair_done = nullptr,
                                                  std::vector<uint64_t>* per_pair = nullptr) {
  static_assert(SG_SIZE > 0, "sub-group size");
  const size_t Q = query_graphs.num_graphs, D = data_graphs.num_graphs;
  const sigmo::detail::Allocation pd = sigmo::detail::lookup(pair_done);
  const bool fill_flags = find_first && pair_done && pd.kind == sigmo::detail::Kind::Shared &&
                          pd.bytes >= sizeof(uint32_t) * Q * D;
  std::vector<uint64_t> flags;
  std::vector<uint64_t>* out = per_pair ? per_pair : (fill_flags ? &flags : nullptr);
  auto be = detail::with_pair_buffers(
      queue, gmcr, total_full_matches, out, true,
      [&](uint64_t* d_total, uint64_t* d_pairs, uint64_t cap, bms2_event* e) {
        return bms2_join_with_partial_matches(
            queue.handle(), &query_graphs, &data_graphs, candidates.handle(), gmcr.handle(),
            uint32_t(end), reinterpret_cast<const int32_t*>(partial_matches),
            num_partial_matches, d_total, find_first ? 1 : 0, d_pairs, cap, e);
      });
  if (fill_flags) {
    std::vector<uint32_t> off, idx;
    gmcr.copyToHost(off, idx);
    for (size_t g = 0; g < D; ++g)
      for (uint32_t k = off[g]; k < off[g + 1]; ++k)
        if ((*out)[k]) pair_done[size_t(idx[k]) * D + g] = 1u;
  }
  return be;
}
}  // namespace join
}  // namespace isomorphism

// sycl::malloc_device / malloc_shared / malloc_host and sycl::free for the buffers the drivers
// pass between calls.  Shared and host allocations are pinned host memory the kernels access
// directly.
template <typename T>
T* malloc_device(size_t count, Queue& queue) {
  void* p = nullptr;
  check(bms2_device_malloc(queue.handle(), sizeof(T) * (count ? count : 1), &p));
  detail::remember(p, sizeof(T) * count, detail::Kind::Device);
  return static_cast<T*>(p);
}
template <typename T>
T* malloc_shared(size_t count, Queue& queue) {
  void* p = nullptr;
  check(bms2_shared_malloc(queue.handle(), sizeof(T) * (count ? count : 1), &p));
  detail::remember(p, sizeof(T) * count, detail::Kind::Shared);
  return static_cast<T*>(p);
}
template <typename T>
T* malloc_host(size_t count, Queue& queue) {
  return malloc_shared<T>(count, queue);
}
inline void free(void* p, Queue& queue) {
  if (!p) return;
  detail::Kind k = detail::Kind::Device;
  detail::forget(p, &k);
  if (k == detail::Kind::Shared)
    bms2_shared_free(queue.handle(), p);
  else
    bms2_device_free(queue.handle(), p);
}
}  // namespace sigmo
