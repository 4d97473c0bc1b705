// Queues, events, device memory and device graph batches (C ABI).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "runtime.hpp"

namespace bms2 {

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  const char* base = std::strrchr(file, '/');
  return set_error(e == hipErrorOutOfMemory ? BMS2_ERR_NOMEM : BMS2_ERR_HIP, "%s failed: %s (%s:%d)",
                   what, hipGetErrorString(e), base ? base + 1 : file, line);
}

int use_device(bms2_queue q) {
  if (!q) return set_error(BMS2_ERR_INVALID, "queue is null");
  BMS2_HIP(hipSetDevice(q->device));
  return BMS2_OK;
}

int EventScope::begin() {
  if (!out) return BMS2_OK;
  ev = new bms2_event_s{};
  BMS2_HIP(hipEventCreate(&ev->start));
  BMS2_HIP(hipEventCreate(&ev->stop));
  BMS2_HIP(hipEventRecord(ev->start, q->stream));
  return BMS2_OK;
}

EventScope::~EventScope() {
  if (!ev) return;
  if (ev->start) (void)hipEventDestroy(ev->start);
  if (ev->stop) (void)hipEventDestroy(ev->stop);
  delete ev;
}

int EventScope::end() {
  if (!out) return BMS2_OK;
  BMS2_HIP(hipEventRecord(ev->stop, q->stream));
  *out = ev;
  ev = nullptr;
  return BMS2_OK;
}

int device_alloc(void** p, uint64_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_error(BMS2_ERR_NOMEM, "hipMalloc of %llu bytes failed: %s",
                     (unsigned long long)bytes, hipGetErrorString(e));
  }
  return BMS2_OK;
}

int queue_scratch(bms2_queue q, uint64_t bytes, void** p) {
  *p = nullptr;
  if (bytes > q->scratch_bytes) {
    if (q->scratch) BMS2_HIP(hipFreeAsync(q->scratch, q->stream));
    q->scratch = nullptr;
    q->scratch_bytes = 0;
    const uint64_t want = std::max<uint64_t>(bytes + bytes / 8, 1u << 20);
    hipError_t e = hipMallocAsync(&q->scratch, want, q->stream);
    if (e != hipSuccess) {
      q->scratch = nullptr;
      (void)hipGetLastError();
      return set_error(BMS2_ERR_NOMEM, "hipMallocAsync of %llu bytes failed: %s",
                       (unsigned long long)want, hipGetErrorString(e));
    }
    q->scratch_bytes = want;
  }
  *p = q->scratch;
  return BMS2_OK;
}

}  // namespace bms2

using namespace bms2;

extern "C" {

int bms2_queue_create(int device, bms2_queue* out) {
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  BMS2_HIP(hipSetDevice(device));
  hipStream_t s;
  BMS2_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = new bms2_queue_s{device, s, true};
  return BMS2_OK;
}

int bms2_queue_wrap(int device, void* hip_stream, bms2_queue* out) {
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  *out = new bms2_queue_s{device, static_cast<hipStream_t>(hip_stream), false};
  return BMS2_OK;
}

int bms2_queue_destroy(bms2_queue q) {
  if (!q) return BMS2_OK;
  if (q->scratch) {
    BMS2_HIP(hipSetDevice(q->device));
    BMS2_HIP(hipFreeAsync(q->scratch, q->stream));
    q->scratch = nullptr;
  }
  if (q->owned) {
    BMS2_HIP(hipSetDevice(q->device));
    BMS2_HIP(hipStreamSynchronize(q->stream));
    BMS2_HIP(hipStreamDestroy(q->stream));
  }
  delete q;
  return BMS2_OK;
}

int bms2_queue_wait(bms2_queue q) {
  if (int rc = use_device(q)) return rc;
  BMS2_HIP(hipStreamSynchronize(q->stream));
  return BMS2_OK;
}

void* bms2_queue_stream(bms2_queue q) { return q ? q->stream : nullptr; }

int bms2_event_elapsed_ms(bms2_event e, float* ms) {
  if (!e || !ms) return set_error(BMS2_ERR_INVALID, "null argument");
  BMS2_HIP(hipEventSynchronize(e->stop));
  BMS2_HIP(hipEventElapsedTime(ms, e->start, e->stop));
  return BMS2_OK;
}

int bms2_event_destroy(bms2_event e) {
  if (!e) return BMS2_OK;
  (void)hipEventDestroy(e->start);
  (void)hipEventDestroy(e->stop);
  delete e;
  return BMS2_OK;
}

int bms2_queue_device_info(bms2_queue q, uint64_t* global_mem_bytes, uint32_t* compute_units,
                           char* name, uint64_t name_capacity) {
  if (int rc = use_device(q)) return rc;
  hipDeviceProp_t prop;
  BMS2_HIP(hipGetDeviceProperties(&prop, q->device));
  if (global_mem_bytes) *global_mem_bytes = uint64_t(prop.totalGlobalMem);
  if (compute_units) *compute_units = uint32_t(prop.multiProcessorCount);
  if (name && name_capacity) {
    std::strncpy(name, prop.name, name_capacity - 1);
    name[name_capacity - 1] = 0;
  }
  return BMS2_OK;
}

int bms2_shared_malloc(bms2_queue q, uint64_t bytes, void** out) {
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  *out = nullptr;
  if (int rc = use_device(q)) return rc;
  hipError_t e = hipHostMalloc(out, bytes ? bytes : 8, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    (void)hipGetLastError();
    return set_error(BMS2_ERR_NOMEM, "hipHostMalloc of %llu bytes failed: %s",
                     (unsigned long long)bytes, hipGetErrorString(e));
  }
  return BMS2_OK;
}

int bms2_shared_free(bms2_queue q, void* ptr) {
  if (!ptr) return BMS2_OK;
  if (int rc = use_device(q)) return rc;
  BMS2_HIP(hipStreamSynchronize(q->stream));
  BMS2_HIP(hipHostFree(ptr));
  return BMS2_OK;
}

int bms2_device_malloc(bms2_queue q, uint64_t bytes, void** out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  return device_alloc(out, bytes);
}

int bms2_device_free(bms2_queue q, void* ptr) {
  if (int rc = use_device(q)) return rc;
  if (ptr) BMS2_HIP(hipFree(ptr));
  return BMS2_OK;
}

int bms2_memcpy_h2d(bms2_queue q, void* dst, const void* src, uint64_t bytes) {
  if (int rc = use_device(q)) return rc;
  if (bytes) BMS2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, q->stream));
  return BMS2_OK;
}

int bms2_memcpy_d2h(bms2_queue q, void* dst, const void* src, uint64_t bytes) {
  if (int rc = use_device(q)) return rc;
  if (bytes) {
    BMS2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, q->stream));
    BMS2_HIP(hipStreamSynchronize(q->stream));
  }
  return BMS2_OK;
}

int bms2_memset_zero(bms2_queue q, void* dst, uint64_t bytes) {
  if (int rc = use_device(q)) return rc;
  if (bytes) BMS2_HIP(hipMemsetAsync(dst, 0, bytes, q->stream));
  return BMS2_OK;
}

int bms2_device_csr_create(bms2_queue q, const bms2_host_csr* host, bms2_device_csr* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  uint32_t max_nodes = 0;
  if (int rc = validate_host_csr(host, BMS2_MAX_GRAPH_NODES, &max_nodes)) return rc;
  // Undirected graphs only: every CSR entry u->v needs a matching v->u with the same
  // multiplicity (the kernels test adjacency in one direction only, graph.hpp:68-73).
  uint32_t max_degree = 0;
  for (uint64_t u = 0; u < host->total_nodes; ++u) {
    uint32_t b = host->row_offsets[u], e = host->row_offsets[u + 1];
    if (e - b > max_degree) max_degree = e - b;
    for (uint32_t k = b; k < e; ++k) {
      uint32_t v = host->column_indices[k];
      if (v == u) return set_error(BMS2_ERR_UNSUPPORTED, "self-loop on node %llu", (unsigned long long)u);
      uint32_t fwd = 0, back = 0;
      for (uint32_t j = b; j < e; ++j) fwd += host->column_indices[j] == v;
      for (uint32_t j = host->row_offsets[v]; j < host->row_offsets[v + 1]; ++j)
        back += host->column_indices[j] == u;
      if (fwd != back)
        return set_error(BMS2_ERR_INVALID, "CSR is not symmetric at edge %llu-%u",
                         (unsigned long long)u, v);
    }
  }
  bms2_device_csr d{};
  d.num_graphs = host->num_graphs;
  d.max_graph_nodes = max_nodes;
  d.max_degree = max_degree;
  d.total_nodes = host->total_nodes;
  d.total_edges = host->total_edges;
  int rc = BMS2_OK;
  if ((rc = device_alloc((void**)&d.graph_offsets, 4ull * (d.num_graphs + 1))) ||
      (rc = device_alloc((void**)&d.row_offsets, 4ull * (d.total_nodes + 1))) ||
      (rc = device_alloc((void**)&d.column_indices, 4ull * d.total_edges)) ||
      (rc = device_alloc((void**)&d.node_labels, d.total_nodes)) ||
      (rc = device_alloc((void**)&d.edge_labels, d.total_edges))) {
    bms2_device_csr_destroy(q, &d);
    return rc;
  }
  hipStream_t s = q->stream;
  BMS2_HIP(hipMemcpyAsync(d.graph_offsets, host->graph_offsets, 4ull * (d.num_graphs + 1),
                          hipMemcpyHostToDevice, s));
  BMS2_HIP(hipMemcpyAsync(d.row_offsets, host->row_offsets, 4ull * (d.total_nodes + 1),
                          hipMemcpyHostToDevice, s));
  if (d.total_edges) {
    BMS2_HIP(hipMemcpyAsync(d.column_indices, host->column_indices, 4ull * d.total_edges,
                            hipMemcpyHostToDevice, s));
    if (host->edge_labels)
      BMS2_HIP(hipMemcpyAsync(d.edge_labels, host->edge_labels, d.total_edges,
                              hipMemcpyHostToDevice, s));
    else
      BMS2_HIP(hipMemsetAsync(d.edge_labels, 0, d.total_edges, s));
  }
  if (d.total_nodes)
    BMS2_HIP(hipMemcpyAsync(d.node_labels, host->node_labels, d.total_nodes,
                            hipMemcpyHostToDevice, s));
  BMS2_HIP(hipStreamSynchronize(s));  // host arrays may be released on return (graph.hpp:307)
  *out = d;
  return BMS2_OK;
}

int bms2_device_csr_destroy(bms2_queue q, bms2_device_csr* g) {
  if (!g) return BMS2_OK;
  if (int rc = use_device(q)) return rc;
  (void)hipFree(g->graph_offsets);
  (void)hipFree(g->row_offsets);
  (void)hipFree(g->column_indices);
  (void)hipFree(g->node_labels);
  (void)hipFree(g->edge_labels);
  std::memset(g, 0, sizeof(*g));
  return BMS2_OK;
}

uint64_t bms2_device_csr_alloc_size(const bms2_device_csr* g) {
  if (!g) return 0;
  return g->total_nodes * 1 + (uint64_t(g->num_graphs) + 1) * 4 + (g->total_nodes + 1) * 4 +
         g->total_edges * 4 + g->total_edges * 1;
}

}  // extern "C"
