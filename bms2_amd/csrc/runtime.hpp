// Queues, events and HIP error plumbing shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

struct bms2_queue_s {
  int device;
  hipStream_t stream;
  bool owned;
  // grow-only scratch ordered on `stream`, released with the queue: the filter's query records
  // (64 B per query node), the refinement's signature planes (8 B per data node) and the join's
  // adjacency rows and depth adjacency (8 B per data node + 128 B per query for graphs of up to
  // 64 nodes; 16 or 32 B per data node + 128 B per query + 8 B per GMCR pair above); it grows to
  // the largest of these plus an eighth, at least 1 MiB
  void* scratch = nullptr;
  uint64_t scratch_bytes = 0;
};

struct bms2_event_s {
  hipEvent_t start;
  hipEvent_t stop;
};

namespace bms2 {

int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define BMS2_HIP(call)                                                              \
  do {                                                                              \
    hipError_t bms2_e_ = (call);                                                    \
    if (bms2_e_ != hipSuccess) return ::bms2::hip_fail(bms2_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define BMS2_CHECK_LAUNCH(name)                                                     \
  do {                                                                              \
    hipError_t bms2_e_ = hipGetLastError();                                        \
    if (bms2_e_ != hipSuccess) return ::bms2::hip_fail(bms2_e_, name, __FILE__, __LINE__); \
  } while (0)

// Makes the queue's device current.
int use_device(bms2_queue q);

// Brackets a stage call with an event pair when `ev` is non-null.  An event pair that end()
// has not handed out (the call failed in between) is destroyed with the scope.
struct EventScope {
  bms2_queue q;
  bms2_event* out;
  bms2_event_s* ev = nullptr;
  int begin();
  int end();
  ~EventScope();
};

int device_alloc(void** p, uint64_t bytes);

// Points *p at q's scratch of at least `bytes`, reallocated in stream order when it grows.  The
// contents are undefined; work on q's stream owns it until the next call.
int queue_scratch(bms2_queue q, uint64_t bytes, void** p);

// Staged join (staged.hip) with an optional matching order (d_jorder [Q][32], null = identity)
// and optional per-pair counts and DFS nodes; enqueued on q, no event.  Its scratch is the
// caller's `scratch` of scratch_bytes (the joins of graphs above 64 nodes then run as many waves
// as their stashes fit) or, when null, the queue's (staged_join_scratch_bytes); `shape`
// (nullable: the defaults) sets the launch of the joins of graphs above 64 nodes.
// Launch shape of the joins of graphs above 64 nodes (autotuned by the fused pipeline): waves
// per CU, waves per workgroup (1, 2 or 4), u64 words per adjacency row for graphs of 65-128
// nodes (2, or 4 to run every graph in the 4-word instance).
struct BigJoinShape {
  uint32_t waves = 12, block_waves = 1, row_words = 2;
};

uint64_t staged_join_scratch_bytes(const bms2_device_csr* query, const bms2_device_csr* data,
                                   uint64_t gmcr_pairs);
int staged_join(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                bms2_candidates cand, bms2_gmcr gmcr, int find_first, const uint8_t* d_jorder,
                uint64_t* d_total, uint64_t* d_pair_counts, uint64_t* d_pair_nodes,
                void* scratch, uint64_t scratch_bytes, const BigJoinShape* shape);

// Zeroes a Signature's label signatures and reachability state and a Candidates matrix (fill
// kernels on the signature's queue), so the staged sequence can run again on the same objects.
int staged_reset(bms2_signature sig, bms2_candidates cand);

}  // namespace bms2
