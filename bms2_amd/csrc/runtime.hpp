// Queues, events and HIP error plumbing shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "device_common.hpp"

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

// Device arrays of the staged signature objects (scope BMS2_SCOPE_DATA or BMS2_SCOPE_QUERY):
// path signatures [n][4], ring counts [n].
const uint64_t* path_signature_bits(bms2_path_signature s, int scope);
const uint64_t* cycle_signature_values(bms2_cycle_signature s, int scope);

// ---- screen and join of the data graphs of 65-256 nodes in the fused pipeline (big.hip)

constexpr int kBigMaxLevels = 8;   // refinement steps of the screen's shell levels
constexpr int kBigRecDwords = 48;  // screening record of one query node: 16 + 4 x kBigMaxLevels
// Control words (u64) shared by the screen and the join: [c] entries of size class c,
// [kBigClasses + c] runs of class c, [2 kBigClasses + c] join cursors, then record words and
// overflow.
constexpr int kBigCtlRecWords = 3 * kBigClasses;
constexpr int kBigCtlOverflow = kBigCtlRecWords + 1;
constexpr int kBigCtlWords = kBigCtlOverflow + 1;

// GMCR entries of the big-graph screen: entry c cap_e + i is the i-th pair of size class c; runs
// of consecutive entries of one data graph (run c cap_r + i is class c's i-th) are the join's
// units of work.
struct BigEntries {
  uint32_t* gidx;            // query of each entry
  uint64_t* rec_off;         // first word of the entry's domains in recs
  uint64_t cap_e;            // entries per class
  uint4* runs;               // {graph (index in the batch of big graphs), entries, first entry}
  uint64_t cap_r;            // runs per class
  uint64_t* recs;            // domains: the class's row words per depth of the matching order
  uint64_t cap_words;
  unsigned long long* ctl;   // kBigCtlWords control words
};

// Shell signatures of levels 1..K of every node of g into out [K][total_nodes].
int big_levels(hipStream_t st, const bms2_device_csr* g, uint32_t K, uint64_t* out);
// Screening records (kBigRecDwords per query node, at slot qpos[u]) from the query batch's path
// signatures, ring counts and shell levels [K][Qn]; the screen writes a node's domain at its
// depth in the matching order jorder ([Q][32], null = identity).
int big_query_records(hipStream_t st, const bms2_device_csr* query, const uint64_t* qpath,
                      const uint64_t* qcyc, const uint64_t* qlev, uint32_t K, const uint32_t* qpos,
                      const uint8_t* jorder, uint32_t* out);
// Screen of the graphs narrow[0 .. n_narrow) (at most 128 nodes) and wide[0 .. n_wide) (batch
// indices) against every query (prefilter records qpre, pre_dwords u32 each; first record and
// node count in words 17 and 18); e.ctl must be zeroed.  Writes the graphs' GMCR bit rows
// valid[gid[b]] (ceil(Q / 64) words each).
int big_screen(hipStream_t st, const bms2_device_csr* data, const uint64_t* dpath,
               const uint64_t* dcyc, const uint64_t* dlev, uint32_t K, const uint32_t* narrow,
               uint32_t n_narrow, const uint32_t* wide, uint32_t n_wide, uint32_t wide_above,
               uint32_t Q, const uint32_t* qpre, uint32_t pre_dwords, const uint32_t* bqrec,
               const BigEntries& e, const uint32_t* gid, uint64_t* valid);
// Joins of the screen's entries (class c: count[c] entries in runs[c] runs), the staged join's
// kernel reading the entries' domains from their records: each pair's count is added into
// d_counts[d_gid[g] Q + q] (zeroed), and the matched pairs, embeddings and DFS nodes into
// d_stats[1..3].  Enqueued on q; the 4-word class runs on q_wide (nullable) beside the 2-word
// class when both have entries, and q waits for it.
uint64_t big_join_scratch_bytes(const bms2_device_csr* query, const bms2_device_csr* data,
                                const BigJoinShape* shape);
int big_join(bms2_queue q, bms2_queue q_wide, const bms2_device_csr* query,
             const bms2_device_csr* data, const BigEntries& e, const uint64_t count[kBigClasses],
             const uint64_t runs[kBigClasses], int find_first, const uint8_t* d_jorder,
             uint64_t* d_counts, const uint32_t* d_gid, uint64_t* d_stats,
             const BigJoinShape* shape);
// Size split of the graphs above 64 nodes under a launch shape: graphs above this many nodes take
// 4-word adjacency rows and domains, the others 2.
inline uint32_t big_wide_above(const BigJoinShape& sh) { return sh.row_words == 4 ? 64u : 128u; }

}  // namespace bms2
